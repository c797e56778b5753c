"""CPU tests of the boundary: the C-ABI library loads and exports every symbol
the header declares (no compute calls), the host-side API logic, the
catalogue generator, and the fit-coefficient tables against the reference's
source text."""
import ctypes
import os
import re

import numpy as np
import pytest

from gwfast_amd import _native as N
from gwfast_amd import catalog, gwfastGlobals as glob, gwfastUtils as U

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'gwfast_hip.h')
REF = '/root/reference/gwfast/waveforms.py'


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(gwf_[a-z_0-9]+)\s*\(', txt)))


@pytest.fixture(scope='module')
def so_path():
    if not os.path.exists(N.LIB_PATH):
        from gwfast_amd import build
        build.build(verbose=False)
    return N.LIB_PATH


def test_library_exports_every_declared_symbol(so_path):
    L = ctypes.CDLL(so_path)
    decl = _declared()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(L, name), name
    # ctypes binding covers exactly the declared ABI
    assert sorted(N.SIGNATURES) == decl


def test_library_binds_without_gpu(so_path):
    L = N.load(so_path)
    assert L.gwf_abi_version() == 1
    assert L.gwf_workspace_bytes(10) > L.gwf_workspace_bytes(0) > 0
    d = (N.GwfDetector * 3)()
    d[0].shape, d[1].shape, d[2].shape = N.SHAPE_T, N.SHAPE_L, N.SHAPE_L
    assert L.gwf_num_arms(d, 3) == 5
    # argument validation happens before any device work
    rc = L.gwf_cov_invert(None, 40, 1, None, None, None, None, None)
    assert rc == -22 and b'P must be' in L.gwf_last_error()


def test_ctypes_signatures_match_header_arity():
    """Every entry point declared in include/gwfast_hip.h has a ctypes
    binding with the same number of arguments (the Python layer calls the
    library through these bindings only)."""
    hdr = open(HEADER).read()
    hdr = re.sub(r'/\*.*?\*/', '', hdr, flags=re.S)
    decls = re.findall(r'\b(?:int|size_t|const char\*)\s+(gwf_\w+)\s*\(([^;]*?)\)\s*;', hdr, flags=re.S)
    assert len(decls) >= 40
    for name, args in decls:
        n = 0 if args.strip() in ('', 'void') else len(args.split(','))
        assert name in N.SIGNATURES, name
        assert len(N.SIGNATURES[name][1]) == n, name
    assert set(N.SIGNATURES) == {d[0] for d in decls}


def test_struct_layouts_match_header():
    # field order and sizes of the ABI structs
    assert ctypes.sizeof(N.GwfWaveform) == 8 * 4 + 2 * 8 + 3 * 8 + 2 * 4 + 8
    assert ctypes.sizeof(N.GwfDetector) == 4 * 4 + 5 * 8 + 2 * 8 + 2 * 4 + 2 * 8


# ------------------------------------------------------------------ host API
def test_parnums_and_nparams():
    from gwfast_amd.waveforms import IMRPhenomD, TaylorF2_RestrictedPN
    assert list(TaylorF2_RestrictedPN().ParNums) == ['Mc', 'eta', 'dL', 'theta', 'phi', 'iota', 'psi', 'tcoal',
                                                     'Phicoal', 'chi1z', 'chi2z']
    t = TaylorF2_RestrictedPN(is_tidal=True)
    assert t.nParams == 13 and t.ParNums['LambdaTilde'] == 11 and t.ParNums['deltaLambda'] == 12
    assert t.objType == 'BNS' and t.is_holomorphic
    d = IMRPhenomD()
    assert d.nParams == 11 and d.fcutPar == 0.2 and d.objType == 'BBH'
    assert len(d.QNMgrid_a) == 1003
    e = TaylorF2_RestrictedPN(is_eccentric=True)
    assert e.nParams == 12 and e.ParNums['ecc'] == 11
    te = TaylorF2_RestrictedPN(is_eccentric=True, is_tidal=True)   # waveforms.py:106-115
    assert te.nParams == 14 and te.ParNums['LambdaTilde'] == 11 and te.ParNums['deltaLambda'] == 12
    assert te.ParNums['ecc'] == 13 and list(te.ParNums.values()) == list(range(14))


def test_gwsignal_validation():
    from gwfast_amd.signal import GWSignal
    from gwfast_amd.waveforms import IMRPhenomD
    with pytest.raises(ValueError, match='valid detector configuration'):
        GWSignal(IMRPhenomD(), psd_path=glob.PSD_ET, detector_shape='X')
    with pytest.raises(ValueError, match='PSD or ASD path'):
        GWSignal(IMRPhenomD())
    s = GWSignal(IMRPhenomD(), psd_path=glob.PSD_ET, detector_shape='T', useEarthMotion=True, noMotion=True)
    assert not s.useEarthMotion  # noMotion wins (signal.py:139-141)
    assert s.noiseCurve[0] == pytest.approx(3.4959517e-17 ** 2)


def test_check_evparams_completion_and_errors():
    ev = {'m1': np.array([30.]), 'm2': np.array([20.]), 'ra': np.array([1.]), 'dec': np.array([0.2]),
          'thetaJN': np.array([0.3]), 'tcoal': np.array([0.])}
    U.check_evparams(ev)
    assert np.allclose(ev['theta'], np.pi / 2 - 0.2) and ev['phi'][0] == 1. and ev['iota'][0] == 0.3
    Mc, eta = U.Mceta_from_m1m2(30., 20.)
    assert ev['Mc'][0] == pytest.approx(Mc) and ev['eta'][0] == pytest.approx(0.24)
    with pytest.raises(ValueError, match='tGPS and tcoal'):
        U.check_evparams({'Mc': 1.})


def test_tidal_conversions_round_trip():
    eta = np.array([0.24, 0.2499, 0.2])
    L1, L2 = np.array([300., 10., 1500.]), np.array([700., 2000., 100.])
    Lt, dL = U.Lamt_delLam_from_Lam12(L1, L2, eta)
    a, b = U.Lam12_from_Lamt_delLam(Lt, dL, eta)
    assert np.allclose(a, L1, rtol=1e-11) and np.allclose(b, L2, rtol=1e-11)
    # equal masses: LambdaTilde = (L1 + L2)/2 ... for L1 = L2 = L it equals L
    Lt, _ = U.Lamt_delLam_from_Lam12(400., 400., 0.25)
    assert Lt == pytest.approx(400., rel=1e-12)


# ---------------------------------------------------------------- catalogue
def test_luminosity_distance_quadrature():
    from scipy.integrate import quad
    z = np.array([0.1, 1.0, 2.0])
    ref = [(1 + zz) * 2.99792458e5 / 67.7 * quad(lambda x: 1 / np.sqrt(0.31 * (1 + x) ** 3 + 0.69), 0, zz)[0] / 1e3
           for zz in z]
    assert np.allclose(catalog.dL_Gpc(z), ref, rtol=1e-12)


def test_bbh_draws_in_ranges():
    ev = catalog.draw_bbh(5000, seed=0)
    assert np.all((ev['eta'] > 0.055) & (ev['eta'] <= 0.25))
    assert np.all((ev['Mc'] > 4.7) & (ev['Mc'] < 210.))
    assert np.all((ev['dL'] > 0.45) & (ev['dL'] < 16.))
    assert abs(np.mean(np.cos(ev['theta']))) < 0.05
    a = catalog.draw_bbh(7, seed=3)
    b = catalog.draw_bbh(7, seed=3)
    assert all(np.array_equal(a[k], b[k]) for k in a)


# ---------------------------------------------------- fit tables vs reference
def _ref_fit_numbers():
    src = open(REF).read().split('\n')[958:1333]
    names = ['sigma1', 'sigma2', 'sigma3', 'sigma4', 'beta1', 'beta2', 'beta3', 'alpha1', 'alpha2', 'alpha3',
             'alpha4', 'alpha5', 'gamma1', 'gamma2', 'gamma3', 'rho1', 'rho2', 'rho3', 'v2']
    out = {}
    for line in src:
        s = line.strip()
        for n in names:
            if s.startswith(n + ' =') and n not in out:
                toks = re.findall(r'([-+]?)\s*(?<![A-Za-z_\d.])((?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?)',
                                  s.split('=', 1)[1])
                out[n] = [(-1 if sg == '-' else 1) * float(m) for sg, m in toks]
    return out


@pytest.mark.skipif(not os.path.exists(REF), reason='reference source not present')
def test_phenomd_fit_tables_match_reference_text():
    """The 19 IMRPhenomD fits (11 numbers each) in the oracle and in the HIP
    source equal the numbers in the reference's text (waveforms.py:1039-1227)."""
    from oracle.waveforms import PHENOMD_FITS
    ref = _ref_fit_numbers()
    assert len(ref) == 19
    for n, nums in ref.items():
        assert list(PHENOMD_FITS[n]) == nums, n
    hip = open(os.path.join(ROOT, 'gwfast_amd', 'csrc', 'gwf_waveforms.h')).read()
    for n, nums in ref.items():
        m = re.search(r'/\* %s \*/ \{([^}]*)\}' % n, hip)
        assert m, n
        vals = [float(x) for x in m.group(1).split(',')]
        assert vals == nums, n


def test_ragged_grid_length_validation():
    """Per-event grid lengths are checked on the host before they reach the
    ragged kernels: each in [2, 2**30], one per event."""
    from gwfast_amd import signal as S
    with pytest.raises(ValueError):
        S._res_arg(np.array([1, 500]), 2, 'cpu')
    with pytest.raises(ValueError):
        S._res_arg(np.array([500, 2 ** 31 - 1]), 2, 'cpu')
    with pytest.raises(ValueError):
        S._res_arg(np.array([500, 600, 700]), 2, 'cpu')
    r0, r = S._res_arg(np.array([2, 500]), 2, 'cpu')
    assert r0 == 0 and r.tolist() == [2, 500]


def test_tidal_conversions_equal_the_pinned_oracle():
    """The host conversion (gwfastUtils.py:398-417) equals the oracle's,
    which the per-body tidal-phase test pins."""
    from oracle import waveforms as OW
    eta = np.array([0.24, 0.2499, 0.2, 0.1])
    L1, L2 = np.array([300., 10., 1500., 0.]), np.array([700., 2000., 100., 50.])
    a = U.Lamt_delLam_from_Lam12(L1, L2, eta)
    b = OW.lamt_dellam_from_lam12(L1, L2, eta)
    assert np.allclose(a[0], b[0], rtol=1e-14, atol=1e-12) and np.allclose(a[1], b[1], rtol=1e-14, atol=1e-12)


# ------------------------------------------------------- parameter rows
def _models():
    from gwfast_amd.waveforms import IMRPhenomD, IMRPhenomD_NRTidalv2, TaylorF2_RestrictedPN
    return [(IMRPhenomD(), 11), (TaylorF2_RestrictedPN(), 11), (TaylorF2_RestrictedPN(is_tidal=True), 13),
            (TaylorF2_RestrictedPN(is_eccentric=True), 12),
            (TaylorF2_RestrictedPN(is_tidal=True, is_eccentric=True), 14), (IMRPhenomD_NRTidalv2(), 13)]


def test_parameter_rows_match_each_model(monkeypatch):
    """GWSignal builds exactly the rows its model's kernels read (11, +2
    tidal, +1 eccentric), and the catalogue driver takes its rows from
    GWSignal."""
    import inspect
    import torch
    from gwfast_amd import _engine as E, forecast
    from gwfast_amd.signal import GWSignal
    monkeypatch.setattr(E, 'to_dev', lambda a, dtype=None: torch.as_tensor(np.ascontiguousarray(a)))
    ev = catalog.draw_bns(3, seed=2)
    ev['ecc'] = np.full(3, 0.05)
    for wf, want in _models():
        assert E.n_rows(wf) == want == wf.nParams
        sig = GWSignal(wf, psd_path=glob.PSD_ET, detector_shape='L', verbose=False)
        assert sig._rows(dict(ev)).shape == (want, 3)
    assert 'sig0._rows(events)' in inspect.getsource(forecast.compute_errs)


def test_parameter_row_checks_reject_bad_rows():
    """The ABI takes the row stride but not the row count: the host rejects
    rows the kernels would read past their end, and other layouts."""
    import torch
    from gwfast_amd import _engine as E
    from gwfast_amd.waveforms import TaylorF2_RestrictedPN
    wf = TaylorF2_RestrictedPN(is_tidal=True)
    with pytest.raises(ValueError, match='needs 13 parameter rows, got 11'):
        E.check_rows(torch.zeros((11, 4), dtype=torch.float64), wf)
    with pytest.raises(ValueError, match='float64'):
        E.check_rows(torch.zeros((13, 4), dtype=torch.float32), wf)
    with pytest.raises(ValueError, match='2-D'):
        E.check_rows(torch.zeros(13, dtype=torch.float64), wf)
    with pytest.raises(ValueError, match='contiguous along events'):
        E.check_rows(torch.zeros((4, 13), dtype=torch.float64).t(), wf)
    with pytest.raises(ValueError, match='on the GPU'):
        E.check_rows(torch.zeros((13, 4), dtype=torch.float64), wf)
    with pytest.raises(ValueError, match='needs 11 parameter rows, got 3'):
        E.check_rows(torch.zeros((3, 4), dtype=torch.float64), need=11)


def test_any_arm_rotation_is_a_turned_x_arm():
    """GWstrain / GWAmplitudes at any rot (signal.py:485) run arm 0 of the
    same site with its x-arm turned by rot: on the oracle, the strain at
    rot = 37 deg equals the strain of the turned detector at rot = 0, for L
    and T shapes."""
    from gwfast_amd.signal import GWSignal, rot_detector
    from gwfast_amd.waveforms import IMRPhenomD
    from oracle import detector as OD, waveforms as OW
    ev = catalog.draw_bbh(3, seed=4)
    f = np.geomspace(5., 200., 30)[:, None] * np.ones((1, 3))
    wf = OW.IMRPhenomD()
    for shape, site in (('L', 'CE1Id'), ('T', 'ETS')):
        d = glob.detectors[site]
        a = OD.Detector(glob.PSD_ET, shape, d['lat'], d['long'], d['xax'], useEarthMotion=True)
        b = OD.Detector(glob.PSD_ET, shape, d['lat'], d['long'], d['xax'] + 37., useEarthMotion=True)
        ha, hb = a.gw_strain(wf, ev, f, rot=37.), b.gw_strain(wf, ev, f, rot=0.)
        assert np.max(np.abs(ha - hb)) <= 1e-12 * np.max(np.abs(ha))
        sig = GWSignal(IMRPhenomD(), psd_path=glob.PSD_ET, detector_shape=shape, det_lat=d['lat'],
                       det_long=d['long'], det_xax=d['xax'], useEarthMotion=True, verbose=False)
        # stand-in device tables: the block only records their addresses
        sig._psd_dev = tuple(type('T', (), {'data_ptr': lambda self: 0})() for _ in range(2))
        det = rot_detector(sig, 37.)
        assert det.xax_deg == pytest.approx(d['xax'] + 37.) and det.compute2arms == 0
