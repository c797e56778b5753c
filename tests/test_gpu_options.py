"""GPU parity of the FisherMatr variable changes and of the duty cycle.

use_m1m2 / use_chi1chi2=False (gwfast/signal.py:818-845, 522-533): the
device applies F' = J^T F J with the exact Jacobian of (m1, m2) -> (Mc, eta)
and (chiS, chiA) -> (chi1z, chi2z); the oracle seeds its dual numbers directly
on m1, m2, chiS, chiA, so the two agree to rounding (<= 1e-9 normalised).

DutyFactor (signal.py:672-761, 809-1073; network.py:36-51): each detector
reseeds numpy's global generator with its seedUse and draws one Bernoulli
mask per arm, in arm order; a masked arm contributes nothing to SNR^2 or to
the Fisher matrix.
"""
import numpy as np
import pytest

from tests._common import NET_C1, bbh, bns, hip_net, norm_err, oracle_net
from tests.test_gpu_parity import _models

pytestmark = pytest.mark.gpu

VARIABLE_SETS = [(True, True), (False, False), (True, False)]


@pytest.mark.parametrize('model', ['PD', 'NRT'])
def test_fisher_variable_change_parity(gpu, model):
    from oracle import detector as OD
    wf, owf = _models(model)
    ev = (bbh if model == 'PD' else bns)(5, seed=21)
    net = hip_net(wf, NET_C1, True)
    odets = oracle_net(NET_C1, True)
    for use_m1m2, use_chi1chi2 in VARIABLE_SETS:
        F = net.FisherMatr(dict(ev), res=800, use_m1m2=use_m1m2, use_chi1chi2=use_chi1chi2)
        Fd = OD.network_fisher(odets, owf, ev, 800, mode='dual', use_m1m2=use_m1m2, use_chi1chi2=use_chi1chi2)
        e = norm_err(F, Fd)
        print(model, 'use_m1m2', use_m1m2, 'use_chi1chi2', use_chi1chi2, 'err %.3e' % e, flush=True)
        assert F.shape == (wf.nParams, wf.nParams, 5)
        assert e <= 1e-9


def test_derivative_rows_in_m1m2_chiSA(gpu):
    """FisherMatr(return_derivatives=True, use_m1m2=True, use_chi1chi2=False)
    rows against the oracle's duals in the new variables."""
    from oracle import detector as OD
    wf, owf = _models('PD')
    ev = bbh(4, seed=22)
    sites = NET_C1[1:2]
    sig = hip_net(wf, sites, True).signals['CE1Id']
    od = oracle_net(sites, True)[0]
    res = 300
    fgrids, _ = od.grid(owf, ev, res)
    Fs, Ds = sig.FisherMatr(dict(ev), res=res, return_derivatives=True, use_m1m2=True, use_chi1chi2=False)
    Do = od.full_dual_derivatives(owf, ev, fgrids, use_m1m2=True, use_chi1chi2=False)
    Do[7] /= 86400.
    scale = np.max(np.abs(Do), axis=(1, 2), keepdims=True)
    assert np.max(np.abs(Ds[0] - Do) / scale) <= 1e-5
    Fo = od.fisher(owf, ev, res, mode='dual', use_m1m2=True, use_chi1chi2=False)
    assert norm_err(Fs[0], Fo) <= 1e-9


def _masks(seed, n_arms, n, p):
    np.random.seed(seed)
    return [np.random.choice([0, 1], n, p=[1. - p, p]) for _ in range(n_arms)]


def test_duty_factor_snr_and_fisher(gpu):
    wf, _ = _models('PD')
    n, res, p = 40, 400, 0.5
    ev = bbh(n, seed=31)
    base = hip_net(wf, NET_C1, True)
    duty = hip_net(wf, NET_C1, True, DutyFactor=p)
    seeds = [101, 202, 303]
    duty._update_all_seeds(seeds, verbose=False)
    arms = {'ETS': ['ETS_0', 'ETS_1', 'ETS_2'], 'CE1Id': ['CE1Id'], 'CE2NM': ['CE2NM']}
    want_mask = {}
    for (name, _), seed in zip(NET_C1, seeds):
        for key, m in zip(arms[name], _masks(seed, len(arms[name]), n, p)):
            want_mask[key] = m.astype(float)
    allm = np.array(list(want_mask.values()))
    assert 0 < allm.mean() < 1

    s0 = base.SNR(dict(ev), res=res, return_all=True)
    s1 = duty.SNR(dict(ev), res=res, return_all=True)
    for key, m in want_mask.items():
        assert np.array_equal(s1[key], s0[key] * m), key
    want_net = np.sqrt(sum((s0[k] * want_mask[k]) ** 2 for k in want_mask))
    assert np.max(np.abs(s1['net'] - want_net) / np.maximum(want_net, 1e-300)) <= 1e-14
    again = duty.SNR(dict(ev), res=res)  # the same seeds, the same masks
    assert np.max(np.abs(again - s1['net']) / np.maximum(s1['net'], 1e-300)) <= 1e-14

    F0 = base.FisherMatr(dict(ev), res=res, return_all=True)
    F1 = duty.FisherMatr(dict(ev), res=res)
    want = sum(F0[k] * want_mask[k] for k in want_mask)
    live = np.einsum('iin->n', want) > 0
    assert np.all(F1[:, :, ~live] == 0)
    assert norm_err(F1[:, :, live], want[:, :, live]) <= 1e-13

    sig = duty.signals['ETS']
    s_sig = sig.SNRInteg(dict(ev), res=res, return_all=True)
    for i in range(3):
        assert np.max(np.abs(s_sig[i] - s1['ETS_%d' % i])) <= 1e-13 * np.max(s1['ETS_%d' % i])


def _spd(P, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, P, P))
    return np.transpose(A @ np.transpose(A, (0, 2, 1)) + P * np.eye(P), (1, 2, 0))


def test_fishertools_variable_changes(gpu):
    """fisherTools.py:410-458, 620-822 on the device against the oracle's
    per-event matrix products (batched (P, P, N) and single (P, P) with
    scalar parameters)."""
    from gwfast_amd import fisherTools as FT
    from gwfast_amd.waveforms import IMRPhenomD
    from oracle import covariance as OC
    ev = bbh(7, seed=41)
    pn = IMRPhenomD().ParNums
    M = _spd(11, 7, 3)
    Mc, eta = ev['Mc'], ev['eta']
    m1, m2 = OC.m1m2_from_mceta_unguarded(Mc, eta)
    Jme, Jmm = OC.jac_m1m2_mceta(Mc, eta), OC.jac_mceta_m1m2(m1, m2)
    Jch = OC.jac_chi1chi2_chieff_dchi(m1, m2)
    T = lambda J: np.swapaxes(J, 0, 1)
    pn_s = dict(pn)
    pn_s['chiS'], pn_s['chiA'] = pn_s.pop('chi1z'), pn_s.pop('chi2z')
    half = np.array([[.5, .5], [.5, -.5]])
    cases = [
        (FT.m1m2_to_Mceta_fish, pn, OC.sandwich(M, 0, 1, Jme, Jme)),
        (FT.m1m2_to_Mceta_cov, pn, OC.sandwich(M, 0, 1, Jmm, Jmm)),
        (FT.Mceta_to_m1m2_fish, pn, OC.sandwich(M, 0, 1, T(Jmm), Jmm)),
        (FT.Mceta_to_m1m2_cov, pn, OC.sandwich(M, 0, 1, T(Jme), Jme)),
        (FT.chi1chi2_to_chieffDeltachi_fish, pn, OC.sandwich(M, 9, 10, T(Jch), Jch)),
        (FT.chiSchiA_to_chi1chi2_fish, pn_s, OC.sandwich(M, 9, 10, half, half)),
        (FT.log_dL_to_dL_derivative_fish, pn, OC.scale_row_col(M, 2, 1. / ev['dL'])),
        (FT.log_dL_to_dL_derivative_cov, pn, OC.scale_row_col(M, 2, ev['dL'])),
    ]
    for fn, parnums, want in cases:
        got = fn(M, parnums, ev)
        err = np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300))
        print(fn.__name__, 'max rel err %.3e' % err, flush=True)
        assert err <= 1e-13, fn.__name__
        one = {k: float(v[2]) for k, v in ev.items()}
        got1 = fn(M[:, :, 2], parnums, one)
        assert got1.shape == (11, 11)
        assert np.max(np.abs(got1 - want[:, :, 2]) / np.maximum(np.abs(want[:, :, 2]), 1e-300)) <= 1e-13, fn.__name__
    assert np.allclose(FT.J_m1m2_Mceta(Mc, eta), Jme, rtol=1e-14, atol=0)
    assert np.allclose(FT.J_Mceta_m1m2(m1, m2), Jmm, rtol=1e-14, atol=1e-300)


def _net_with_motions(wf, motions):
    from gwfast_amd import gwfastGlobals as glob
    from gwfast_amd.network import DetNet
    from gwfast_amd.signal import GWSignal
    sigs = {}
    for (name, psd), m in zip(NET_C1, motions):
        d = glob.detectors[name]
        sigs[name] = GWSignal(wf, psd_path=psd, detector_shape=d['shape'], det_lat=d['lat'], det_long=d['long'],
                              det_xax=d['xax'], useEarthMotion=m == 'earth', noMotion=m == 'none', fmin=2.,
                              verbose=False)
    return DetNet(sigs, verbose=False)


@pytest.mark.parametrize('model,motion', [('PD', 'earth'), ('NRT', 'earth'), ('PD', 'mixed'), ('NRT', 'mixed')])
def test_masked_fisher_kernel(gpu, model, motion):
    """gwf_fisher_masked: per event, the sum of the Fisher matrices of the
    arms whose bit is set, against per-arm launches; all bits set gives the
    fused network Fisher, no bit gives zero.  'mixed': rotating-Earth,
    fixed-Earth and motionless sites in one network (the pure-row consumer
    for 11 parameters, the matrix-core Gram for 13)."""
    from gwfast_amd import signal as S
    wf, _ = _models(model)
    n, res = 48, 500
    ev = (bbh if model == 'PD' else bns)(n, seed=51)
    net = hip_net(wf, NET_C1, True) if motion == 'earth' else _net_with_motions(wf, ('earth', 'fixed', 'none'))
    sigs = list(net.signals.values())
    sigs[0]._complete(ev)
    rows = sigs[0]._rows(ev)
    per_arm = [S.fisher_device(sigs, wf, rows, res, arm=k).cpu().numpy() for k in range(5)]
    full = S.fisher_device(sigs, wf, rows, res).cpu().numpy()
    rng = np.random.default_rng(52)
    bits = rng.integers(0, 2, size=(5, n))
    bits[:, 0] = 0
    bits[:, 1] = 1
    bits[:3, 2] = 1          # the whole triangle, no L detector
    bits[:, 3] = [1, 0, 1, 0, 1]
    Fm = S.fisher_masked_device(sigs, wf, rows, res, S.mask_words([bits[:3], bits[3:4], bits[4:5]])).cpu().numpy()
    want = sum(per_arm[k] * bits[k] for k in range(5))
    assert np.all(Fm[:, :, bits.sum(axis=0) == 0] == 0)
    live = np.nonzero(bits.any(axis=0))[0]
    e = norm_err(Fm[:, :, live], want[:, :, live])
    print(model, motion, 'masked vs per-arm sum %.3e' % e, flush=True)
    assert e <= 1e-13
    assert norm_err(Fm[:, :, 1:2], full[:, :, 1:2]) <= 1e-14


def test_catalogue_forecast_with_duty_factor(gpu, tmp_path):
    """The driver's duty cycle (calculate_forecasts_from_catalog.py:423-556):
    per-detector seeds, per-arm masks on the SNRs, detection on the masked
    network SNR, masked per-arm Fisher sum of the detected events."""
    from gwfast_amd import forecast as FC
    from gwfast_amd.waveforms import IMRPhenomD
    ev = bbh(120, seed=61)
    ev['dL'] = ev['dL'] * 4.0
    wf = IMRPhenomD()
    net = hip_net(wf, NET_C1, True)
    seeds, p, res = [7, 8, 9], 0.6, 400
    FC.run_catalog(ev, net, str(tmp_path), batch_size=120, res=res, duty_factor=p, seeds=seeds, verbose=False)
    s_all = net.SNR(dict(ev), res=res, return_all=True)
    masks = {}
    for (name, _), seed in zip(NET_C1, seeds):
        keys = ['ETS_0', 'ETS_1', 'ETS_2'] if name == 'ETS' else [name]
        for key, m in zip(keys, _masks(seed, len(keys), 120, p)):
            masks[key] = m
    snr = np.sqrt(sum((s_all[k] * masks[k]) ** 2 for k in masks))
    got = np.loadtxt(tmp_path / 'snrs_0_to_120.txt')
    assert np.allclose(got, snr, rtol=1e-14, atol=0)
    det = np.nonzero(snr > 12.0)[0]
    assert 0 < len(det) < 120
    assert np.array_equal(np.atleast_1d(np.loadtxt(tmp_path / 'idxs_det_0_to_120.txt')).astype(int), det)
    F0 = net.FisherMatr({k: v[det] for k, v in ev.items()}, res=res, return_all=True)
    want = sum(F0[k] * masks[k][det] for k in masks)
    F = np.load(tmp_path / 'fishers_0_to_120.npy')
    assert norm_err(F, want) <= 1e-13


def _cov_cases():
    """(name, F (P, P, N), kwargs) covering every inversion method."""
    rng = np.random.default_rng(3)
    P = 6
    X = rng.standard_normal((4, P, P))
    A = X @ np.transpose(X, (0, 2, 1)) + 3 * P * np.eye(P)
    s = np.geomspace(1, 1e4, P)
    good = np.transpose(A, (1, 2, 0)) * s[:, None, None] * s[None, :, None]
    r = 1 - 1e-8
    B = np.eye(5)
    B[0, 1] = B[1, 0] = r
    d = np.linspace(0.5, 3.0, 5)
    blk = (d[:, None] * B * d[None, :])[:, :, None]
    rng = np.random.default_rng(0)            # symmetric, indefinite, LU pivoting is a 3-cycle
    Xs = rng.standard_normal((4, 4))
    S = (Xs + Xs.T) / 2
    np.fill_diagonal(S, np.abs(np.diag(S)) * 0.1 + 0.01)
    cyc = S[:, :, None]
    out = []
    for m in ('cho', 'inv', 'svd', 'svd_reg', 'lu'):
        out.append(('good-' + m, good, dict(invMethodIn=m)))
    out.append(('svd_reg-block', blk, dict(invMethodIn='svd_reg', svals_thresh=1e-6)))
    out.append(('svd-trunc-block', blk, dict(invMethodIn='svd', truncate=True, svals_thresh=1e-6, condNumbMax=1e3)))
    for alt in ('svd', 'inv', 'lu', 'svd_reg'):
        out.append(('cyc-cho-alt-' + alt, cyc, dict(invMethodIn='cho', alt_method=alt)))
    out.append(('cyc-lu', cyc, dict(invMethodIn='lu')))
    return out


@pytest.mark.parametrize('case', _cov_cases(), ids=[c[0] for c in _cov_cases()])
def test_covmatr_methods_vs_oracle(gpu, case):
    """CovMatr with every invMethodIn / alt_method / truncation against the
    oracle's mpmath restatement (fisherTools.py:32-196)."""
    from gwfast_amd import fisherTools as FT
    from oracle import covariance as OC
    name, F, kw = case
    C, eps, cond, meth = FT.CovMatr(F, return_details=True, **kw)
    Co, epso, metho = OC.cov_matr(F, inv_method=kw.get('invMethodIn', 'cho'), alt_method=kw.get('alt_method', 'svd'),
                                  truncate=kw.get('truncate', False), svals_thresh=kw.get('svals_thresh', 1e-15),
                                  cond_max=kw.get('condNumbMax', 1e50))
    assert np.array_equal(meth, metho), (meth, metho)
    scale = np.max(np.abs(Co), axis=(0, 1))
    err = np.max(np.abs(C - Co) / scale)
    print(name, 'cov err %.3e, eps %s vs %s' % (err, eps, epso), flush=True)
    assert err <= 1e-9
    assert np.allclose(eps, epso, rtol=1e-3, atol=1e-10)


def _graded_spd(rng, P, n, spread):
    """n symmetric positive-definite P x P matrices (P, P, n) with diagonal
    scales spread over `spread` decades, like Fisher matrices."""
    X = rng.standard_normal((n, P, P))
    A = X @ np.transpose(X, (0, 2, 1)) + P * np.eye(P)
    s = np.geomspace(1.0, 10.0 ** spread, P)
    return np.transpose(A, (1, 2, 0)) * s[:, None, None] * s[None, :, None]


@pytest.mark.parametrize('P,alt', [(11, 'svd'), (12, 'svd'), (13, 'svd'), (12, 'inv'), (12, 'lu'), (12, 'svd_reg')])
def test_covmatr_lane_path_vs_oracle(gpu, P, alt):
    """'cho' at 11-13 parameters runs one event per lane; the events it
    leaves (an indefinite matrix takes alt_method, an all-NaN one gives NaN)
    are finished by the wavefront kernel.  A mixed batch against the
    oracle's mpmath restatement (fisherTools.py:53-196)."""
    from gwfast_amd import fisherTools as FT
    from oracle import covariance as OC
    rng = np.random.default_rng(50 + P)
    F = _graded_spd(rng, P, 6, 5)
    B = np.diag(np.linspace(1.0, 2.0, P))
    B[0, 1] = B[1, 0] = 5.0                     # indefinite 2 x 2 block
    F[:, :, 2] = B
    F[:, :, 4] = np.nan
    C, eps, cond, meth = FT.CovMatr(F, return_details=True, alt_method=alt)
    Co, epso, metho = OC.cov_matr(F, alt_method=alt)
    assert np.array_equal(meth, metho), (meth, metho)
    assert meth[2] == OC.METHOD_IDS[alt] and meth[4] == -1 and np.all(meth[[0, 1, 3, 5]] == 0)
    assert np.all(np.isnan(C[:, :, 4])) and np.isnan(eps[4]) and np.isnan(cond[4])
    ok = meth >= 0
    scale = np.max(np.abs(Co[:, :, ok]), axis=(0, 1))
    err = np.max(np.abs(C[:, :, ok] - Co[:, :, ok]) / scale)
    print(P, alt, 'cov err %.3e' % err, flush=True)
    assert err <= 1e-9
    assert np.allclose(eps[ok], epso[ok], rtol=1e-3, atol=1e-10)


@pytest.mark.parametrize('P', [11, 12, 13])
def test_check_fisher_lane_and_wavefront_paths(gpu, P):
    """CheckFisher at 11 and 12 parameters runs one event per lane, at 13 one
    wavefront per event; an all-NaN event goes to the wavefront kernel.
    Eigenvalues against LAPACK (numpy.linalg.eigh) and the oracle's mpmath
    restatement (fisherTools.py:216-279), eigenvector residuals and
    orthonormality, ascending order, condition numbers."""
    from gwfast_amd import fisherTools as FT
    from oracle import covariance as OC
    rng = np.random.default_rng(60 + P)
    n = 70                                      # more than one wave of lanes
    F = _graded_spd(rng, P, n, 6)
    F[:, :, 65] = np.nan
    ev, vec, cond = FT.CheckFisher(F)
    assert np.all(np.isnan(ev[65])) and np.all(np.isnan(vec[65])) and np.isnan(cond[65])
    keep = np.arange(n) != 65
    Ft = np.transpose(F[:, :, keep], (2, 0, 1))
    wn = np.linalg.eigvalsh(Ft)
    lam_max = np.max(np.abs(wn), axis=1, keepdims=True)
    assert np.max(np.abs(ev[keep] - wn) / lam_max) <= 1e-12
    assert np.all(np.diff(ev[keep], axis=1) >= 0)
    V = vec[keep]
    r = Ft @ V - V * ev[keep][:, None, :]
    assert np.max(np.abs(r) / lam_max[:, :, None]) <= 1e-10
    assert np.max(np.abs(np.transpose(V, (0, 2, 1)) @ V - np.eye(P))) <= 1e-12
    assert np.allclose(cond[keep], np.max(np.abs(ev[keep]), axis=1) / np.min(np.abs(ev[keep]), axis=1), rtol=0,
                       atol=0)
    evo, _, condo = OC.check_fisher(F[:, :, :3])
    assert np.max(np.abs(ev[:3] - evo) / lam_max[:3]) <= 1e-12
    # the smallest eigenvalue carries an absolute error of order eps * lam_max
    assert np.all(np.abs(cond[:3] - condo) <= 1e-14 * cond[:3] * condo)


@pytest.mark.parametrize('model', ['TF2', 'PD'])
def test_fisher_linear_spacing_and_df(gpu, model):
    """FisherMatr(spacing='lin') (signal.py:893-894) against the oracle on
    linear grids, and the df option (884-892): one grid length, the largest
    floor(1 + (fcut - fmin)/df) over the events."""
    from oracle import detector as OD
    wf, owf = _models(model)
    ev = bbh(6, seed=71)
    sites = NET_C1 if model == 'PD' else NET_C1[1:2]
    net = hip_net(wf, sites, True)
    odets = oracle_net(sites, True)
    F = net.FisherMatr(dict(ev), res=1500, spacing='lin')
    Fd = OD.network_fisher(odets, owf, ev, 1500, mode='dual', spacing='lin')
    e = norm_err(F, Fd)
    print(model, 'lin-spaced Fisher err %.3e' % e, flush=True)
    assert e <= 1e-10
    assert norm_err(F, net.FisherMatr(dict(ev), res=1500)) > 1e-6   # a different grid gives a different matrix
    df = 2.0
    fcut = owf.fcut(ev)
    want_res = int(np.amax(np.floor(1 + (fcut - 2.0) / df)))
    Fdf = net.FisherMatr(dict(ev), res=None, df=df, spacing='lin')
    assert norm_err(Fdf, net.FisherMatr(dict(ev), res=want_res, spacing='lin')) == 0.0
    sig = next(iter(net.signals.values()))
    Fs, Ds = sig.FisherMatr(dict(ev), res=300, spacing='lin', return_derivatives=True)
    f, _ = odets[0].grid(owf, ev, 300, spacing='lin')
    Do = odets[0].full_dual_derivatives(owf, ev, f)
    Do[7] /= 86400.
    scale = np.max(np.abs(Do), axis=(1, 2), keepdims=True)
    assert np.max(np.abs(Ds[0] - Do) / scale) <= 1e-5


def test_gwstrain_single_components_and_m1m2(gpu):
    """GWstrain(return_single_comp=...) (signal.py:629-652) against the
    oracle's amplitudes and phase, and is_m1m2 (522-524)."""
    from gwfast_amd.gwfastUtils import m1m2_from_Mceta
    wf, owf = _models('PD')
    ev = bbh(5, seed=81)
    sites = NET_C1[:1]
    sig = hip_net(wf, sites, True).signals['ETS']
    od = oracle_net(sites, True)[0]
    f, _ = od.grid(owf, ev, 400)
    args = (ev['Mc'], ev['eta'], ev['dL'], ev['theta'], ev['phi'], ev['iota'], ev['psi'], ev['tcoal'], ev['Phicoal'],
            ev['chi1z'], ev['chi2z'])
    Ap, Ac, Psi = od.strain_components(owf, ev, f, rot=60.)
    got = {c: sig.GWstrain(f, *args, rot=60., is_chi1chi2=True, return_single_comp=c)
           for c in ('Ap', 'Ac', 'Psip', 'Psic', 'At', 'Psit')}
    amax = np.max(np.hypot(Ap, Ac))
    assert np.max(np.abs(got['Ap'] - Ap)) <= 1e-12 * amax
    assert np.max(np.abs(got['Ac'] - Ac)) <= 1e-12 * amax
    assert np.max(np.abs(got['At'] - np.hypot(Ap, Ac))) <= 1e-12 * amax
    tol = 1e-13 * np.max(np.abs(Psi)) + 1e-6      # |Psi| reaches 1e9 rad: compare to its rounding
    assert np.max(np.abs(got['Psip'] - Psi)) <= tol
    assert np.max(np.abs(got['Psic'] - (Psi + np.pi / 2))) <= tol
    assert np.max(np.abs(got['Psit'] - (Psi + np.arctan2(Ac, Ap)))) <= tol
    with pytest.raises(ValueError):
        sig.GWstrain(f, *args, rot=60., is_chi1chi2=True, return_single_comp='Phi')
    m1, m2 = m1m2_from_Mceta(ev['Mc'], ev['eta'])
    h1 = sig.GWstrain(f, *args, is_chi1chi2=True)
    h2 = sig.GWstrain(f, m1, m2, *args[2:], is_chi1chi2=True, is_m1m2=True)
    assert np.max(np.abs(h1 - h2)) <= 1e-6 * np.max(np.abs(h1))


def test_wf_overlap_and_amplitudes(gpu):
    """GWSignal / DetNet.WFOverlap (signal.py:1759-1935, network.py:198-232)
    against the oracle: identical signals give 1, a coalescence-phase shift d
    gives cos d, mixed models match the oracle's overlap; GWAmplitudes
    (signal.py:425-466) against the oracle's amplitudes."""
    from gwfast_amd.waveforms import TaylorF2_RestrictedPN
    from oracle import detector as OD
    from oracle import waveforms as OW
    wf, owf = _models('PD')
    ev = bbh(6, seed=91)
    net = hip_net(wf, NET_C1, True)
    odets = oracle_net(NET_C1, True)
    assert np.allclose(net.WFOverlap(wf, wf, dict(ev), dict(ev), res=500), 1.0, rtol=1e-12, atol=0)
    ev2 = dict(ev)
    ev2['Phicoal'] = ev['Phicoal'] + 0.7
    ev2['dL'] = ev['dL'] * 2.0
    got = net.WFOverlap(wf, wf, dict(ev), dict(ev2), res=500)
    assert np.allclose(got, np.cos(0.7), rtol=0, atol=1e-6)
    tf2, otf2 = TaylorF2_RestrictedPN(), OW.TaylorF2()
    ev3 = dict(ev)
    ev3['Mc'] = ev['Mc'] * (1 + 1e-4)
    got = net.WFOverlap(wf, tf2, dict(ev), dict(ev3), res=500)
    want = OD.network_overlap(odets, owf, otf2, ev, ev3, 500)
    print('mixed-model overlap', got, want, flush=True)
    assert np.allclose(got, want, rtol=0, atol=1e-6)
    sig = net.signals['CE1Id']
    f, _ = odets[1].grid(owf, ev, 300)
    Ap, Ac = sig.GWAmplitudes(dict(ev), f)
    Apo, Aco = odets[1].gw_amplitudes(owf, ev, f)
    amax = np.max(np.hypot(Apo, Aco))
    assert np.max(np.abs(Ap - Apo)) <= 1e-12 * amax and np.max(np.abs(Ac - Aco)) <= 1e-12 * amax


def test_eccentric_taylorf2(gpu):
    """TaylorF2_RestrictedPN(is_eccentric=True) (waveforms.py:814-845): the
    phase, the SNR and the 12-parameter Fisher (ParNums['ecc'] = 11) against
    the oracle, both with v0 = min v over the grid and with fRef_ecc."""
    from gwfast_amd.waveforms import TaylorF2_RestrictedPN
    from oracle import detector as OD
    from oracle import waveforms as OW
    ev = bbh(6, seed=97)
    ev['ecc'] = np.linspace(0.01, 0.15, 6)
    for fref in (None, 5.0):
        wf = TaylorF2_RestrictedPN(is_eccentric=True, fRef_ecc=fref)
        owf = OW.TaylorF2(is_eccentric=True, fRef_ecc=fref)
        assert wf.ParNums['ecc'] == 11 and wf.nParams == 12
        f = np.geomspace(2.0, 0.9 * owf.fcut(ev), 500)
        ph = wf.Phi(f, **ev)
        pho = owf.Phi(f, ev)
        assert np.max(np.abs(ph - pho)) <= 1e-13 * np.max(np.abs(pho)) + 1e-9
        sites = NET_C1
        net = hip_net(wf, sites, True)
        odets = oracle_net(sites, True)
        snr = net.SNR(dict(ev), res=1000)
        snro = OD.network_snr(odets, owf, ev, 1000)
        assert np.max(np.abs(snr / snro - 1)) <= 1e-11
        F = net.FisherMatr(dict(ev), res=1000)
        assert F.shape == (12, 12, 6)
        Fd = OD.network_fisher(odets, owf, ev, 1000, mode='dual')
        e = norm_err(F, Fd)
        print('eccentric TF2 fRef_ecc=%s Fisher err %.3e' % (fref, e), flush=True)
        assert e <= 1e-10
        assert np.all(np.diag(F[:, :, 0]) > 0) and F[11, 11, 0] > 0
    ev0 = dict(ev)
    ev0['ecc'] = np.zeros(6)
    wf = TaylorF2_RestrictedPN(is_eccentric=True)
    wfc = TaylorF2_RestrictedPN()
    net, netc = hip_net(wf, NET_C1[1:2], True), hip_net(wfc, NET_C1[1:2], True)
    F0, Fc = net.FisherMatr(dict(ev0), res=800), netc.FisherMatr(dict(ev0), res=800)
    assert norm_err(F0[:11, :11], Fc) <= 1e-13      # e0 = 0: the circular Fisher on the other rows
    with pytest.raises(ValueError):
        net.SNR({k: v for k, v in ev.items() if k != 'ecc'}, res=200)


@pytest.mark.parametrize('model', ['PD', 'NRT'])
def test_fisher_all_arms_one_launch(gpu, model):
    """gwf_fisher_arms (FisherMatr(return_all=True), signal.py:1087-1098,
    network.py:125-135): every physical arm's Fisher from one launch equals
    the single-arm launches; their sum is the network Fisher; DetNet's
    return_all dict carries them per detector and arm, and agrees with the
    oracle per arm."""
    from gwfast_amd import signal as S
    from oracle import detector as OD
    wf, owf = _models(model)
    n, res = 40, 600
    ev = (bbh if model == 'PD' else bns)(n, seed=61)
    net = hip_net(wf, NET_C1, True)
    sigs = list(net.signals.values())
    sigs[0]._complete(ev)
    rows = sigs[0]._rows(ev)
    Fa = S.fisher_arms_device(sigs, wf, rows, res).cpu().numpy()
    assert Fa.shape == (5, wf.nParams, wf.nParams, n)
    for k in range(5):
        Fk = S.fisher_device(sigs, wf, rows, res, arm=k).cpu().numpy()
        e = norm_err(Fa[k], Fk)
        print(model, 'arm', k, 'one-launch vs single-arm %.3e' % e, flush=True)
        assert e <= 1e-13
    full = S.fisher_device(sigs, wf, rows, res).cpu().numpy()
    assert norm_err(Fa.sum(axis=0), full) <= 1e-13
    Fl = S.fisher_arms_device(sigs, wf, rows, res, lin=True).cpu().numpy()   # spacing='lin' grids
    for k in (0, 3):
        assert norm_err(Fl[k], S.fisher_device(sigs, wf, rows, res, arm=k, lin=True).cpu().numpy()) <= 1e-13
    assert norm_err(Fl[0], Fa[0]) > 1e-6
    allF = net.FisherMatr(dict(ev), res=res, return_all=True)
    assert sorted(allF) == sorted(['ETS_0', 'ETS_1', 'ETS_2', 'CE1Id', 'CE2NM', 'net'])
    for k, key in enumerate(['ETS_0', 'ETS_1', 'ETS_2', 'CE1Id', 'CE2NM']):
        assert norm_err(allF[key], Fa[k]) <= 1e-15
    odets = oracle_net(NET_C1[1:2], True)
    Fo = OD.network_fisher(odets, owf, ev, res, mode='dual')
    assert norm_err(allF['CE1Id'], Fo) <= 1e-10


def test_fisher_all_arms_ragged(gpu):
    """gwf_fisher_arms with per-event grid lengths (the event queue of the
    pair kernel): every event's per-arm matrices equal a uniform launch at
    that event's length (the two are separately compiled kernels: equal to
    rounding, measured 3e-15); DetNet's return_all takes the same path."""
    from gwfast_amd import signal as S
    from gwfast_amd.waveforms import IMRPhenomD
    wf = IMRPhenomD()
    n = 30
    ev = bbh(n, seed=67)
    lens = np.array([64, 65, 300, 777, 1000], dtype=np.int32)
    res = lens[np.arange(n) % len(lens)]
    net = hip_net(wf, NET_C1, True)
    sigs = list(net.signals.values())
    sigs[0]._complete(ev)
    rows = sigs[0]._rows(ev)
    Fa = S.fisher_arms_device(sigs, wf, rows, res).cpu().numpy()
    for r in lens:
        sel = np.where(res == r)[0]
        Fu = S.fisher_arms_device(sigs, wf, rows[:, sel].contiguous(), int(r)).cpu().numpy()
        for k in range(5):
            assert norm_err(Fa[k][:, :, sel], Fu[k]) <= 1e-13, (r, k)
    allF = net.FisherMatr(dict(ev), res=res, return_all=True)
    for k, key in enumerate(['ETS_0', 'ETS_1', 'ETS_2', 'CE1Id', 'CE2NM']):
        assert norm_err(allF[key], Fa[k]) <= 1e-15


def test_pair_kernel_expired_wait_gives_nan(gpu):
    """A wait of the producer / consumer kernel that runs out of polls ends
    its wave and the launch reports NaN (forced from the host with a limit of
    one poll, which the consumer's first wait cannot meet); the default limit
    then gives the normal result again."""
    from gwfast_amd import _native as N
    from gwfast_amd import signal as S
    from gwfast_amd.waveforms import IMRPhenomD
    wf = IMRPhenomD()
    ev = bbh(16, seed=71)
    net = hip_net(wf, NET_C1, True)
    sigs = list(net.signals.values())
    sigs[0]._complete(ev)
    rows = sigs[0]._rows(ev)
    good = S.fisher_device(sigs, wf, rows, 800).cpu().numpy()
    L = N.lib()
    try:
        N.check(L.gwf_set_queue_spin_limit(1), L)
        bad = S.fisher_device(sigs, wf, rows, 800).cpu().numpy()
    finally:
        N.check(L.gwf_set_queue_spin_limit(0), L)
    assert np.all(np.isnan(bad))
    again = S.fisher_device(sigs, wf, rows, 800).cpu().numpy()
    assert np.array_equal(again, good)


def test_tidal_eccentric_taylorf2(gpu):
    """TaylorF2_RestrictedPN(is_tidal=True, is_eccentric=True): 14 parameters
    (ParNums LambdaTilde 11, deltaLambda 12, ecc 13; waveforms.py:106-115):
    phase, SNR and network Fisher against the oracle; e0 = 0 gives the tidal
    Fisher on the first 13 rows."""
    from gwfast_amd.waveforms import TaylorF2_RestrictedPN
    from oracle import detector as OD
    from oracle import waveforms as OW
    ev = bns(6, seed=99)
    ev['ecc'] = np.linspace(0.01, 0.15, 6)
    wf = TaylorF2_RestrictedPN(is_tidal=True, is_eccentric=True)
    owf = OW.TaylorF2(is_tidal=True, is_eccentric=True)
    assert wf.nParams == 14 and wf.ParNums['ecc'] == 13
    f = np.geomspace(2.0, 0.9 * owf.fcut(ev), 500)
    ph = wf.Phi(f, **ev)
    pho = owf.Phi(f, ev)
    assert np.max(np.abs(ph - pho)) <= 1e-13 * np.max(np.abs(pho)) + 1e-9
    net = hip_net(wf, NET_C1, True)
    odets = oracle_net(NET_C1, True)
    snr = net.SNR(dict(ev), res=1000)
    assert np.max(np.abs(snr / OD.network_snr(odets, owf, ev, 1000) - 1)) <= 1e-11
    F = net.FisherMatr(dict(ev), res=1000)
    assert F.shape == (14, 14, 6)
    e = norm_err(F, OD.network_fisher(odets, owf, ev, 1000, mode='dual'))
    print('tidal+eccentric TF2 Fisher err %.3e' % e, flush=True)
    assert e <= 1e-10
    ev0 = dict(ev)
    ev0['ecc'] = np.zeros(6)
    F0 = net.FisherMatr(dict(ev0), res=1000)
    Ft = hip_net(TaylorF2_RestrictedPN(is_tidal=True), NET_C1, True).FisherMatr(dict(ev0), res=1000)
    assert norm_err(F0[:13, :13], Ft) <= 1e-13


def test_gwstrain_any_rotation(gpu):
    """GWstrain and GWAmplitudes at rot = 37 deg (any angle, signal.py:485)
    against the oracle's pattern functions at that rotation, on the ET
    triangle and an L detector."""
    wf, owf = _models('PD')
    ev = bbh(4, seed=83)
    for sites in (NET_C1[:1], NET_C1[1:2]):
        sig = next(iter(hip_net(wf, sites, True).signals.values()))
        od = oracle_net(sites, True)[0]
        f, _ = od.grid(owf, ev, 300)
        args = (ev['Mc'], ev['eta'], ev['dL'], ev['theta'], ev['phi'], ev['iota'], ev['psi'], ev['tcoal'],
                ev['Phicoal'], ev['chi1z'], ev['chi2z'])
        Ap, Ac, Psi = od.strain_components(owf, ev, f, rot=37.)
        got_ap = sig.GWstrain(f, *args, rot=37., is_chi1chi2=True, return_single_comp='Ap')
        got_ac = sig.GWstrain(f, *args, rot=37., is_chi1chi2=True, return_single_comp='Ac')
        amax = np.max(np.hypot(Ap, Ac))
        print(sites[0][0], 'rot 37: Ap err %.3e Ac err %.3e' % (np.max(np.abs(got_ap - Ap)) / amax,
                                                               np.max(np.abs(got_ac - Ac)) / amax), flush=True)
        assert np.max(np.abs(got_ap - Ap)) <= 1e-12 * amax
        assert np.max(np.abs(got_ac - Ac)) <= 1e-12 * amax
        a2, c2 = sig.GWAmplitudes(dict(ev), f, rot=37.)
        assert np.max(np.abs(a2 - Ap)) <= 1e-12 * amax and np.max(np.abs(c2 - Ac)) <= 1e-12 * amax
        h = sig.GWstrain(f, *args, rot=37., is_chi1chi2=True)
        ho = od.gw_strain(owf, ev, f, rot=37.)
        # the strain's phase reaches 1e9 rad: compare to its rounding
        assert np.max(np.abs(h - ho)) <= 1e-6 * np.max(np.abs(ho))


def test_catalogue_forecast_eccentric_model(gpu, tmp_path):
    """The catalogue driver with eccentric TaylorF2 (the driver's 'tf2_ecc',
    calculate_forecasts_from_catalog.py:68; 12 parameters, row 11 = ecc):
    the chunk's Fisher matrices equal DetNet.FisherMatr on the detected
    events."""
    from gwfast_amd import forecast as FC
    from gwfast_amd.waveforms import TaylorF2_RestrictedPN
    wf = TaylorF2_RestrictedPN(is_eccentric=True, fRef_ecc=5.)
    ev = bbh(40, seed=85)
    ev['ecc'] = np.linspace(0.0, 0.1, 40)
    net = hip_net(wf, NET_C1, True)
    suffs = FC.run_catalog(dict(ev), net, str(tmp_path), batch_size=40, res=400, verbose=False)
    snr = net.SNR(dict(ev), res=400)
    det = np.nonzero(snr > 12.0)[0]
    assert len(det) > 0
    F = np.load(tmp_path / ('fishers%s.npy' % suffs[0]))
    Fr = net.FisherMatr({k: v[det] for k, v in ev.items()}, res=400)
    print('ecc catalogue: %d detected, Fisher %s' % (len(det), F.shape), flush=True)
    assert F.shape == (12, 12, len(det)) and np.array_equal(F, Fr)
