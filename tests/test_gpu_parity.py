"""GPU parity of the HIP path against the CPU oracle (FP64).

Tolerances: SNR <= 1e-10 relative; Fisher <= 1e-6 normalised
(|dF_ij| / sqrt(F_ii F_jj)) against the oracle's reference-faithful
construction, and <= 1e-10 against the oracle's all-dual construction (the
same mathematics without the reference's phase-association rounding).
"""
import numpy as np
import pytest

from tests._common import NET_C0, NET_C1, bbh, bns, hip_net, norm_err, oracle_net

pytestmark = pytest.mark.gpu


def _models(name, **kw):
    from gwfast_amd import waveforms as W
    from oracle import waveforms as OW
    if name == 'TF2':
        return W.TaylorF2_RestrictedPN(**kw), OW.TaylorF2(**kw)
    if name == 'NRT':
        # the kernels cut the taper's end point (see IMRPhenomD_NRTidalv2)
        return W.IMRPhenomD_NRTidalv2(**kw), OW.IMRPhenomD_NRTidalv2(taper_end_inclusive=True, **kw)
    return W.IMRPhenomD(**kw), OW.IMRPhenomD(**kw)


CASES = [
    # id, model, model kwargs, sites, earth motion, n events, res, event draw
    ('C0-TF2-L', 'TF2', {}, NET_C0, False, 16, 1000, bbh),
    ('C0-TF2-L-spinHO', 'TF2', {'use_3p5PN_SpinHO': True}, NET_C0, False, 6, 500, bbh),
    ('C1-PhenomD-net', 'PD', {}, NET_C1, True, 8, 2000, bbh),
    ('C1-PhenomD-net-static', 'PD', {}, NET_C1, False, 6, 700, bbh),
    ('C3-TF2-tidal-ET', 'TF2', {'is_tidal': True}, NET_C1[:1], True, 6, 1500, bns),
    ('NRTidalv2-net', 'NRT', {}, NET_C1, True, 8, 1500, bns),
]


@pytest.mark.parametrize('case', CASES, ids=[c[0] for c in CASES])
def test_snr_parity(gpu, case):
    cid, model, kw, sites, earth, n, res, draw = case
    from oracle import detector as OD
    wf, owf = _models(model, **kw)
    ev = draw(n, seed=3)
    net = hip_net(wf, sites, earth)
    snr = net.SNR(dict(ev), res=res)
    ref = OD.network_snr(oracle_net(sites, earth), owf, ev, res)
    err = np.max(np.abs(snr - ref) / ref)
    print(cid, 'SNR max rel err', err, flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize('case', CASES, ids=[c[0] for c in CASES])
def test_fisher_parity(gpu, case):
    cid, model, kw, sites, earth, n, res, draw = case
    from oracle import detector as OD
    wf, owf = _models(model, **kw)
    ev = draw(n, seed=5)
    net = hip_net(wf, sites, earth)
    F = net.FisherMatr(dict(ev), res=res)
    odets = oracle_net(sites, earth)
    Fd = OD.network_fisher(odets, owf, ev, res, mode='dual')
    Ff = OD.network_fisher(odets, owf, ev, res, mode='faithful')
    e_d, e_f = norm_err(F, Fd), norm_err(F, Ff)
    print(cid, 'Fisher err vs dual %.3e vs faithful %.3e' % (e_d, e_f), flush=True)
    assert F.shape == (wf.nParams, wf.nParams, n)
    assert e_d <= 1e-10
    assert e_f <= 1e-6


@pytest.mark.parametrize('model', ['PD', 'NRT'])
@pytest.mark.parametrize('motions', [('earth', 'fixed', 'none'), ('earth', 'none', 'earth'), ('fixed', 'none', 'fixed')],
                         ids=['earth-fixed-none', 'earth-none', 'fixed-none'])
def test_fisher_mixed_site_motion(gpu, motions, model):
    """One network whose sites move differently (rotating Earth, Earth fixed
    at tcoal, no motion).  The 11-parameter consumer factors the Mc, eta and
    tcoal rows through the sites' common d t / dX, which a network with both
    rotating- and fixed-Earth sites lacks: that one takes the pure-row path;
    the others the factored one (a site without motion adds no rotation term).
    Both against the oracle with the same per-site motion; 11 parameters
    (IMRPhenomD) and 13 (IMRPhenomD_NRTidalv2, the matrix-core consumers)."""
    from gwfast_amd import gwfastGlobals as glob
    from gwfast_amd.network import DetNet
    from gwfast_amd.signal import GWSignal
    from oracle import detector as OD
    wf, owf = _models(model)
    ev = (bbh if model == 'PD' else bns)(6, seed=13)
    sigs, odets = {}, []
    for (name, psd), m in zip(NET_C1, motions):
        d = glob.detectors[name]
        kw = dict(useEarthMotion=m == 'earth', noMotion=m == 'none', fmin=2.)
        sigs[name] = GWSignal(wf, psd_path=psd, detector_shape=d['shape'], det_lat=d['lat'], det_long=d['long'],
                              det_xax=d['xax'], verbose=False, **kw)
        odets.append(OD.Detector(psd, d['shape'], d['lat'], d['long'], d['xax'], **kw))
    net = DetNet(sigs, verbose=False)
    F = net.FisherMatr(dict(ev), res=900)
    Fd = OD.network_fisher(odets, owf, ev, 900, mode='dual')
    e_d = norm_err(F, Fd)
    print(model, motions, 'Fisher err vs dual %.3e' % e_d, flush=True)
    assert e_d <= 1e-10


def test_fisher_per_arm_sums_to_network(gpu):
    wf, _ = _models('PD')
    ev = bbh(5, seed=7)
    net = hip_net(wf, NET_C1, True)
    allF = net.FisherMatr(dict(ev), res=400, return_all=True)
    assert set(allF) == {'ETS_0', 'ETS_1', 'ETS_2', 'CE1Id', 'CE2NM', 'net'}
    tot = net.FisherMatr(dict(ev), res=400)
    assert norm_err(allF['net'], tot) <= 1e-13


def test_snr_return_all(gpu):
    from oracle import detector as OD
    wf, owf = _models('PD')
    ev = bbh(5, seed=8)
    net = hip_net(wf, NET_C1, True)
    s = net.SNR(dict(ev), res=300, return_all=True)
    od = oracle_net(NET_C1, True)
    arms = od[0].snr(owf, ev, 300, return_all=True)
    for i in range(3):
        assert np.max(np.abs(s['ETS_%d' % i] - arms[i]) / arms[i]) <= 1e-10
    assert np.max(np.abs(s['CE1Id'] - od[1].snr(owf, ev, 300)) / s['CE1Id']) <= 1e-10


@pytest.mark.parametrize('model', ['TF2', 'PD'])
def test_strain_and_derivatives(gpu, model):
    """GWstrain and FisherMatr(return_derivatives=True) against the oracle's
    strain and derivative rows, and the plain (materialised-J) Fisher against
    the fused one."""
    from oracle import detector as OD
    wf, owf = _models(model)
    ev = bbh(4, seed=9)
    sites = NET_C1[:1]
    net = hip_net(wf, sites, True)
    sig = net.signals['ETS']
    od = oracle_net(sites, True)[0]
    res = 300
    fgrids, _ = od.grid(owf, ev, res)
    h = sig.GWstrain(fgrids, ev['Mc'], ev['eta'], ev['dL'], ev['theta'], ev['phi'], ev['iota'], ev['psi'],
                     ev['tcoal'], ev['Phicoal'], ev['chi1z'], ev['chi2z'], rot=60., is_chi1chi2=True)
    ho = od.gw_strain(owf, ev, fgrids, rot=60.)
    assert np.max(np.abs(h - ho)) <= 1e-6 * np.max(np.abs(ho))
    Fs, Ds = sig.FisherMatr(dict(ev), res=res, return_derivatives=True)
    Do = od.signal_derivatives(owf, ev, fgrids, rot=0.)
    Do[7] /= 86400.
    scale = np.max(np.abs(Do), axis=(1, 2), keepdims=True)
    assert np.max(np.abs(Ds[0] - Do) / scale) <= 1e-5
    Ffused = sig.FisherMatr(dict(ev), res=res, return_all=True)
    for k in range(3):
        assert norm_err(Fs[k], Ffused[k]) <= 1e-9


@pytest.mark.parametrize('model', ['TF2', 'PD'])
def test_waveform_methods(gpu, model):
    wf, owf = _models(model)
    ev = bbh(5, seed=12)
    od = oracle_net(NET_C0, False)[0]
    f, _ = od.grid(owf, ev, 256)
    for meth in ('Ampl', 'tau_star'):
        g = getattr(wf, meth)(f, **ev)
        o = getattr(owf, meth)(f, ev)
        nz = o != 0
        assert np.all(g[~nz] == 0), meth  # IMRPhenomD is cut at Mf = 0.2 (waveforms.py:1250)
        assert np.max(np.abs(g[nz] - o[nz]) / np.abs(o[nz])) <= 1e-11, meth
    g = wf.Phi(f, **ev)
    o = owf.Phi(f, ev)
    assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o)) + 1e-9
    assert np.allclose(wf.fcut(**ev), owf.fcut(ev), rtol=1e-14, atol=0)


def test_covariance_parity(gpu):
    from gwfast_amd import fisherTools as FT
    from oracle import covariance as OC
    wf, _ = _models('PD')
    ev = bbh(12, seed=13)
    net = hip_net(wf, NET_C1, True)
    F = net.FisherMatr(dict(ev), res=500)
    C, eps, cond, meth = FT.CovMatr(F, return_details=True)
    Co, epso, metho = OC.cov_matr(F)
    # compare marginalised errors sigma_i = sqrt(Cov_ii) where the reference used cholesky
    ok = (meth == 0) & (metho == 0)
    assert ok.sum() >= 8
    s = np.sqrt(np.abs(np.einsum('iin->in', C)))[:, ok]
    so = np.sqrt(np.abs(np.einsum('iin->in', Co)))[:, ok]
    rel = np.abs(s - so) / so
    print('sigma rel err max', rel.max(), 'cond range', cond.min(), cond.max(), flush=True)
    assert np.max(rel / np.maximum(1.0, cond[ok] * 1e-13)) <= 1e-6
    assert np.all(np.isfinite(eps))
    # inversion error agrees with the standalone computation
    eps2 = FT.compute_inversion_error(F, C)
    assert np.allclose(eps, eps2, rtol=1e-12, atol=0)


def test_check_fisher_parity(gpu):
    from gwfast_amd import fisherTools as FT
    from oracle import covariance as OC
    wf, _ = _models('TF2')
    ev = bbh(6, seed=14)
    net = hip_net(wf, NET_C0, False)
    F = net.FisherMatr(dict(ev), res=400)
    ev_, vec, cond = FT.CheckFisher(F)
    evo, veco, condo = OC.check_fisher(F)
    lam_max = np.max(np.abs(evo), axis=1, keepdims=True)
    assert np.max(np.abs(ev_ - evo) / lam_max) <= 1e-12
    # F v = lambda v for every returned pair
    for k in range(F.shape[-1]):
        Fk = F[:, :, k]
        r = Fk @ vec[k] - vec[k] * ev_[k][None, :]
        assert np.max(np.abs(r)) <= 1e-10 * lam_max[k, 0]


def test_cov_failure_modes(gpu):
    """All-NaN input -> NaN; indefinite -> eigen fallback; zero diagonal with a
    singular matrix -> NaN (fisherTools.py:63-66, 98-105, 183-187)."""
    from gwfast_amd import fisherTools as FT
    P = 11
    F = np.zeros((P, P, 3))
    F[:, :, 0] = np.nan
    A = np.diag(np.arange(1.0, P + 1))
    A[0, 1] = A[1, 0] = 5.0  # indefinite 2x2 block
    F[:, :, 1] = A
    F[:, :, 2] = np.diag(np.r_[0.0, np.ones(P - 1)])
    C, eps, cond, meth = FT.CovMatr(F, return_details=True)
    assert np.all(np.isnan(C[:, :, 0])) and meth[0] == -1
    assert meth[1] == 1 and np.allclose(C[:, :, 1] @ A, np.eye(P), atol=1e-12)
    assert np.all(np.isnan(C[:, :, 2])) and meth[2] == -1


def test_full_size_properties(gpu):
    """Config-1 shapes (1e4 events, 2000 frequencies, 5 interferometers):
    symmetry, the closed forms F[dL,dL] = SNR^2/dL^2 and F[Phic,Phic] = SNR^2
    (signal.py: the strain is linear in 1/dL and exp(-i Phic)), and bitwise
    determinism of two runs."""
    import torch
    from gwfast_amd import _engine as E
    wf, _ = _models('PD')
    n, res = 10000, 2000
    ev = bbh(n, seed=21)
    net = hip_net(wf, NET_C1, True)
    sig0 = next(iter(net.signals.values()))
    sig0._complete(ev)
    rows = E.to_dev(E.param_rows(ev, False))
    F1 = net.fisher_device(rows, res)
    F2 = net.fisher_device(rows, res)
    snr = net.snr_device(rows, res)
    torch.cuda.synchronize()
    assert torch.equal(F1, F2)
    F = F1.cpu().numpy()
    snr = snr.cpu().numpy()
    assert np.all(np.isfinite(F))
    assert np.max(np.abs(F - F.transpose(1, 0, 2))) == 0.0
    dL = ev['dL']
    assert np.max(np.abs(F[2, 2] * dL ** 2 / snr ** 2 - 1.0)) <= 1e-10
    assert np.max(np.abs(F[8, 8] / snr ** 2 - 1.0)) <= 1e-10
    assert np.max(np.abs(F[2, 8]) / np.sqrt(F[2, 2] * F[8, 8])) <= 1e-12


def test_ragged_grids_match_uniform(gpu):
    """Per-event grid lengths in one persistent launch give, for every event,
    the Fisher and SNR of a uniform launch at that event's length (the same
    arithmetic per event, in separately compiled kernels whose multiply-adds
    may contract differently: equal to rounding, <= 1e-14 normalised)."""
    wf, _ = _models('TF2', is_tidal=True, which_ISCO='Kerr')
    ev = bns(24, seed=31)
    rng = np.random.default_rng(0)
    res = rng.integers(500, 3000, size=24).astype(np.int32)
    net = hip_net(wf, NET_C1, True)
    F = net.FisherMatr(dict(ev), res=res)
    snr = net.SNR(dict(ev), res=res)
    for r in np.unique(res)[:6]:
        sel = np.where(res == r)[0]
        sub = {k: v[sel] for k, v in ev.items()}
        Fu = net.FisherMatr(dict(sub), res=int(r))
        su = net.SNR(dict(sub), res=int(r))
        assert norm_err(F[:, :, sel], Fu) <= 1e-14
        assert np.max(np.abs(snr[sel] - su) / su) <= 1e-14


C3_LENS = np.array([500, 511, 512, 513, 575, 576, 577, 1000, 2048, 4095, 4096, 4097, 7777, 9999, 10000, 10000],
                   dtype=np.int32)


def test_c3_ragged_tidal_network_vs_oracle(gpu):
    """Config C3 on its own workload: tidal TaylorF2 (13 parameters, Kerr
    ISCO cut: 1.3-4.2 kHz for this BNS population) on ETS + CE1Id + CE2NM
    with Earth rotation, one grid length per event from 500 to 10^4 points
    (64-point chunk edges included) in one launch; every event against the
    oracle run at its own length."""
    from oracle import detector as OD
    wf, owf = _models('TF2', is_tidal=True, which_ISCO='Kerr')
    n = len(C3_LENS)
    ev = bns(n, seed=41)
    net = hip_net(wf, NET_C1, True)
    odets = oracle_net(NET_C1, True)
    F = net.FisherMatr(dict(ev), res=C3_LENS)
    snr = net.SNR(dict(ev), res=C3_LENS)
    assert F.shape == (13, 13, n)
    worst = [0., 0., 0.]
    for e in range(n):
        one = {k: v[e:e + 1] for k, v in ev.items()}
        r = int(C3_LENS[e])
        so = OD.network_snr(odets, owf, one, r)
        Fd = OD.network_fisher(odets, owf, one, r, mode='dual')
        Ff = OD.network_fisher(odets, owf, one, r, mode='faithful')
        errs = (abs(snr[e] - so[0]) / so[0], norm_err(F[:, :, e:e + 1], Fd), norm_err(F[:, :, e:e + 1], Ff))
        worst = [max(a, b) for a, b in zip(worst, errs)]
        print('C3 event %d res %d: SNR %.2f rel err %.2e, Fisher vs dual %.2e vs faithful %.2e' % ((e, r, so[0]) + errs),
              flush=True)
    assert worst[0] <= 1e-10
    assert worst[1] <= 1e-10
    assert worst[2] <= 1e-6


def test_c3_full_size_properties(gpu):
    """Config C3 at full size: 10^5 tidal BNS events, grid lengths uniform in
    [500, 10^4], one ragged launch: finite, exactly symmetric,
    F[dL,dL] dL^2 / SNR^2 = 1 and F[Phic,Phic] / SNR^2 = 1 against the ragged
    SNR, F[dL,Phic] = 0, and bitwise determinism of two launches."""
    import torch
    from gwfast_amd import _engine as E
    wf, _ = _models('TF2', is_tidal=True, which_ISCO='Kerr')
    n = 100_000
    ev = bns(n, seed=11)
    res = np.random.default_rng(11).integers(500, 10_001, n).astype(np.int32)
    net = hip_net(wf, NET_C1, True)
    rows = E.to_dev(E.param_rows(ev, True))
    F1 = net.fisher_device(rows, res)
    F2 = net.fisher_device(rows, res)
    snr = net.snr_device(rows, res)
    torch.cuda.synchronize()
    assert torch.equal(F1, F2)
    F = F1.cpu().numpy()
    snr = snr.cpu().numpy()
    assert np.all(np.isfinite(F)) and np.all(snr > 0)
    assert np.max(np.abs(F - F.transpose(1, 0, 2))) == 0.0
    dL = ev['dL']
    e1 = np.max(np.abs(F[2, 2] * dL ** 2 / snr ** 2 - 1.0))
    e2 = np.max(np.abs(F[8, 8] / snr ** 2 - 1.0))
    e3 = np.max(np.abs(F[2, 8]) / np.sqrt(F[2, 2] * F[8, 8]))
    print('C3 10^5 events: dL identity %.2e, Phic identity %.2e, dL-Phic %.2e' % (e1, e2, e3), flush=True)
    assert e1 <= 1e-10 and e2 <= 1e-10 and e3 <= 1e-12


def test_ragged_chunk_edges(gpu):
    """Ragged grid lengths at the 64-point chunk edges of the producer /
    consumer pairs (2, 3, 63, 64, 65, 127, 128, 129 and 1000 points), more
    events than the resident pairs, IMRPhenomD on the 5-interferometer
    network: every event equals a uniform launch at its own length."""
    from gwfast_amd.waveforms import IMRPhenomD
    lens = np.array([2, 3, 63, 64, 65, 127, 128, 129, 1000], dtype=np.int32)
    n = 1500
    res = lens[np.arange(n) % len(lens)]
    ev = bbh(n, seed=33)
    net = hip_net(IMRPhenomD(), NET_C1, True)
    F = net.FisherMatr(dict(ev), res=res)
    assert np.all(np.isfinite(F))
    for r in lens:
        sel = np.where(res == r)[0][:40]
        sub = {k: v[sel] for k, v in ev.items()}
        Fu = net.FisherMatr(dict(sub), res=int(r))
        assert norm_err(F[:, :, sel], Fu) <= 1e-14, r  # separately compiled kernels: equal to rounding


# ------------------------------------------------------------ post-Fisher
def test_post_fisher_batch_vs_oracle(gpu):
    """fixParams / addPrior (bit-exact gathers), sigma and the 90% sky area
    against the oracle on a real network Fisher."""
    from gwfast_amd import fisherTools as FT
    from gwfast_amd.waveforms import IMRPhenomD
    from oracle import covariance as OC
    ev = bbh(16, seed=71)
    wf = IMRPhenomD()
    net = hip_net(wf, NET_C1, True)
    F = net.FisherMatr(dict(ev), res=1000)
    Fr, nums = FT.fixParams(F, wf.ParNums, ['Phicoal', 'psi'])
    Fo, nums_o = OC.fix_params(F, wf.ParNums, ['Phicoal', 'psi'])
    assert nums == nums_o and np.array_equal(Fr, Fo)
    Fp = FT.addPrior(F, [1.0, 4.0], wf.ParNums, ['chi2z', 'chi1z'])
    assert np.array_equal(Fp, OC.add_prior(F, [1.0, 4.0], wf.ParNums, ['chi2z', 'chi1z']))
    C, _ = FT.CovMatr(Fr)
    area = FT.compute_localization_region(C, nums, ev['theta'])
    area_o = OC.localization_region(C, nums, ev['theta'])
    assert np.max(np.abs(area / area_o - 1)) < 1e-13
    sig, _ = FT.sigma_sky_area_device(FT._dev(C), ev['theta'], nums['theta'], nums['phi'])
    assert np.allclose(sig.cpu().numpy(), np.sqrt(np.einsum('iin->in', C)), rtol=1e-15, atol=0)
    assert FT.compute_localization_region(C, nums, ev['theta'], units='bogus') is None


def test_catalogue_forecast_files_and_resume(gpu, tmp_path):
    """The driver's chain and file set (calculate_forecasts_from_catalog.py:
    222-247, 421-630, 738-741) on 300 events in 3 chunks, against the same
    chain called directly."""
    from gwfast_amd import forecast as FC, fisherTools as FT
    from gwfast_amd.waveforms import IMRPhenomD
    ev = bbh(300, seed=81)
    ev['dL'] = ev['dL'] * 4.0   # push part of the catalogue below SNR 12
    wf = IMRPhenomD()
    net = hip_net(wf, NET_C1, True)
    suffs = FC.run_catalog(ev, net, str(tmp_path), batch_size=100, params_fix=['Phicoal'], res=500, verbose=False)
    assert suffs == ['_0_to_100', '_100_to_200', '_200_to_300']
    snr = net.SNR(dict(ev), res=500)
    det = np.nonzero(snr > 12.0)[0]
    assert 0 < len(det) < 300
    have = [s for s in suffs if (tmp_path / ('fishers%s.npy' % s)).exists()]
    got_idx = np.concatenate([np.atleast_1d(np.loadtxt(tmp_path / ('idxs_det%s.txt' % s))) for s in have])
    assert np.array_equal(got_idx.astype(int), det)
    snrs = np.concatenate([np.loadtxt(tmp_path / ('snrs%s.txt' % s)) for s in suffs])
    assert np.allclose(snrs, snr, rtol=1e-14, atol=0)
    F = np.concatenate([np.load(tmp_path / ('fishers%s.npy' % s)) for s in have], axis=2)
    Fr, nums = FT.fixParams(net.FisherMatr({k: v[det] for k, v in ev.items()}, res=500), wf.ParNums, ['Phicoal'])
    assert F.shape == (10, 10, len(det)) and np.array_equal(F, Fr)
    C = np.concatenate([np.load(tmp_path / ('covs%s.npy' % s)) for s in have], axis=2)
    Cr, eps = FT.CovMatr(Fr)
    assert np.array_equal(C, Cr)
    err = np.concatenate([np.loadtxt(tmp_path / ('errors%s.txt' % s)).reshape(10, -1) for s in have], axis=1)
    assert np.allclose(err, np.sqrt(np.einsum('iin->in', Cr)), rtol=1e-15)
    area = np.concatenate([np.atleast_1d(np.loadtxt(tmp_path / ('sky_area%s.txt' % s))) for s in have])
    assert np.allclose(area, FT.compute_localization_region(Cr, nums, ev['theta'][det]), rtol=1e-14)
    # resume: an existing chunk is skipped (its file is left untouched)
    p = tmp_path / 'snrs_100_to_200.txt'
    p.write_text('sentinel\n')
    FC.run_catalog(ev, net, str(tmp_path), batch_size=100, params_fix=['Phicoal'], res=500, resume=True,
                   verbose=False)
    assert p.read_text() == 'sentinel\n'


def test_catalogue_forecast_return_all(gpu, tmp_path):
    """--return_all (calculate_forecasts_from_catalog.py:421, 507-513,
    578-590, 337-380): per-arm SNRs of every event and per-arm Fisher
    matrices (params_fix applied) of the detected ones, keyed as DetNet's
    return_all; their network sums equal snrs / fishers of the same run and
    DetNet.FisherMatr(return_all=True); with a duty factor the arms carry the
    driver's masks."""
    from gwfast_amd import forecast as FC, fisherTools as FT
    from gwfast_amd.waveforms import IMRPhenomD
    ev = bbh(120, seed=83)
    ev['dL'] = ev['dL'] * 4.0
    wf = IMRPhenomD()
    net = hip_net(wf, NET_C1, True)
    suffs = FC.run_catalog(ev, net, str(tmp_path), batch_size=60, params_fix=['Phicoal'], res=500, verbose=False,
                           return_all=True)
    keys = ['ETS_0', 'ETS_1', 'ETS_2', 'CE1Id', 'CE2NM', 'net']
    snr_all = net.SNR(dict(ev), res=500, return_all=True)
    for s, (i0, i1) in zip(suffs, ((0, 60), (60, 120))):
        A = np.load(tmp_path / ('all_snrs%s.npz' % s))
        assert sorted(A.files) == sorted(keys)
        for k in keys:
            want = snr_all[k][i0:i1] if k != 'net' else snr_all['net'][i0:i1]
            assert np.allclose(A[k], want, rtol=1e-13, atol=0), k
        assert np.allclose(A['net'], np.loadtxt(tmp_path / ('snrs%s.txt' % s)), rtol=1e-15, atol=0)
        fp = tmp_path / ('all_fishers%s.npz' % s)
        if not fp.exists():
            continue
        B = np.load(fp)
        det = np.atleast_1d(np.loadtxt(tmp_path / ('idxs_det%s.txt' % s))).astype(int)
        assert B['net'].shape == (10, 10, len(det))
        assert np.allclose(B['net'], sum(B[k] for k in keys[:-1]), rtol=1e-14, atol=0)
        assert norm_err(B['net'], np.load(tmp_path / ('fishers%s.npy' % s))) == 0.0
        Fd = net.FisherMatr({k: v[det] for k, v in ev.items()}, res=500, return_all=True)
        for k in keys:
            Fr, _ = FT.fixParams(Fd[k], wf.ParNums, ['Phicoal'])
            assert norm_err(B[k], Fr) <= 1e-13, k
    # duty cycle: masked arms
    seeds = [11, 12, 13]
    FC.run_catalog(ev, net, str(tmp_path / 'df'), batch_size=120, res=500, verbose=False, return_all=True,
                   duty_factor=0.6, seeds=seeds)
    masks = FC.duty_masks(net, seeds, 0.6, 120)
    A = np.load(tmp_path / 'df' / 'all_snrs_0_to_120.npz')
    for k in keys[:-1]:
        assert np.allclose(A[k], snr_all[k] * masks[k], rtol=1e-13, atol=0), k
    B = np.load(tmp_path / 'df' / 'all_fishers_0_to_120.npz')
    det = np.atleast_1d(np.loadtxt(tmp_path / 'df' / 'idxs_det_0_to_120.txt')).astype(int)
    for k in keys[:-1]:
        zero = masks[k][det] == 0
        assert np.all(B[k][:, :, zero] == 0), k


def test_nrtidal_waveform_methods(gpu):
    wf, owf = _models('NRT')
    ev = bns(5, seed=13)
    od = oracle_net(NET_C0, False)[0]
    f, _ = od.grid(owf, ev, 256)
    for meth in ('Ampl', 'tau_star'):
        g = getattr(wf, meth)(f, **ev)
        o = getattr(owf, meth)(f, ev)
        nz = o != 0
        assert np.all(g[~nz] == 0), meth
        assert np.max(np.abs(g[nz] - o[nz]) / np.abs(o[nz])) <= 1e-11, meth
    g = wf.Phi(f, **ev)
    o = owf.Phi(f, ev)
    assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o)) + 1e-9
    assert np.allclose(wf.fcut(**ev), owf.fcut(ev), rtol=1e-14, atol=0)


def test_nrtidal_gw170817_known_answers(gpu):
    """gwfast_tutorial.ipynb cells 12 and 22 on the GPU: LIGO-Virgo O2 SNR of
    GW170817 '33.16' and 90% sky area after fixing deltaLambda '19.9 deg^2'."""
    import os
    from gwfast_amd import fisherTools as FT
    from gwfast_amd.network import DetNet
    from gwfast_amd.signal import GWSignal
    from gwfast_amd.waveforms import IMRPhenomD_NRTidalv2
    from gwfast_amd import gwfastGlobals as glob
    from tests.test_oracle_pins import GOLDEN_PSD, O2_PSDS, gw170817_tutorial
    wf = IMRPhenomD_NRTidalv2()
    sigs = {n: GWSignal(wf, psd_path=os.path.join(GOLDEN_PSD, O2_PSDS[n]), detector_shape=glob.detectors[n]['shape'],
                        det_lat=glob.detectors[n]['lat'], det_long=glob.detectors[n]['long'],
                        det_xax=glob.detectors[n]['xax'], useEarthMotion=False, fmin=10., verbose=False)
            for n in ('L1', 'H1', 'Virgo')}
    net = DetNet(sigs, verbose=False)
    ev = gw170817_tutorial()
    snr = net.SNR(dict(ev))
    assert abs(snr[0] - 33.16) < 0.005
    F = net.FisherMatr(dict(ev))
    Ff, nums = FT.fixParams(F, wf.ParNums, ['deltaLambda'])
    C, _ = FT.CovMatr(Ff)
    area = FT.compute_localization_region(C, nums, ev['theta'])
    assert abs(area[0] - 19.9) < 0.05
