"""Pins of the CPU oracle (no GPU): reference-held known answers, textbook
formulas, geometric identities and derivative cross-checks.  Each test is
chosen so that a plausible transcription mistake in the oracle (dropped term,
wrong sign, swapped operand) fails it.
"""
import numpy as np
import pytest
from scipy.optimize import minimize

from gwfast_amd import gwfastGlobals as glob
from oracle import dual as D
from oracle import detector as OD
from oracle import waveforms as OW


def gw170817_like():
    # notebooks/new_features_tutorial.ipynb:72-84 (dL from Planck18 is irrelevant for fcut)
    z = np.array([0.00980])
    return {'Mc': np.array([1.1859]) * (1. + z), 'dL': np.array([0.04]), 'theta': np.array([np.pi / 2. + 0.408084]),
            'phi': np.array([3.44616]), 'iota': np.array([2.545065595974997]), 'psi': np.array([0.]),
            'tcoal': np.array([0.]), 'eta': np.array([0.24786618323504223]), 'Phicoal': np.array([0.]),
            'chi1z': np.array([0.005136138323169717]), 'chi2z': np.array([0.003235146993487445]),
            'Lambda1': np.array([368.17802383555687]), 'Lambda2': np.array([586.5487031450857])}


# ------------------------------------------------------------ known answers
def test_tf2_fcut_known_answers():
    """new_features_tutorial.ipynb:115: 'Schwarzschild ... 1590.1 Hz',
    'Kerr ... 3433.5 Hz' (printed with %.1f)."""
    ev = gw170817_like()
    assert '%.1f' % OW.TaylorF2(is_tidal=True).fcut(ev)[0] == '1590.1'
    assert '%.1f' % OW.TaylorF2(is_tidal=True, which_ISCO='Kerr').fcut(ev)[0] == '3433.5'


def test_et_optimal_location_known_answer():
    """gwfast_tutorial.ipynb:662-671: DetNet({ETS}).optimal_location(0.) =
    [0.8636426, 0.1643505]; the network version sums Fp^2 + Fc^2 over the
    three arms at psi = 0 and minimises -sqrt of it from (1, 1) within bounds
    (network.py:174-196)."""
    d = glob.detectors['ETS']
    det = OD.Detector(glob.PSD_ET, 'T', d['lat'], d['long'], d['xax'])

    def negF(p):
        th, ph = p
        s = 0.
        for i in range(3):
            Fp, Fc = det.pattern_function(th, ph, 0., 0., rot=i * 60.)
            s += Fp ** 2 + Fc ** 2
        return -np.sqrt(s)

    x = minimize(negF, [1., 1.], bounds=((0., np.pi), (0., 2. * np.pi))).x
    assert np.allclose(x, [0.8636426, 0.1643505], atol=2e-6)


# ------------------------------------------------------------ TaylorF2 phase
def tf2_textbook_nonspinning(f, Mc, eta):
    """3.5PN TaylorF2 phase of arXiv:0907.0700 eq. (3.18) with (3.3) coefficients,
    written with canonical fractions, reference frequency at v = 1 (so the
    v_lso logarithm becomes ln v) and the constant pi - pi/4 of the
    reference's LAL convention."""
    g = np.euler_gamma
    M = Mc / eta ** 0.6 * glob.GMsun_over_c3
    v = (np.pi * M * f) ** (1. / 3.)
    a = [1.0, 0.0,
         3715. / 756. + 55. / 9. * eta,
         -16. * np.pi,
         15293365. / 508032. + 27145. / 504. * eta + 3085. / 72. * eta ** 2,
         np.pi * (38645. / 756. - 65. / 9. * eta) * (1. + 3. * np.log(v)),
         (11583231236531. / 4694215680. - 640. / 3. * np.pi ** 2 - 6848. / 21. * g - 6848. / 21. * np.log(4. * v)
          + (-15737765635. / 3048192. + 2255. / 12. * np.pi ** 2) * eta + 76055. / 1728. * eta ** 2
          - 127825. / 1296. * eta ** 3),
         np.pi * (77096675. / 254016. + 378515. / 1512. * eta - 74045. / 756. * eta ** 2)]
    s = sum(a[k] * v ** k for k in range(8))
    return 3. / (128. * eta * v ** 5) * s + np.pi - np.pi / 4.


def test_tf2_phase_matches_textbook_nonspinning():
    f = np.geomspace(2., 300., 50)[:, None]
    ev = {'Mc': np.array([1.2, 10., 30.]), 'eta': np.array([0.2499, 0.2, 0.12]),
          'chi1z': np.zeros(3), 'chi2z': np.zeros(3)}
    ph = OW.TaylorF2().Phi(f, ev)
    ref = tf2_textbook_nonspinning(f, ev['Mc'], ev['eta'])
    assert np.max(np.abs(ph - ref) / np.abs(ref)) < 1e-13


def test_tf2_spin_orbit_15pn():
    """The aligned-spin 1.5PN coefficient is
    113/3 delta chi_a + (113/3 - 76 eta/3) chi_s (arXiv:1107.1267 eq. 5.3,
    delta = sqrt(1 - 4 eta)).  The spin-dependent part of the phase times
    128 eta v^2 / 3 lies exactly in span{1, v, v^2, v^2 ln v, v^3, v^4}
    (1.5 to 3.5PN), so a least-squares fit returns the v^0 coefficient."""
    eta, chi1, chi2 = 0.2, 0.6, -0.3
    wf = OW.TaylorF2()
    Mc = 10.
    M = Mc / eta ** 0.6 * glob.GMsun_over_c3
    vs = np.linspace(0.05, 0.4, 25)
    f = (vs ** 3 / (np.pi * M))[:, None]
    ev1 = {'Mc': np.array([Mc]), 'eta': np.array([eta]), 'chi1z': np.array([chi1]), 'chi2z': np.array([chi2])}
    ev0 = dict(ev1, chi1z=np.zeros(1), chi2z=np.zeros(1))
    v = (np.pi * M * f[:, 0]) ** (1. / 3.)
    d = (wf.Phi(f, ev1) - wf.Phi(f, ev0))[:, 0] * 128. * eta * v ** 2 / 3.
    B = np.stack([np.ones_like(v), v, v ** 2, v ** 2 * np.log(v), v ** 3, v ** 4], axis=1)
    coef = np.linalg.lstsq(B, d, rcond=None)[0]
    chi_s, chi_a, delta = 0.5 * (chi1 + chi2), 0.5 * (chi1 - chi2), np.sqrt(1 - 4 * eta)
    expect = 113. / 3. * delta * chi_a + (113. / 3. - 76. * eta / 3.) * chi_s
    assert abs(coef[0] - expect) < 1e-8 * abs(expect)
    # 2.5PN: the v^2 ln v coefficient is 3 x (minus the spin-orbit part of five)
    so5 = -((732985. / 2268. - 24260. * eta / 81. - 340. * eta ** 2 / 9.) * chi_s
            + (732985. / 2268. + 140. * eta / 9.) * delta * chi_a)
    assert abs(coef[3] - 3 * so5) < 1e-6 * abs(so5)


def test_tau_star_is_spa_time_of_phase():
    """Stationary phase: with h ~ exp(i (2 pi f tc - Phi_GW)) (signal.py:483),
    -dPhi_GW/df = 2 pi tau(f) up to terms beyond 3.5PN; at v ~ 0.05 the
    mismatch is < 1e-6 relative for non-spinning TaylorF2."""
    wf = OW.TaylorF2()
    ev = {'Mc': np.array([5.]), 'eta': np.array([0.22]), 'chi1z': np.zeros(1), 'chi2z': np.zeros(1)}
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    f0 = (0.05 ** 3 / (np.pi * M))
    h = f0 * 1e-5
    f = np.array([f0 - h, f0 + h])[:, None] * np.ones((1, 1))
    dphi = (wf.Phi(f[1:], ev) - wf.Phi(f[:1], ev)) / (2 * h)
    tau = wf.tau_star(np.array([[f0[0]]]), ev)
    assert abs(-dphi[0, 0] / (2 * np.pi * tau[0, 0]) - 1.) < 1e-6


# ----------------------------------------------------------- IMRPhenomD
def _pd_ev(n=4, seed=1):
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(n, seed=seed)
    return ev


def test_phenomd_remnant_equal_mass_nonspinning():
    """Equal-mass non-spinning remnant: a_f = 0.6864, E_rad = 0.0484 of M
    (arXiv:1508.07250 fits, waveforms.py:1256-1297)."""
    wf = OW.IMRPhenomD()
    eta = np.array([0.2499999])
    z = np.zeros(1)
    assert abs(wf._finalspin(eta, z, z)[0] - 0.6864) < 5e-4
    assert abs(wf._radiatednrg(eta, z, z)[0] - 0.0484) < 5e-4


def test_phenomd_newtonian_amplitude_matches_taylorf2():
    """At low frequency IMRPhenomD's amplitude tends to the Newtonian one, which
    is TaylorF2's amplitude; the two differ only by the two Gpc conversions
    (gwfastGlobals.py:68 vs 111, ~9e-8 relative)."""
    ev = _pd_ev()
    f = np.array([[1e-3]]) * np.ones((1, 4))
    r = OW.IMRPhenomD().Ampl(f, ev) / OW.TaylorF2().Ampl(f, ev)
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    assert np.all(np.abs(r - 1.) < 1e-6 + 50 * (M * 1e-3) ** (2. / 3.))


def _pd_region_edges(wf, ev):
    M = ev['Mc'] / ev['eta'] ** 0.6
    fring, fdamp = wf._ringdown(ev['eta'], ev['chi1z'], ev['chi2z'])
    return M * glob.GMsun_over_c3, fring


@pytest.mark.parametrize('seed', [1, 2])
def test_phenomd_phase_c1_at_joins(seed):
    """Phase is C1 at Mf = 0.018 and Mf = fring/2 (waveforms.py:1098-1129)."""
    wf = OW.IMRPhenomD(fRef=10.)
    ev = _pd_ev(4, seed)
    MG, fring = _pd_region_edges(wf, ev)
    for xj in (0.018 * np.ones(4), 0.5 * fring):
        fj = xj / MG
        eps = 1e-7
        f = np.array([fj * (1 - 2 * eps), fj * (1 - eps), fj * (1 + eps), fj * (1 + 2 * eps)])
        p = wf.Phi(f, ev)
        left_slope = (p[1] - p[0]) / (f[1] - f[0])
        right_slope = (p[3] - p[2]) / (f[3] - f[2])
        jump = (p[2] - p[1]) - left_slope * (f[2] - f[1])
        scale = np.abs(left_slope * fj)
        assert np.all(np.abs(jump) < 1e-6 * scale)
        assert np.all(np.abs(left_slope - right_slope) < 1e-4 * np.abs(left_slope))


@pytest.mark.parametrize('seed', [1, 2, 3])
def test_phenomd_amplitude_c1_at_joins(seed):
    """Amplitude is C1 at Mf = 0.014 and at fpeak: the intermediate quartic
    matches value and slope of the inspiral and merger-ringdown models there
    (collocation, waveforms.py:1199-1245)."""
    wf = OW.IMRPhenomD()
    ev = _pd_ev(4, seed)
    M = ev['Mc'] / ev['eta'] ** 0.6
    MG = M * glob.GMsun_over_c3
    fring, fdamp = wf._ringdown(ev['eta'], ev['chi1z'], ev['chi2z'])
    eta = ev['eta']
    xi = -1.0 + (0.5 * (ev['chi1z'] + ev['chi2z']) * (1.0 - eta * 76.0 / 113.0)
                 + np.sqrt(1 - 4 * eta) * 0.5 * (ev['chi1z'] - ev['chi2z']))
    g2 = OW._fit('gamma2', eta, eta * eta, xi)
    g3 = OW._fit('gamma3', eta, eta * eta, xi)
    fpeak = np.where(g2 >= 1.0, np.abs(fring - fdamp * g3 / g2),
                     fring + fdamp * (-1.0 + np.sqrt(1.0 - g2 * g2)) * g3 / g2)
    for xj in (0.014 * np.ones(4), fpeak):
        fj = xj / MG
        eps = 1e-7
        f = np.array([fj * (1 - 2 * eps), fj * (1 - eps), fj * (1 + eps), fj * (1 + 2 * eps)])
        a = wf.Ampl(f, ev) * f ** (7. / 6.)  # remove the smooth f^-7/6
        ls = (a[1] - a[0]) / (f[1] - f[0])
        rs = (a[3] - a[2]) / (f[3] - f[2])
        assert np.all(np.abs((a[2] - a[1]) - ls * (f[2] - f[1])) < 1e-7 * np.abs(a[1]))
        # one-sided difference quotients sit 1.5 eps from the join, so their
        # difference carries a curvature term ~3 eps a'' f; a C1 defect (wrong
        # d1/d2 or delta_i) shows up at O(0.1 |a|/f)
        assert np.all(np.abs(ls - rs) < 1e-4 * np.abs(a[1] / fj))


# ------------------------------------------------------------- projection
def _tensor_pattern(lat, lon, xax_deg, zeta, ra, dec, t):
    """F+^2 + Fx^2 from detector tensor D = (u u - v v)/2 contracted with the
    TT polarisation tensors, arms at xax -/+ zeta/2 measured counter-clockwise
    from local East (the angle convention of signal.py:114-116)."""
    L = lon + 2 * np.pi * t
    E_ = np.array([-np.sin(L), np.cos(L), 0.])
    N_ = np.array([-np.sin(lat) * np.cos(L), -np.sin(lat) * np.sin(L), np.cos(lat)])
    b = np.deg2rad(xax_deg)
    # bisector at angle xax from East towards North; arms at -/+ zeta/2 about it
    u = np.cos(b - zeta / 2) * E_ + np.sin(b - zeta / 2) * N_
    v = np.cos(b + zeta / 2) * E_ + np.sin(b + zeta / 2) * N_
    Dt = 0.5 * (np.outer(u, u) - np.outer(v, v))
    n = np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])
    m = np.cross(n, [0., 0., 1.])
    m /= np.linalg.norm(m)
    k = np.cross(n, m)
    ep = np.outer(m, m) - np.outer(k, k)
    ec = np.outer(m, k) + np.outer(k, m)
    return np.sum(Dt * ep) ** 2 + np.sum(Dt * ec) ** 2


@pytest.mark.parametrize('name', ['Virgo', 'ETS', 'CE1Id', 'H1'])
def test_pattern_functions_vs_detector_tensor(name):
    d = glob.detectors[name]
    det = OD.Detector(glob.PSD_ET, d['shape'], d['lat'], d['long'], d['xax'])
    zeta = det.angbtwArms
    rng = np.random.default_rng(4)
    for _ in range(20):
        th, ph, t, psi = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), rng.uniform(), rng.uniform(0, np.pi)
        Fp, Fc = det.pattern_function(th, ph, t, psi)
        ref = _tensor_pattern(det.det_lat_rad, det.det_long_rad, d['xax'], zeta, ph, 0.5 * np.pi - th, t)
        assert abs(Fp ** 2 + Fc ** 2 - ref) < 1e-12


def test_pattern_identities():
    d = glob.detectors['ETS']
    det = OD.Detector(glob.PSD_ET, 'T', d['lat'], d['long'], d['xax'])
    th, ph, t = 1.1, 4.0, 0.3
    s = [det.pattern_function(th, ph, t, 0.7, rot=r) for r in (0., 60., 120.)]
    assert abs(sum(x[0] for x in s)) < 1e-15 and abs(sum(x[1] for x in s)) < 1e-15
    Fp, Fc = det.pattern_function(th, ph, t, 0.7)
    Fp2, Fc2 = det.pattern_function(th, ph, t, 0.7 + np.pi / 2)
    assert abs(Fp + Fp2) < 1e-15 and abs(Fc + Fc2) < 1e-15
    for psi in (0.1, 1.3, 2.9):
        a, b = det.pattern_function(th, ph, t, psi)
        assert abs(a * a + b * b - (Fp * Fp + Fc * Fc)) < 1e-15


def test_delt_loc_unit_vectors():
    d = glob.detectors['CE2NM']
    det = OD.Detector(glob.PSD_CE, 'L', d['lat'], d['long'], d['xax'])
    rng = np.random.default_rng(5)
    for _ in range(10):
        th, ph, t = np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), rng.uniform()
        dec = 0.5 * np.pi - th
        n = np.array([np.cos(dec) * np.cos(ph), np.cos(dec) * np.sin(ph), np.sin(dec)])
        L = det.det_long_rad + 2 * np.pi * t
        r = np.array([np.cos(det.det_lat_rad) * np.cos(L), np.cos(det.det_lat_rad) * np.sin(L), np.sin(det.det_lat_rad)])
        assert abs(det.delt_loc(th, ph, t) - (-glob.REarth * n @ r / glob.clight)) < 1e-17
    # source at the detector's zenith: -R/c
    th = 0.5 * np.pi - det.det_lat_rad
    assert abs(det.delt_loc(th, det.det_long_rad, 0.) + glob.REarth / glob.clight) < 1e-15


# ------------------------------------------------------------ derivatives
def test_dual_vs_complex_step_taylorf2():
    """TaylorF2 is holomorphic (waveforms.py:741): complex-step derivatives of
    the phase and amplitude are exact to rounding."""
    wf = OW.TaylorF2()
    ev = {'Mc': np.array([3., 25.]), 'eta': np.array([0.24, 0.16]), 'dL': np.array([1., 2.]),
          'chi1z': np.array([0.3, -0.5]), 'chi2z': np.array([-0.2, 0.7])}
    f = np.geomspace(3., 200., 40)[:, None] * np.ones((1, 2))
    keys = ['Mc', 'eta', 'chi1z', 'chi2z']
    evd = dict(ev)
    for k, key in enumerate(keys):
        evd[key] = D.Dual.seed(ev[key], k, 4)
    ph = wf.Phi(f, evd)
    am = wf.Ampl(f, evd)
    for k, key in enumerate(keys):
        h = 1e-30
        evc = {kk: v.astype(complex) for kk, v in ev.items()}
        evc[key] = evc[key] + 1j * h
        dph = np.imag(wf.Phi(f, evc)) / h
        dam = np.imag(wf.Ampl(f, evc)) / h
        assert np.max(np.abs(ph.d[k] - dph)) <= 1e-12 * np.max(np.abs(dph)) + 1e-300
        assert np.max(np.abs(am.d[k] - dam)) <= 1e-12 * np.max(np.abs(dam)) + 1e-300


def test_dual_vs_central_fd_phenomd():
    wf = OW.IMRPhenomD(fRef=5.)
    ev = _pd_ev(3, 6)
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    f = np.geomspace(0.002, 0.15, 60)[:, None] / M[None, :]
    keys = ['Mc', 'eta', 'chi1z', 'chi2z']
    evd = dict(ev)
    for k, key in enumerate(keys):
        evd[key] = D.Dual.seed(ev[key], k, 4)
    ph = wf.Phi(f, evd)
    am = wf.Ampl(f, evd)
    for k, key in enumerate(keys):
        h = 1e-7 * np.maximum(np.abs(ev[key]), 1e-2)
        evp, evm = dict(ev), dict(ev)
        evp[key] = ev[key] + h
        evm[key] = ev[key] - h
        dph = (wf.Phi(f, evp) - wf.Phi(f, evm)) / (2 * h)
        dam = (wf.Ampl(f, evp) - wf.Ampl(f, evm)) / (2 * h)
        assert np.max(np.abs(ph.d[k] - dph)) <= 2e-5 * np.max(np.abs(dph)), key
        assert np.max(np.abs(am.d[k] - dam)) <= 2e-5 * np.max(np.abs(dam)), key


@pytest.mark.parametrize('model,earth,shape', [('TF2', False, 'L'), ('PD', True, 'T'), ('PD', True, 'L')])
def test_closed_form_rows_vs_duals(model, earth, shape):
    """The reference's closed-form derivative rows (signal.py:1432-1577) equal
    the dual-number derivatives of GWstrain, up to the phase-association
    rounding of exp(i Psi) (|Psi| up to ~1e8 rad here)."""
    wf = OW.TaylorF2() if model == 'TF2' else OW.IMRPhenomD()
    name = 'Virgo' if shape == 'L' else 'ETS'
    d = glob.detectors[name]
    det = OD.Detector(glob.PSD_ET, shape, d['lat'], d['long'], d['xax'], useEarthMotion=earth)
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(3, seed=8)
    f, _ = det.grid(wf, ev, 200)
    a = det.signal_derivatives(wf, ev, f, rot=60. if shape == 'T' else 0.)
    b = det.full_dual_derivatives(wf, ev, f, rot=60. if shape == 'T' else 0.)
    for row in range(a.shape[0]):
        scale = np.max(np.abs(b[row]))
        assert np.max(np.abs(a[row] - b[row])) <= 2e-7 * scale, row


# ---------------------------------------------------------------- integrals
def test_fisher_closed_forms_oracle():
    """F[dL,dL] = SNR^2/dL^2, F[Phic,Phic] = SNR^2, F[dL,Phic] = 0
    (gwfast_tutorial.ipynb:296-304 checks the first at 2.2e-16)."""
    wf = OW.IMRPhenomD()
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(3, seed=9)
    d = glob.detectors['CE1Id']
    det = OD.Detector(glob.PSD_CE, 'L', d['lat'], d['long'], d['xax'], useEarthMotion=True)
    F = det.fisher(wf, ev, 300)
    snr = det.snr(wf, ev, 300)
    assert np.max(np.abs(F[2, 2] * ev['dL'] ** 2 / snr ** 2 - 1)) < 1e-12
    assert np.max(np.abs(F[8, 8] / snr ** 2 - 1)) < 1e-12
    assert np.max(np.abs(F[2, 8]) / np.sqrt(F[2, 2] * F[8, 8])) < 1e-12
    assert np.all(np.linalg.eigvalsh(F.transpose(2, 0, 1) / np.sqrt(np.einsum('iin->ni', F)[:, :, None] *
                                                                      np.einsum('iin->ni', F)[:, None, :])) > -1e-9)


def test_geomspace_jax_grid():
    g = OD.geomspace_jax(np.array([2., 3.]), np.array([100., 4000.]), 1000)
    assert g.shape == (1000, 2)
    assert np.allclose(g, np.geomspace([2., 3.], [100., 4000.], 1000), rtol=1e-13)
    assert np.all(np.diff(g, axis=0) > 0)


def test_jnp_interp_semantics():
    xp = np.array([0., 1., 3.])
    fp = np.array([1., 3., 7.])
    x = D.Dual.seed(np.array([-1., 0., 0.5, 1., 2., 3., 4.]), 0, 1)
    y = D.interp(x, xp, fp)
    assert np.allclose(y.v, [1., 1., 2., 3., 5., 7., 7.])
    # slope of the right-hand segment at a knot, of the last segment at the end, 0 outside
    assert np.allclose(y.d[0], [0., 2., 2., 2., 2., 2., 0.])


# ------------------------------------------------------------ post-Fisher
PARNUMS11 = {'Mc': 0, 'eta': 1, 'dL': 2, 'theta': 3, 'phi': 4, 'iota': 5, 'psi': 6, 'tcoal': 7, 'Phicoal': 8,
             'chi1z': 9, 'chi2z': 10}


def _spd(n, P, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(0, 1, (n, P, P))
    F = np.einsum('nij,nkj->ikn', A, A) + 0.5 * np.eye(P)[:, :, None]
    return F


def test_fix_params_is_conditional_covariance():
    """Fixing parameters = conditioning: inv(F_kept) = C_kk - C_kf C_ff^-1 C_fk."""
    from oracle.covariance import fix_params
    F = _spd(5, 11, 1)
    Fr, nums = fix_params(F, PARNUMS11, ['psi', 'Phicoal', 'tcoal'])
    keep = [i for i in range(11) if i not in (6, 7, 8)]
    assert nums == {'Mc': 0, 'eta': 1, 'dL': 2, 'theta': 3, 'phi': 4, 'iota': 5, 'chi1z': 6, 'chi2z': 7}
    for e in range(5):
        C = np.linalg.inv(F[:, :, e])
        f = [6, 7, 8]
        cond = C[np.ix_(keep, keep)] - C[np.ix_(keep, f)] @ np.linalg.solve(C[np.ix_(f, f)], C[np.ix_(f, keep)])
        assert np.allclose(np.linalg.inv(Fr[:, :, e]), cond, rtol=1e-9, atol=1e-12)


def test_add_prior_adds_diagonal_only():
    """Values land on the SORTED parameter indices (fisherTools.py:395-398
    sorts the indices, not the values)."""
    from oracle.covariance import add_prior
    F = _spd(3, 11, 2)
    G = add_prior(F, [1.0, 2.0], PARNUMS11, ['iota', 'eta'])
    D = G - F
    assert np.allclose(D[1, 1], 1.0) and np.allclose(D[5, 5], 2.0)
    D[1, 1] = D[5, 5] = 0
    assert np.all(D == 0)


def test_localization_region_closed_form():
    """Uncorrelated sky errors: area = -ln(1-p) 2 pi s_th s_ph |sin th|."""
    from oracle.covariance import localization_region
    C = np.zeros((11, 11, 2))
    C[3, 3] = [1e-4, 4e-6]
    C[4, 4] = [9e-4, 1e-6]
    th = np.array([0.3, 2.0])
    a = localization_region(C, PARNUMS11, th, 90, 'Sterad')
    assert np.allclose(a, np.log(10) * 2 * np.pi * np.sqrt(C[3, 3] * C[4, 4]) * np.abs(np.sin(th)), rtol=1e-14)
    assert np.allclose(localization_region(C, PARNUMS11, th), a * (180 / np.pi) ** 2, rtol=1e-14)


# ------------------------------------------------------ IMRPhenomD_NRTidalv2
GOLDEN_PSD = __import__('os').path.join(__import__('os').path.dirname(__file__), 'golden', 'psds')
O2_PSDS = {'L1': '2017-08-06_DCH_C02_L1_O2_Sensitivity_strain_asd.txt',
           'H1': '2017-06-10_DCH_C02_H1_O2_Sensitivity_strain_asd.txt',
           'Virgo': 'Hrec_hoft_V1O2Repro2A_16384Hz.txt'}


def gw170817_tutorial():
    """notebooks/gwfast_tutorial.ipynb cell 10 (printed values): Planck18
    dL(z=0.0098) = 0.04374755 Gpc and GMST(tGPS 1187008882.4) = 0.43432288 d."""
    z = np.array([0.00980])
    return {'Mc': np.array([1.1859]) * (1. + z), 'dL': np.array([0.04374755]),
            'theta': np.array([np.pi / 2. + 0.4080839999999999]), 'phi': np.array([3.4461599999999994]),
            'iota': np.array([2.545065595974997]), 'psi': np.array([0.]), 'tcoal': np.array([0.43432288]),
            'eta': np.array([0.24786618323504223]), 'Phicoal': np.array([0.]),
            'chi1z': np.array([0.005136138323169717]), 'chi2z': np.array([0.003235146993487445]),
            'Lambda1': np.array([368.17802383555687]), 'Lambda2': np.array([586.5487031450857])}


def _lv_o2_oracle():
    import os
    return [OD.Detector(os.path.join(GOLDEN_PSD, O2_PSDS[n]), glob.detectors[n]['shape'], glob.detectors[n]['lat'],
                        glob.detectors[n]['long'], glob.detectors[n]['xax'], useEarthMotion=False, fmin=10.)
            for n in ('L1', 'H1', 'Virgo')]


def test_nrtidal_gw170817_known_answers():
    """gwfast_tutorial.ipynb cells 12 and 22: LIGO-Virgo O2 network SNR of
    GW170817 with IMRPhenomD_NRTidalv2 '33.16', and its 90% sky area after
    fixing deltaLambda '19.9 deg^2'."""
    from oracle import covariance as OC
    ev = gw170817_tutorial()
    dets = _lv_o2_oracle()
    wf = OW.IMRPhenomD_NRTidalv2()
    snr = OD.network_snr(dets, wf, ev, 1000)
    assert abs(snr[0] - 33.16) < 0.005
    F = OD.network_fisher(dets, wf, ev, 1000, mode='faithful')
    PN = {'Mc': 0, 'eta': 1, 'dL': 2, 'theta': 3, 'phi': 4, 'iota': 5, 'psi': 6, 'tcoal': 7, 'Phicoal': 8,
          'chi1z': 9, 'chi2z': 10, 'LambdaTilde': 11, 'deltaLambda': 12}
    Ff, nums = OC.fix_params(F, PN, ['deltaLambda'])
    C, eps, _ = OC.cov_matr(Ff)
    area = OC.localization_region(C, nums, ev['theta'])
    assert abs(area[0] - 19.9) < 0.05


def test_nrtidal_reduces_to_phenomd_without_tides_and_spins():
    """Lambda = 0, chi = 0: NRTidalv2's phase is IMRPhenomD's and its
    amplitude is IMRPhenomD's below the taper (quadrupole moment 1, no tidal
    or spin-cubed terms)."""
    ev = _pd_ev(3, 5)
    ev['chi1z'] = ev['chi2z'] = np.zeros(3)
    ev['Lambda1'] = ev['Lambda2'] = np.zeros(3)
    nrt, pd = OW.IMRPhenomD_NRTidalv2(fRef=10.), OW.IMRPhenomD(fRef=10.)
    fm = nrt.fcut(ev) / 1.2
    f = np.geomspace(1e-2, 1., 50)[:, None] * fm[None, :] * 0.9
    assert np.allclose(nrt.Phi(f, ev), pd.Phi(f, ev), rtol=1e-13, atol=1e-9)
    assert np.allclose(nrt.Ampl(f, ev), pd.Ampl(f, ev), rtol=1e-14, atol=0)


def test_nrtidal_leading_tidal_phase_is_5pn_term():
    """At low frequency the tidal phase tends to the 5PN term
    -(117/256) (LambdaTilde/eta) v^5 (Flanagan & Hinderer 2008), whatever the
    mass ratio: kappa2T = 3/16 LambdaTilde."""
    from gwfast_amd import catalog
    ev = catalog.draw_bns(3, seed=4)
    ev['chi1z'] = ev['chi2z'] = np.zeros(3)  # quadrupole terms enter with chi^2
    wf = OW.IMRPhenomD_NRTidalv2()
    ev0 = dict(ev)
    ev0['Lambda1'] = ev0['Lambda2'] = np.zeros(3)
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    x = 1e-6
    f = np.array([[x]]) / M[None, :]
    d = wf.Phi(f, ev) - wf.Phi(f, ev0)
    Lt, _ = OW.lamt_dellam_from_lam12(ev['Lambda1'], ev['Lambda2'], ev['eta'])
    lead = -(117. / 256.) * Lt / ev['eta'] * (np.pi * x) ** (5. / 3.)
    # the phase's fRef reference point carries no tidal term; the Pade's
    # relative corrections are O((pi x)^(2/3)) ~ 2e-4 here
    assert np.allclose(d[0], lead, rtol=2e-3)


def test_planck_taper_values_and_y_derivative():
    y = np.array([0.05])
    xs = y * np.array([0.5, 1.0, 1.05, 1.1, 1.15, 1.2, 1.3])
    v = OW.planck_taper(xs, y)
    assert v[0] == 1.0 and v[1] == 1.0 and v[-1] == 0.0
    assert np.all(np.diff(v[1:5]) < 0)
    assert v[5] == 1.0  # the reference's expression at exactly 1.2 y (exp(+inf))
    assert OW.planck_taper(xs, y, end_inclusive=True)[5] == 0.0
    yd = D.Dual.seed(y, 0, 1)
    for x in (1.05, 1.1, 1.15):
        h = 1e-8 * y
        fd = (OW.planck_taper(x * y, y + h) - OW.planck_taper(x * y, y - h)) / (2 * h)
        an = OW.planck_taper(np.array([x * y[0]]), yd).d[0]
        assert np.allclose(an, fd, rtol=1e-6)


def test_nrtidal_quadrupole_continuous_and_octupole_baseline():
    lo = OW._quadmon(np.array([1.0 - 1e-12]))
    hi = OW._quadmon(np.array([1.0]))
    assert abs(lo[0] - hi[0]) < 1e-6
    assert abs(hi[0] - np.exp(0.1940)) < 1e-15


def test_nrtidal_amplitude_derivatives_vs_finite_differences():
    """Tidal amplitude and taper tangents (LambdaTilde, deltaLambda, eta)
    against central differences of the amplitude; the taper's x-tangent is
    dropped as in the reference's custom JVP, so the check sits below the
    taper (pure tidal amplitude) and inside it for the Lambda rows only."""
    from gwfast_amd import catalog
    ev = catalog.draw_bns(2, seed=6)
    wf = OW.IMRPhenomD_NRTidalv2()
    Lt, dL = OW.lamt_dellam_from_lam12(ev['Lambda1'], ev['Lambda2'], ev['eta'])
    fm = wf.fcut(ev) / 1.2
    f = np.array([0.3, 0.7, 1.05, 1.12])[:, None] * fm[None, :]

    def amp(lt, dl):
        e = dict(ev)
        e['Lambda1'], e['Lambda2'] = OW.lam12_from_lamt_dellam(lt, dl, ev['eta'])
        return wf.Ampl(f, e)
    e = dict(ev)
    L1, L2 = OW.lam12_from_lamt_dellam(D.Dual.seed(Lt, 0, 2), D.Dual.seed(dL, 1, 2), ev['eta'])
    e['Lambda1'], e['Lambda2'] = L1, L2
    A = wf.Ampl(f, e)
    for k, (lt, dl) in enumerate([(Lt * 1e-6, 0 * dL), (0 * Lt, np.maximum(np.abs(dL), 1.) * 1e-6)]):
        fd = (amp(Lt + lt, dL + dl) - amp(Lt - lt, dL - dl)) / (2 * (lt + dl))
        assert np.allclose(A.d[k], fd, rtol=1e-5, atol=1e-9 * np.abs(A.v).max())


# ------------------------------------------------- variable changes (m1, m2), (chiS, chiA)
def _m12_from_mceta(Mc, eta):
    M = Mc / eta ** 0.6
    d = np.sqrt(np.maximum(1. - 4. * eta, 0.))
    return 0.5 * M * (1. + d), 0.5 * M * (1. - d)


def _mceta_from_m12(m1, m2):
    M = m1 + m2
    return (m1 * m2) ** 0.6 / M ** 0.2, m1 * m2 / M ** 2


def test_m1m2_chiSA_seeding_vs_central_fd():
    """Duals seeded on (m1, m2, chiS, chiA) give the derivatives of Phi and
    Ampl that central differences in those variables give (chirp mass and
    symmetric mass ratio from textbook definitions, chi1z = chiS + chiA,
    chi2z = chiS - chiA)."""
    wf = OW.IMRPhenomD(fRef=5.)
    ev = _pd_ev(3, 11)
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    f = np.geomspace(0.002, 0.15, 60)[:, None] / M[None, :]
    evd = dict(ev)
    OD._seed_masses_spins(evd, ev, 4, True, False)
    ph, am = wf.Phi(f, evd), wf.Ampl(f, evd)
    m1, m2 = _m12_from_mceta(ev['Mc'], ev['eta'])
    x0 = [m1, m2, 0.5 * (ev['chi1z'] + ev['chi2z']), 0.5 * (ev['chi1z'] - ev['chi2z'])]

    def at(x):
        e = dict(ev)
        e['Mc'], e['eta'] = _mceta_from_m12(x[0], x[1])
        e['chi1z'], e['chi2z'] = x[2] + x[3], x[2] - x[3]
        return wf.Phi(f, e), wf.Ampl(f, e)

    for k in range(4):
        h = 1e-7 * np.maximum(np.abs(x0[k]), 1e-2)
        xp, xm = list(x0), list(x0)
        xp[k] = x0[k] + h
        xm[k] = x0[k] - h
        (pp, ap), (pm, amm) = at(xp), at(xm)
        dph, dam = (pp - pm) / (2 * h), (ap - amm) / (2 * h)
        assert np.max(np.abs(ph.d[k] - dph)) <= 2e-5 * np.max(np.abs(dph)), k
        assert np.max(np.abs(am.d[k] - dam)) <= 2e-5 * np.max(np.abs(dam)), k


def test_fisher_in_m1m2_chiSA_is_pullback():
    """The Fisher matrix in (m1, m2, chiS, chiA) is J^T F J of the one in
    (Mc, eta, chi1z, chi2z), J the Jacobian of the variable change, here by
    central differences of the textbook mass relations."""
    wf = OW.IMRPhenomD()
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(3, seed=12)
    d = glob.detectors['CE1Id']
    det = OD.Detector(glob.PSD_CE, 'L', d['lat'], d['long'], d['xax'], useEarthMotion=True)
    F = det.fisher(wf, ev, 300)
    G = det.fisher(wf, ev, 300, use_m1m2=True, use_chi1chi2=False)
    m1, m2 = _m12_from_mceta(ev['Mc'], ev['eta'])
    n = len(m1)
    J = np.zeros((11, 11, n))
    for i in range(11):
        J[i, i] = 1.
    h1, h2 = 1e-6 * m1, 1e-6 * m2
    dm1 = (np.array(_mceta_from_m12(m1 + h1, m2)) - np.array(_mceta_from_m12(m1 - h1, m2))) / (2 * h1)
    dm2 = (np.array(_mceta_from_m12(m1, m2 + h2)) - np.array(_mceta_from_m12(m1, m2 - h2))) / (2 * h2)
    J[0, 0], J[1, 0] = dm1
    J[0, 1], J[1, 1] = dm2
    J[9, 9], J[9, 10], J[10, 9], J[10, 10] = 1., 1., 1., -1.
    want = np.einsum('ain,abn,bjn->ijn', J, F, J)
    scale = np.sqrt(np.abs(np.einsum('iin->in', want)))
    rel = np.abs(G - want) / (scale[:, None, :] * scale[None, :, :])
    assert np.max(rel) < 1e-6


def test_fishertools_mass_spin_jacobians():
    """fisherTools.py:461-589, 712-774 restated: d(m1, m2)/d(Mc, eta) by
    central differences of the component masses, d(Mc, eta)/d(m1, m2) of the
    textbook definitions, the two mutually inverse, and d(chi1, chi2)/
    d(chi_eff, Delta chi) the inverse of the derivative of
    chi_eff = (m1 chi1 + m2 chi2)/M, Delta chi = chi1 - chi2."""
    from oracle import covariance as OC
    ev = _pd_ev(6, 13)
    Mc, eta = ev['Mc'], np.minimum(ev['eta'], 0.2499)
    K = OC.jac_m1m2_mceta(Mc, eta)
    hM, he = 1e-6 * Mc, 1e-7 * eta
    d_Mc = (np.array(OC.m1m2_from_mceta_unguarded(Mc + hM, eta)) - OC.m1m2_from_mceta_unguarded(Mc - hM, eta)) / (2 * hM)
    d_eta = (np.array(OC.m1m2_from_mceta_unguarded(Mc, eta + he)) - OC.m1m2_from_mceta_unguarded(Mc, eta - he)) / (2 * he)
    assert np.allclose(K[:, 0], d_Mc, rtol=1e-7, atol=0)
    assert np.allclose(K[:, 1], d_eta, rtol=1e-5, atol=0)
    m1, m2 = OC.m1m2_from_mceta_unguarded(Mc, eta)
    J = OC.jac_mceta_m1m2(m1, m2)
    h1, h2 = 1e-6 * m1, 1e-6 * m2
    assert np.allclose(J[:, 0], (np.array(_mceta_from_m12(m1 + h1, m2)) - _mceta_from_m12(m1 - h1, m2)) / (2 * h1),
                       rtol=1e-7, atol=1e-14)
    assert np.allclose(J[:, 1], (np.array(_mceta_from_m12(m1, m2 + h2)) - _mceta_from_m12(m1, m2 - h2)) / (2 * h2),
                       rtol=1e-7, atol=1e-14)
    assert np.allclose(np.einsum('abn,bcn->acn', K, J), np.eye(2)[:, :, None], atol=1e-12)
    S = OC.jac_chi1chi2_chieff_dchi(m1, m2)
    M = m1 + m2
    inv = np.array([[m1 / M, m2 / M], [np.ones_like(M), -np.ones_like(M)]])
    assert np.allclose(np.einsum('abn,bcn->acn', S, inv), np.eye(2)[:, :, None], atol=1e-14)


def test_fishertools_mceta_to_m1m2_fish_is_fisher_in_m1m2():
    """R^T F R with R = d(Mc, eta)/d(m1, m2) on the mass block
    (fisherTools.py:666-687) is the Fisher matrix differentiated in (m1, m2)
    directly; and the dL row of a Fisher in log dL, whose [dL, dL] entry is
    SNR^2, is taken back to SNR^2/dL^2 by the division of
    fisherTools.py:435-458."""
    from oracle import covariance as OC
    wf = OW.IMRPhenomD()
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(3, seed=14)
    d = glob.detectors['CE1Id']
    det = OD.Detector(glob.PSD_CE, 'L', d['lat'], d['long'], d['xax'], useEarthMotion=True)
    F = det.fisher(wf, ev, 300, mode='dual')
    G = det.fisher(wf, ev, 300, mode='dual', use_m1m2=True)
    J = OC.jac_mceta_m1m2(*OC.m1m2_from_mceta_unguarded(ev['Mc'], ev['eta']))
    H = OC.sandwich(F, 0, 1, np.swapaxes(J, 0, 1), J)
    sc = np.sqrt(np.abs(np.einsum('iin->in', G)))
    assert np.max(np.abs(H - G) / (sc[:, None] * sc[None, :])) < 1e-9
    snr = det.snr(wf, ev, 300)
    Flog = np.zeros_like(F)
    Flog[2, 2] = snr ** 2
    back = OC.scale_row_col(Flog, 2, 1. / ev['dL'])
    assert np.allclose(back[2, 2], F[2, 2], rtol=1e-12, atol=0)


# --------------------------------------------------- CovMatr inversion methods
def _two_block(r, P=5, d=None):
    """Unit-diagonal A with a [[1, r], [r, 1]] block (eigenvalues 1 +- r,
    eigenvectors (1, +-1)/sqrt 2) and identity elsewhere; F = D A D."""
    A = np.eye(P)
    A[0, 1] = A[1, 0] = r
    d = np.linspace(0.5, 3.0, P) if d is None else d
    return A, d[:, None] * A * d[None, :], d


def test_cov_methods_agree_on_well_conditioned_matrices():
    from oracle import covariance as OC
    rng = np.random.default_rng(3)
    P, n = 6, 4
    X = rng.standard_normal((n, P, P))
    A = X @ np.transpose(X, (0, 2, 1)) + 3 * P * np.eye(P)   # diagonally dominant: no LU row swaps
    F = np.transpose(A, (1, 2, 0)) * np.geomspace(1, 1e4, P)[:, None, None] * np.geomspace(1, 1e4, P)[None, :, None]
    want = np.transpose(np.linalg.inv(np.transpose(F, (2, 0, 1))), (1, 2, 0))
    for m in ('cho', 'inv', 'svd', 'svd_reg', 'lu'):
        C, eps, meth = OC.cov_matr(F, inv_method=m)
        d = np.sqrt(np.einsum('iin->in', want))
        assert np.max(np.abs(C - want) / (d[:, None] * d[None, :])) < 1e-12, m
        assert np.all(meth == OC.METHOD_IDS[m]) and np.all(eps < 1e-10), m


def test_cov_svd_reg_and_truncation_closed_forms():
    """svd_reg drops singular values <= thresh of the scaled matrix; 'svd'
    with truncate floors them at thresh * max and takes eps against the
    truncated scaled matrix (fisherTools.py:120-149, 189)."""
    from oracle import covariance as OC
    r, thr = 1 - 1e-8, 1e-6
    A, F, d = _two_block(r)
    lam1, lam2 = 1 + r, 1 - r
    q1, q2 = np.array([1, 1, 0, 0, 0]) / np.sqrt(2), np.array([1, -1, 0, 0, 0]) / np.sqrt(2)
    rest = np.diag([0, 0, 1., 1., 1.])
    Di = np.diag(1 / d)
    cc = np.outer(q1, q1) / lam1 + rest
    C, eps, meth = OC.cov_matr(F[:, :, None], inv_method='svd_reg', svals_thresh=thr)
    assert np.allclose(C[:, :, 0], Di @ cc @ Di, rtol=1e-7, atol=1e-12) and meth[0] == 3
    C, eps, meth = OC.cov_matr(F[:, :, None], inv_method='svd', truncate=True, svals_thresh=thr, cond_max=1e3)
    floor = lam1 * thr
    cc = np.outer(q1, q1) / lam1 + np.outer(q2, q2) / floor + rest
    tF = np.outer(q1, q1) * lam1 + np.outer(q2, q2) * floor + rest
    Cw = Di @ cc @ Di
    assert np.allclose(C[:, :, 0], Cw, rtol=1e-6, atol=0)
    assert abs(eps[0] / np.max(np.abs(Cw @ tF - np.eye(5))) - 1) < 1e-6
    # below the condition-number bound nothing is truncated: the plain inverse
    C, eps, meth = OC.cov_matr(F[:, :, None], inv_method='svd', truncate=True, svals_thresh=thr, cond_max=1e50)
    assert np.allclose(C[:, :, 0], np.linalg.inv(F), rtol=1e-6, atol=0)


def test_cov_lu_composition_as_written():
    """'lu' composes U^-1 (P L)^-1 with mpmath's P A = L U
    (fisherTools.py:156-161): the inverse times P^-2, the inverse itself when
    the row swaps form an involution; 'inv' is the inverse."""
    import mpmath
    from oracle import covariance as OC
    A = np.array([[0.1, 1, 0.2], [0.3, 0.1, 2.0], [3.0, 0.2, 0.1]])
    M = mpmath.matrix(A.tolist())
    Pm, _, _ = mpmath.lu(M)
    Pn = np.array(Pm.tolist(), dtype=float)
    assert not np.allclose(Pn @ Pn, np.eye(3))          # a 3-cycle
    cc, _ = OC.invert_scaled(M, 'lu')
    cc = np.array(cc.tolist(), dtype=float)
    assert np.allclose(cc @ Pn @ Pn, np.linalg.inv(A), rtol=1e-12, atol=1e-14)
    ci, _ = OC.invert_scaled(M, 'inv')
    assert np.allclose(np.array(ci.tolist(), dtype=float), np.linalg.inv(A), rtol=1e-12, atol=1e-14)


def test_linspace_jax_grid():
    """Linear Fisher grids (signal.py:893-894): end points exact, uniform
    steps, equal to numpy's linspace to rounding."""
    a, b = np.array([2.0, 3.5]), np.array([120.0, 2000.0])
    g = OD.linspace_jax(a, b, 101)
    assert np.array_equal(g[0], a) and np.array_equal(g[-1], b)
    assert np.allclose(np.diff(g, axis=0), (b - a) / 100, rtol=1e-12, atol=0)
    assert np.allclose(g, np.linspace(a, b, 101), rtol=1e-15, atol=0)


def test_overlap_closed_forms():
    """WFOverlap (signal.py:1759-1935): 1 for identical signals, cos(d) for a
    coalescence-phase shift d (h is proportional to exp(-i Phicoal)),
    independent of the distances, and SNR1 equal to the SNR integral."""
    wf = OW.IMRPhenomD()
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(3, seed=15)
    d = glob.detectors['ETS']
    det = OD.Detector(glob.PSD_ET, 'T', d['lat'], d['long'], d['xax'], useEarthMotion=True)
    ov, s1, s2 = det.wf_overlap(wf, wf, ev, ev, 400)
    assert np.allclose(ov / (s1 * s2), 1.0, rtol=1e-12, atol=0)
    assert np.allclose(s1, det.snr(wf, ev, 400), rtol=1e-12, atol=0)
    ev2 = dict(ev)
    ev2['Phicoal'] = ev['Phicoal'] + 0.7
    ev2['dL'] = ev['dL'] * 3.0
    ov, s1, s2 = det.wf_overlap(wf, wf, ev, ev2, 400)
    assert np.allclose(ov / (s1 * s2), np.cos(0.7), rtol=1e-9, atol=0)


# ---------------------------------------------------------- eccentric TaylorF2
def _ecc_ev(n=4, seed=16):
    from gwfast_amd import catalog
    ev = catalog.draw_bbh(n, seed=seed)
    ev['ecc'] = np.linspace(0.02, 0.2, n)
    return ev


def test_eccentric_tf2_limits():
    """waveforms.py:814-845 restated: e0 = 0 gives the circular phase; at
    low frequency the correction tends to the leading low-eccentricity term
    -(2355/1462) e0^2 (v0/v)^(19/3) 3/(128 eta v^5) (Krolak et al. 1995;
    Yunes & Arun 2009); it is even in e0, so d Phi/d e0 = 2 (Phi - Phi_circ)/e0."""
    wf, wc = OW.TaylorF2(is_eccentric=True), OW.TaylorF2()
    ev = _ecc_ev()
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    f = np.geomspace(1e-7, 1e-3, 80)[:, None] / M[None, :]
    e0 = dict(ev)
    e0['ecc'] = np.zeros_like(ev['ecc'])
    assert np.array_equal(wf.Phi(f, e0), wc.Phi(f, ev))
    dphi = wf.Phi(f, ev) - wc.Phi(f, ev)
    v = (np.pi * M * f) ** (1. / 3.)
    v0 = v[0]
    lead = 3. / (128. * ev['eta']) * (-2355. / 1462.) * ev['ecc'] ** 2 * (v0 / v) ** (19. / 3.) / v ** 5
    rel = np.abs(dphi / lead - 1.)
    assert np.all(rel[:5] < 10. * v[:5] ** 2)   # v < 8e-3: only the O(v^2) 1PN terms remain
    evd = dict(ev)
    evd['ecc'] = D.Dual.seed(ev['ecc'], 0, 1)
    g = wf.Phi(f, evd)
    # dphi is a difference of two phases: allow its cancellation error
    tol = 1e-10 * np.abs(2. * dphi / ev['ecc']) + 1e-14 * np.abs(wc.Phi(f, ev)) * 2. / ev['ecc']
    assert np.all(np.abs(g.d[0] - 2. * dphi / ev['ecc']) <= tol)


def test_eccentric_tf2_dual_vs_complex_step():
    """The eccentric phase is holomorphic in (Mc, eta, e0) as the reference
    declares TaylorF2 (waveforms.py:741): dual tangents equal complex-step
    derivatives; v0 = amin(v) follows Mc and eta along its arg-min."""
    wf = OW.TaylorF2(is_eccentric=True)
    ev = _ecc_ev(3, 17)
    M = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    f = np.geomspace(0.001, 0.02, 50)[:, None] / M[None, :]
    keys = ['Mc', 'eta', 'ecc']
    evd = dict(ev)
    for k, key in enumerate(keys):
        evd[key] = D.Dual.seed(ev[key], k, 3)
    ph = wf.Phi(f, evd)
    for k, key in enumerate(keys):
        h = 1e-20 * np.abs(ev[key])
        evc = {q: np.asarray(v, dtype=complex) for q, v in ev.items()}
        evc[key] = ev[key] + 1j * h
        dph = np.imag(wf.Phi(f.astype(complex), evc)) / h
        assert np.max(np.abs(ph.d[k] - dph)) <= 1e-10 * np.max(np.abs(dph)), key


def test_tidal_eccentric_tf2_is_additive():
    """waveforms.py:814-862: the eccentric and the tidal phase terms enter
    the TaylorF2 phase as independent additions, so with both on the phase is
    Phi_tidal + (Phi_ecc - Phi_circ); the 14-parameter Fisher rows follow
    ParNums (LambdaTilde 11, deltaLambda 12, ecc 13)."""
    from gwfast_amd import catalog
    ev = catalog.draw_bns(4, seed=21)
    ev['ecc'] = np.array([0.01, 0.05, 0.1, 0.2])
    f = np.geomspace(2.0, 0.9 * OW.TaylorF2(is_tidal=True).fcut(ev), 300)
    both = OW.TaylorF2(is_tidal=True, is_eccentric=True).Phi(f, ev)
    tid = OW.TaylorF2(is_tidal=True).Phi(f, ev)
    ecc = OW.TaylorF2(is_eccentric=True).Phi(f, ev)
    circ = OW.TaylorF2().Phi(f, ev)
    assert np.max(np.abs(both - (tid + ecc - circ))) <= 1e-12 * np.max(np.abs(both))
    assert OW.TaylorF2(is_tidal=True, is_eccentric=True).nParams == 14


def test_tidal_eccentric_faithful_equals_dual():
    """The 14-parameter TaylorF2 Fisher built the reference's way (AD rows for
    Mc, eta, chi1z, chi2z, LambdaTilde, deltaLambda, ecc spliced with the
    closed-form rows, signal.py:1143-1299) equals the all-dual construction:
    the ecc row lands at ParNums['ecc'] = 13."""
    from tests._common import NET_C1, oracle_net, norm_err
    from gwfast_amd import catalog
    ev = catalog.draw_bns(3, seed=5)
    ev['ecc'] = np.array([0.02, 0.08, 0.15])
    owf = OW.TaylorF2(is_tidal=True, is_eccentric=True)
    d = oracle_net(NET_C1[1:2], True)
    Fd = OD.network_fisher(d, owf, ev, 400, mode='dual')
    Ff = OD.network_fisher(d, owf, ev, 400, mode='faithful')
    assert Fd.shape == (14, 14, 3)
    assert norm_err(Ff, Fd) <= 1e-8


# ------------------------------------------- tidal TaylorF2, t0 and phiRef
def _vfh_tidal_phase(v, eta, m1, m2, L1, L2):
    """Tidal phase of a binary summed body by body, in the per-body form of
    Vines, Flanagan & Hinderer (2011, PRD 83, 084051, eq. 3.5): with
    X_A = m_A / M, 5PN -24 (12 - 11 X_A) X_A^4 Lambda_A and 6PN
    -(5/28) (3179 - 919 X_A - 2286 X_A^2 + 260 X_A^3) X_A^4 Lambda_A, times
    3 / (128 eta v^5).  Wade et al. (2014, PRD 89, 103012; cited at
    waveforms.py:851) rewrite exactly this sum with LambdaTilde and
    deltaLambda."""
    M = m1 + m2
    s = 0.
    for L, X in ((L1, m1 / M), (L2, m2 / M)):
        s = s + L * X ** 4 * (-24. * (12. - 11. * X) * v ** 10
                              - (5. / 28.) * (3179. - 919. * X - 2286. * X * X + 260. * X ** 3) * v ** 12)
    return 3. / (128. * eta * v ** 5) * s


def test_tf2_tidal_phase_is_per_body_vfh_sum():
    """TaylorF2's tidal phase (waveforms.py:850-858) minus the point-particle
    phase equals the per-body 5PN + 6PN tidal phase written in Lambda1,
    Lambda2 and the component masses.  This pins the -39/2 and -3115/64
    coefficients, the 6595/364 sqrt(1 - 4 eta) deltaLambda term and the
    LambdaTilde / deltaLambda combinations of gwfastUtils.py:398-417 at
    once; the inverse map (gwfastUtils.py:419-444) must return the inputs."""
    from gwfast_amd import catalog
    ev = catalog.draw_bns(8, seed=21)
    ev['eta'][0] = 0.25 - 1e-9           # near-equal masses
    ev['eta'][1] = 0.20                  # mass ratio ~ 0.38
    M = ev['Mc'] / ev['eta'] ** 0.6
    Seta = np.sqrt(1. - 4. * ev['eta'])
    m1, m2 = 0.5 * M * (1. + Seta), 0.5 * M * (1. - Seta)
    fc = OW.TaylorF2(is_tidal=True).fcut(ev)
    f = np.geomspace(20., 1., 60)[:, None] * fc[None, :] / 21.   # v from ~0.1 to ~0.4
    d = OW.TaylorF2(is_tidal=True).Phi(f, ev) - OW.TaylorF2().Phi(f, ev)
    v = (np.pi * M * glob.GMsun_over_c3 * f) ** (1. / 3.)
    want = _vfh_tidal_phase(v, ev['eta'], m1, m2, ev['Lambda1'], ev['Lambda2'])
    scale = np.max(np.abs(want), axis=0)
    assert np.all(scale > 1.)                # the tidal phase is radians-sized here
    # the difference of two ~1e2-1e4 rad phases carries ~1e-12 rad of rounding
    assert np.max(np.abs(d - want) / scale) < 1e-10
    Lt, dLm = OW.lamt_dellam_from_lam12(ev['Lambda1'], ev['Lambda2'], ev['eta'])
    L1, L2 = OW.lam12_from_lamt_dellam(Lt, dLm, ev['eta'])
    assert np.allclose(L1, ev['Lambda1'], rtol=1e-9, atol=1e-9)
    assert np.allclose(L2, ev['Lambda2'], rtol=1e-9, atol=1e-9)
    # equal masses: LambdaTilde = Lambda, deltaLambda = 0 (Wade et al. 2014)
    Lt, dLm = OW.lamt_dellam_from_lam12(np.array([400.]), np.array([400.]), np.array([0.25]))
    assert abs(Lt[0] - 400.) < 1e-12 and abs(dLm[0]) < 1e-12


def _pd_branch_events():
    """Events on both sides of IMRPhenomD's gamma2 = 1 switch (gamma2 >= 1
    only at small eta and spins near +1)."""
    n = 6
    ev = {'Mc': np.array([12., 8., 30., 25., 40., 20.]), 'eta': np.array([0.05, 0.055, 0.2, 0.2499, 0.12, 0.05]),
          'dL': np.ones(n), 'chi1z': np.array([0.99, 0.99, 0.3, -0.6, 0.8, -0.9]),
          'chi2z': np.array([0.99, 0.99, -0.2, -0.6, 0.1, -0.9])}
    return ev


def _ansatz_mrd(x, co):
    """The merger-ringdown phase ansatz of Khan et al. (2016, PRD 93 044007,
    eq. 14): (alpha1 x - alpha2 / x + (4/3) alpha3 x^(3/4)
    + alpha4 arctan((x - alpha5 fring) / fdamp)) / eta."""
    a1, a2, a3, a4, a5 = (D.value(co[k]) for k in ('alpha1', 'alpha2', 'alpha3', 'alpha4', 'alpha5'))
    fr, fd = D.value(co['fring']), D.value(co['fdamp'])
    return (a1 * x - a2 / x + (4. / 3.) * a3 * x ** 0.75 + a4 * np.arctan((x - a5 * fr) / fd)) / D.value(co['eta'])


def test_phenomd_t0_is_mrd_time_at_fpeak():
    """IMRPhenomD's time shift t0 (waveforms.py:1132-1136) is the derivative
    of the merger-ringdown phase ansatz at the amplitude peak fpeak, on both
    gamma2 branches; fpeak maximises the Lorentzian ringdown amplitude
    exp(-gamma2 (x - fring) / (gamma3 fdamp)) / ((x - fring)^2
    + (gamma3 fdamp)^2) where gamma2 < 1.  End to end, the final phase's slope at fpeak is C2MRD."""
    from scipy.optimize import minimize_scalar
    wf = OW.IMRPhenomD()
    ev = _pd_branch_events()
    MG = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    f = np.geomspace(2., 1000., 16)[:, None] / (MG[None, :] / MG.min())
    co = wf.phase_coeffs(f, ev)
    g2 = D.value(co['gamma2'])
    assert np.any(g2 >= 1.) and np.any(g2 < 1.)
    fpk, t0 = D.value(co['fpeak']), D.value(co['t0'])
    h = 1e-5 * fpk
    fd_t0 = (_ansatz_mrd(fpk + h, co) - _ansatz_mrd(fpk - h, co)) / (2. * h)
    assert np.max(np.abs(fd_t0 / t0 - 1.)) < 1e-8
    fr, fdm, g3 = D.value(co['fring']), D.value(co['fdamp']), D.value(co['gamma3'])
    for e in np.nonzero(g2 < 1.)[0]:
        a = fdm[e] * g3[e]
        r = minimize_scalar(lambda x: g2[e] * (x - fr[e]) / a + np.log((x - fr[e]) ** 2 + a * a),
                            bounds=(fr[e] - a, fr[e] + 5 * a), method='bounded', options={'xatol': 1e-14})
        assert abs(r.x / fpk[e] - 1.) < 1e-7
    # (the ansatz also has a local minimum below fring - gamma3 fdamp, where
    # it rises again towards low frequency; the peak is the local maximum.
    # For gamma2 >= 1 the ansatz has no maximum and fpeak = |fring - gamma3
    # fdamp / gamma2| is the model's convention, the same point as the
    # amplitude's collocation peak.)
    # the total phase's slope at fpeak (Mf units) is C2MRD
    evf = dict(ev)
    fp_hz = fpk / MG
    ff = np.array([fp_hz * (1. - 1e-5), fp_hz * (1. + 1e-5)])
    c2 = D.value(co['C2MRD'])
    assert np.all(fpk < wf.fcutPar)
    for cut in (True, False):
        p = OW.IMRPhenomD(apply_fcut=cut).Phi(ff, evf)
        slope = (p[1] - p[0]) / (MG * (ff[1] - ff[0]))
        assert np.max(np.abs(slope - c2) / np.abs(t0)) < 1e-6


def test_phenomd_phase_vanishes_at_fref():
    """The reference phase phiRef (waveforms.py:1139-1147) is the phase at
    fRef: the returned phase is exactly zero there, with fRef the grid
    minimum by default or a given frequency; t0 (f - fRef) leaves the
    phase's slope unchanged by the choice of fRef."""
    ev = _pd_ev(4, seed=5)
    MG = ev['Mc'] / ev['eta'] ** 0.6 * glob.GMsun_over_c3
    fc = OW.IMRPhenomD().fcut(ev)
    f = np.geomspace(2., 1., 300)[:, None] * fc[None, :] * (np.array([1.0, 0.9, 0.5, 0.2])[None, :] + 0.)
    f = np.sort(f, axis=0)
    p = OW.IMRPhenomD().Phi(f, ev)
    assert np.all(np.abs(p[0]) <= 1e-12 * np.max(np.abs(p), axis=0))
    assert np.max(np.abs(p[-5])) > 1.
    wf10 = OW.IMRPhenomD(fRef=10.)
    f10 = np.concatenate([np.full((1, 4), 10.), f], axis=0)
    p10 = wf10.Phi(f10, ev)
    assert np.all(np.abs(p10[0]) <= 1e-12 * np.max(np.abs(p10), axis=0))
    # changing fRef adds a constant: the slopes agree
    p_default = OW.IMRPhenomD().Phi(f10, ev)
    dd = (p10 - p_default) - (p10 - p_default)[0]
    assert np.max(np.abs(dd)) <= 1e-9 * np.max(np.abs(p_default))
    del MG
