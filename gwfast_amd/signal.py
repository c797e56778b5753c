"""Single-detector signal, SNR and Fisher matrix with gwfast's GWSignal API.

Mirrors gwfast/signal.py:37-160 (constructor), 485-655 (GWstrain),
658-777 (SNRInteg) and 782-1099 (FisherMatr).  Every computation runs in the
HIP library: SNR and Fisher in one fused pass per event (waveform, detector
projection with Earth rotation, strain tangents, trapezoid integral), the
strain and its derivative rows only when they are asked for.
"""
import numpy as np

from . import _engine as E
from . import _native as N
from .gwfastUtils import check_evparams, Lam12_from_Lamt_delLam, Mceta_from_m1m2


class GWSignal(object):
    """gwfast/signal.py:37-160."""

    def __init__(self, wf_model, psd_path=None, detector_shape='T', det_lat=40.44, det_long=9.45, det_xax=0.,
                 verbose=True, is_ASD=True, useEarthMotion=False, noMotion=False, fmin=2., fmax=None,
                 IntTablePath=None, DutyFactor=None, compute2arms=True, jitCompileDerivs=False):
        if (detector_shape != 'L') and (detector_shape != 'T'):
            raise ValueError('Enter valid detector configuration')
        if psd_path is None:
            raise ValueError('Enter a valid PSD or ASD path')
        if DutyFactor is not None and not (0. <= DutyFactor <= 1.):
            raise ValueError('DutyFactor must be between 0 and 1')
        self.wf_model = wf_model
        self.verbose = verbose
        self.detector_shape = detector_shape
        self.det_lat, self.det_long, self.det_xax = float(det_lat), float(det_long), float(det_xax)
        self.det_lat_rad = det_lat * np.pi / 180.
        self.det_long_rad = det_long * np.pi / 180.
        self.det_xax_rad = det_xax * np.pi / 180.
        self.IntTablePath = IntTablePath
        self.DutyFactor = DutyFactor
        noise = np.loadtxt(psd_path, usecols=(0, 1))
        self.strainFreq = np.ascontiguousarray(noise[:, 0])
        self.noiseCurve = np.ascontiguousarray(noise[:, 1] ** 2 if is_ASD else noise[:, 1])
        if np.any(np.diff(self.strainFreq) <= 0):
            raise ValueError('PSD frequencies must be strictly increasing')
        self.useEarthMotion = useEarthMotion
        self.noMotion = noMotion
        if self.noMotion and self.useEarthMotion:
            self.useEarthMotion = False
        self.fmin = fmin
        self.fmax = fmax
        self.angbtwArms = 0.5 * np.pi if detector_shape == 'L' else np.pi / 3.
        self.compute2arms = compute2arms
        self.jitCompileDerivs = jitCompileDerivs
        # duty-cycle seed (signal.py:152-153): numpy's global generator, as the reference
        np.random.seed(None)
        self.seedUse = np.random.randint(2 ** 32 - 1, size=1)
        self._psd_dev = None

    def _clear_cache(self):
        """signal.py:216-219 rebuilds compiled derivative functions; here the device
        copy of the PSD table is what is cached."""
        self._psd_dev = None

    def _update_seed(self, seed=None):
        """signal.py:221-232."""
        np.random.seed(None)
        self.seedUse = np.random.randint(2 ** 32 - 1, size=1) if seed is None else seed

    def _duty_masks(self, n):
        """The reference's duty-cycle masks (signal.py:672-673, 729-761,
        932-936): numpy's global generator reseeded with seedUse, then one
        Bernoulli(DutyFactor) draw of n per arm, in arm order.  None without a
        DutyFactor."""
        if self.DutyFactor is None:
            return None
        np.random.seed(self.seedUse)
        return np.array([np.random.choice([0, 1], n, p=[1. - self.DutyFactor, self.DutyFactor])
                         for _ in range(self.n_arms)], dtype=np.float64)

    # ------------------------------------------------------------------ ABI
    def c_struct(self, lin=False):
        if self._psd_dev is None:
            self._psd_dev = _shared_psd(self.strainFreq, self.noiseCurve)
        d = N.GwfDetector()
        d.shape = N.SHAPE_L if self.detector_shape == 'L' else N.SHAPE_T
        d.earth_motion = int(self.useEarthMotion)
        d.no_motion = int(self.noMotion)
        d.compute2arms = int(self.compute2arms)
        d.lat_deg, d.long_deg, d.xax_deg = self.det_lat, self.det_long, self.det_xax
        d.fmin = float(self.fmin)
        d.fmax = float(self.fmax) if self.fmax is not None else 0.0
        d.psd_f = self._psd_dev[0].data_ptr()
        d.psd_S = self._psd_dev[1].data_ptr()
        d.n_psd = len(self.strainFreq)
        d.psd_f_first = float(self.strainFreq[0])
        d.psd_f_last = float(self.strainFreq[-1])
        d.lin_spacing = int(bool(lin))
        return d

    @property
    def n_arms(self):
        return 1 if self.detector_shape == 'L' else 3

    def _complete(self, evParams):
        """Input completion of signal.py:684-713."""
        check_evparams(evParams)
        if 'chi1z' not in evParams:
            try:
                evParams['chi1z'] = evParams['chiS'] + evParams['chiA']
                evParams['chi2z'] = evParams['chiS'] - evParams['chiA']
            except KeyError:
                raise ValueError('Two among chi1z, chi2z and chiS, chiA have to be provided.')
        if self.wf_model.is_tidal and 'Lambda1' not in evParams:
            try:
                evParams['Lambda1'], evParams['Lambda2'] = Lam12_from_Lamt_delLam(
                    evParams['LambdaTilde'], evParams['deltaLambda'], evParams['eta'])
            except KeyError:
                raise ValueError('Two among Lambda1, Lambda2 and LambdaTilde and deltaLambda have to be provided.')
        return evParams

    def _rows(self, evParams):
        return E.to_dev(E.param_rows(evParams, self.wf_model.is_tidal, self.wf_model.is_eccentric))

    # --------------------------------------------------------------- public
    def SNRInteg(self, evParams, res=1000, return_all=False):
        """signal.py:658-777 (with the DutyFactor masks when set)."""
        self._complete(evParams)
        snrsq = snr_sq_device([self], self._rows(evParams), res).cpu().numpy()
        masks = self._duty_masks(snrsq.shape[1])
        if masks is not None:
            snrsq = snrsq * masks
        if return_all:
            return 2 * np.sqrt(snrsq) if self.detector_shape == 'T' else 2 * np.sqrt(snrsq[0])
        return 2 * np.sqrt(snrsq.sum(axis=0))

    def FisherMatr(self, evParams, res=1000, df=None, spacing='geom', use_m1m2=False, use_chi1chi2=True,
                   use_prec_ang=True, computeDerivFinDiff=False, computeAnalyticalDeriv=True, return_all=False,
                   return_derivatives=False, return_SNR_derivatives=False, **kwargs):
        """signal.py:782-1099.  Derivatives are exact (forward-mode) for every
        row, which equals the reference's JAX + closed-form rows up to
        rounding; the finite-difference option is not provided."""
        _check_fisher_options(res, df, spacing, use_m1m2, use_chi1chi2, computeDerivFinDiff)
        self._complete(evParams)
        res = resolve_res(self, evParams, res, df)
        lin = spacing == 'lin'
        rows = self._rows(evParams)
        n = rows.shape[1]
        masks = self._duty_masks(n)
        P = self.wf_model.nParams
        if return_derivatives or return_SNR_derivatives:
            if np.ndim(res) != 0:
                raise NotImplementedError('derivative outputs need one grid length for all events')
            allF, allD = self._fisher_with_derivs(evParams, rows, int(res), return_derivatives, lin)
            for k in range(len(allF)):
                allF[k] = reparam_host(allF[k], evParams, use_m1m2, use_chi1chi2)
                allD[k] = reparam_rows_host(allD[k], evParams, use_m1m2, use_chi1chi2)
                if masks is not None:
                    allF[k] = allF[k] * masks[k]
                    allD[k] = allD[k] * (masks[k][None, :, None] if allD[k].ndim == 3 else masks[k][None, :])
            return allF, allD
        if return_all:
            return self._fisher_per_arm(rows, res, lin, use_m1m2, use_chi1chi2, masks)
        if masks is not None:
            F = fisher_masked_device([self], self.wf_model, rows, res, mask_words([masks]), lin=lin)
        else:
            F = fisher_device([self], self.wf_model, rows, res, lin=lin)
        reparam_device(F, rows, use_m1m2, use_chi1chi2)
        return F.cpu().numpy()

    def _fisher_per_arm(self, rows, res, lin, use_m1m2, use_chi1chi2, masks, Fa=None):
        """FisherMatr(return_all=True): the list of this detector's per-arm
        matrices (host), each times its duty-cycle mask.  Fa: this detector's
        (n_arms, P, P, N) slice of a gwf_fisher_arms launch, else one launch
        here."""
        if Fa is None and self.n_arms <= ARMS_PER_LAUNCH:
            Fa = fisher_arms_device([self], self.wf_model, rows, res, lin=lin)
        Fs = []
        for k in range(self.n_arms):
            F = Fa[k] if Fa is not None else fisher_device([self], self.wf_model, rows, res, arm=k, lin=lin)
            reparam_device(F, rows, use_m1m2, use_chi1chi2)
            Fs.append(F.cpu().numpy() * (masks[k] if masks is not None else 1.0))
        return Fs

    def _fisher_with_derivs(self, evParams, rows, res, return_derivatives, lin=False):
        n = rows.shape[1]
        f, S = grid_device(self, rows, res, lin)
        P = self.wf_model.nParams
        allF, allD = [], []
        for k in range(self.n_arms):
            h, J = strain_device(self, rows, f, k, want_h=not return_derivatives)
            F = gram_device(J, f, S, P)
            allF.append(F.cpu().numpy())
            Jc = J.cpu().numpy()
            Jc = Jc[..., 0] + 1j * Jc[..., 1]
            if return_derivatives:
                allD.append(Jc)
            else:
                hc = h.cpu().numpy()
                hc = hc[..., 0] + 1j * hc[..., 1]
                fg, Sg = f.cpu().numpy(), S.cpu().numpy()
                D = np.zeros((P, n))
                for a in range(P):
                    D[a] = 4. * np.trapezoid((hc * np.conjugate(Jc[a]).T).real / Sg, fg, axis=0)
                allD.append(D)
        return allF, allD

    def GWstrain(self, f, Mc, eta, dL, theta, phi, iota, psi, tcoal, Phicoal, chiS, chiA, chi1x=0., chi2x=0.,
                 chi1y=0., chi2y=0., LambdaTilde=0., deltaLambda=0., ecc=0., rot=0., is_m1m2=False,
                 is_chi1chi2=False, is_prec_ang=False, is_Lam1Lam2=False, return_single_comp=None):
        """signal.py:485-655: complex strain on the grid f (N_f, N_ev) for the
        arm at rotation rot (degrees; 0, 60, 120 are a triangle's arms); is_m1m2 reads (Mc, eta) as
        (m1, m2); return_single_comp in 'Ap', 'Ac', 'Psip', 'Psic', 'At',
        'Psit' returns that real component instead (629-652)."""
        if is_prec_ang:
            raise NotImplementedError('is_prec_ang needs a precessing model, which is not on this path')
        comps = ('Ap', 'Ac', 'Psip', 'Psic', 'At', 'Psit')
        if return_single_comp is not None and return_single_comp not in comps:
            raise ValueError('Single component to return has to be among Ap, Ac, Psip, Psic')
        if is_m1m2:
            Mc, eta = Mceta_from_m1m2(np.asarray(Mc, dtype=np.float64), np.asarray(eta, dtype=np.float64))
        chi1z, chi2z = (chiS, chiA) if is_chi1chi2 else (np.asarray(chiS) + chiA, np.asarray(chiS) - chiA)
        ev = {'Mc': Mc, 'eta': eta, 'dL': dL, 'theta': theta, 'phi': phi, 'iota': iota, 'psi': psi,
              'tcoal': tcoal, 'Phicoal': Phicoal, 'chi1z': chi1z, 'chi2z': chi2z}
        if self.wf_model.is_tidal:
            if is_Lam1Lam2:
                ev['Lambda1'], ev['Lambda2'] = LambdaTilde, deltaLambda
            else:
                ev['LambdaTilde'], ev['deltaLambda'] = LambdaTilde, deltaLambda
        if self.wf_model.is_eccentric:
            ev['ecc'] = ecc
        ev = {k: np.atleast_1d(np.asarray(v, dtype=np.float64)) for k, v in ev.items()}
        det = rot_detector(self, rot)
        rows = self._rows(ev)
        fd = _grid_dev(f, rows.shape[1])
        if return_single_comp is not None:
            t = E.torch()
            c = strain_components_device(self, rows, fd, 0, det=det)
            Ap, Ac, Psi = c[0], c[1], c[2]
            out = {'Ap': lambda: Ap, 'Ac': lambda: Ac, 'Psip': lambda: Psi, 'Psic': lambda: Psi + np.pi * 0.5,
                   'At': lambda: t.hypot(Ap, Ac), 'Psit': lambda: Psi + t.atan2(Ac, Ap)}[return_single_comp]()
            return out.cpu().numpy()
        h, _ = strain_device(self, rows, fd, 0, want_h=True, want_J=False, det=det)
        hc = h.cpu().numpy()
        return hc[..., 0] + 1j * hc[..., 1]

    def GWAmplitudes(self, evParams, f, rot=0.):
        """signal.py:425-466: plus and cross amplitudes (Ap, Ac) of the arm at
        rotation rot on the grid f (N_f, N_ev)."""
        ev = self._complete(dict(evParams))
        rows = self._rows(ev)
        c = strain_components_device(self, rows, _grid_dev(f, rows.shape[1]), 0, det=rot_detector(self, rot))
        return c[0].cpu().numpy(), c[1].cpu().numpy()

    def WFOverlap(self, WF1, WF2, evParams1, evParams2, res=1000, return_separate=False, **kwargs):
        """signal.py:1759-1935: normalised overlap of the strains of WF1 on
        evParams1 and WF2 on evParams2, summed over the arms, on one grid from
        fmin to the larger cut; return_separate gives (overlap integral,
        SNR1, SNR2)."""
        t = E.torch()
        ev1, ev2 = _complete_for(WF1, evParams1), _complete_for(WF2, evParams2)
        check_evparams(ev1)
        check_evparams(ev2)
        fcut1 = np.asarray(WF1.fcut(**ev1))
        fcut2 = np.asarray(WF2.fcut(**ev2))
        fcutUse = np.where(fcut1 > fcut2, fcut1, fcut2)
        if self.fmax is not None:
            # as the reference writes it (signal.py:1858-1859): fcut1 where below fmax
            fcutUse = np.where(fcutUse > self.fmax, self.fmax, fcut1)
        r1 = E.to_dev(E.param_rows(ev1, WF1.is_tidal, WF1.is_eccentric))
        r2 = E.to_dev(E.param_rows(ev2, WF2.is_tidal, WF2.is_eccentric))
        n = r1.shape[1]
        f = t.empty((int(res), n), dtype=t.float64, device=r1.device)
        S = t.empty_like(f)
        E.call('gwf_grid_fcut', ctypes_byref(self.c_struct()), N.ptr(E.to_dev(fcutUse)), n, int(res), N.ptr(f), N.ptr(S),
               N.stream_ptr())
        out = t.empty((3, n), dtype=t.float64, device=r1.device)
        for k in range(self.n_arms):
            h1, _ = strain_device(self, r1, f, k, want_J=False, wf=WF1)
            h2, _ = strain_device(self, r2, f, k, want_J=False, wf=WF2)
            E.call('gwf_overlap', N.ptr(h1), N.ptr(h2), N.ptr(f), N.ptr(S), n, int(res), int(k > 0), N.ptr(out),
                   N.stream_ptr())
        o = out.cpu().numpy()
        overlap_int, SNRh1, SNRh2 = o[0], np.sqrt(o[1]), np.sqrt(o[2])
        if return_separate:
            return overlap_int, SNRh1, SNRh2
        return overlap_int / (SNRh1 * SNRh2)

    def GWPhase(self, evParams, f):
        """signal.py:468-483: 2 pi f tcoal - Phicoal - Phi_GW."""
        return (2. * np.pi * f * (evParams['tcoal'] * 3600. * 24.) - evParams['Phicoal']
                - self.wf_model.Phi(f, **evParams))


_PSD_DEV = {}


def _shared_psd(f, S):
    """One device copy per distinct PSD table and device: detectors with the
    same noise curve (e.g. two CE sites) then share it, and a network kernel
    interpolates each distinct table once per frequency."""
    key = (E.device().index, f.shape[0], hash(f.tobytes()), hash(S.tobytes()))
    hit = _PSD_DEV.get(key)
    if hit is None or not (np.array_equal(hit[2], f) and np.array_equal(hit[3], S)):
        hit = (E.to_dev(f), E.to_dev(S), f.copy(), S.copy())
        _PSD_DEV[key] = hit
    return hit[0], hit[1]


def reparam_device(F, rows, use_m1m2, use_chi1chi2):
    """In place: Fisher (P, P, N) device tensor into (m1, m2) and / or
    (chiS, chiA) rows (signal.py:818-845), by the exact Jacobian."""
    if not use_m1m2 and use_chi1chi2:
        return F
    E.check_rows(rows, need=11)
    E.call('gwf_fisher_reparam', N.ptr(rows), rows.stride(0), N.ptr(F), F.shape[0], F.shape[2], int(bool(use_m1m2)),
           int(not use_chi1chi2), N.stream_ptr())
    return F


def _jac_blocks(evParams, use_m1m2, use_chi1chi2):
    """Host copies of the 2x2 Jacobian blocks used by gwf_fisher_reparam."""
    Mc = np.atleast_1d(np.asarray(evParams['Mc'], dtype=np.float64))
    eta = np.atleast_1d(np.asarray(evParams['eta'], dtype=np.float64))
    n = Mc.shape[0]
    A = np.zeros((2, 2, n))
    A[0, 0] = A[1, 1] = 1.
    if use_m1m2:
        M = Mc / eta ** (3. / 5.)
        Seta = np.sqrt(np.where(eta < 0.25, 1.0 - 4.0 * eta, 0.))
        m1, m2 = 0.5 * M * (1.0 + Seta), 0.5 * M * (1.0 - Seta)
        A = np.array([[Mc * (3. / (5. * m1) - 1. / (5. * M)), Mc * (3. / (5. * m2) - 1. / (5. * M))],
                      [eta * (1. / m1 - 2. / M), eta * (1. / m2 - 2. / M)]])
    B = np.zeros((2, 2, n))
    B[0, 0] = B[1, 1] = 1.
    if not use_chi1chi2:
        B[0, 0], B[0, 1], B[1, 0], B[1, 1] = 1., 1., 1., -1.
    return A, B


def reparam_host(F, evParams, use_m1m2, use_chi1chi2):
    """The same change of variables on a host Fisher (P, P, N)."""
    if not use_m1m2 and use_chi1chi2:
        return F
    A, B = _jac_blocks(evParams, use_m1m2, use_chi1chi2)
    J = np.zeros((F.shape[0], F.shape[0], F.shape[2]))
    for i in range(F.shape[0]):
        J[i, i] = 1.
    J[0:2, 0:2] = A
    J[9:11, 9:11] = B
    return np.einsum('kin,kln,ljn->ijn', J, F, J)


def reparam_rows_host(D, evParams, use_m1m2, use_chi1chi2):
    """Derivative rows (P, N[, N_f]) in the new variables: D'_j = sum_k J_kj D_k."""
    if not use_m1m2 and use_chi1chi2:
        return D
    A, B = _jac_blocks(evParams, use_m1m2, use_chi1chi2)
    out = np.array(D, copy=True)
    ex = (lambda a: a[:, None]) if D.ndim == 3 else (lambda a: a)
    for (r0, blk) in ((0, A), (9, B)):
        d0, d1 = D[r0], D[r0 + 1]
        out[r0] = ex(blk[0, 0]) * d0 + ex(blk[1, 0]) * d1
        out[r0 + 1] = ex(blk[0, 1]) * d0 + ex(blk[1, 1]) * d1
    return out


def resolve_res(sig, evParams, res, df):
    """signal.py:884-892: res, or from the step df the largest
    floor(1 + (fcut - fmin)/df) over the events (one grid length for all)."""
    if res is not None:
        return res
    if df is None:
        raise ValueError('Provide either resolution in frequency or step size.')
    fcut = np.asarray(sig.wf_model.fcut(**evParams))
    if sig.fmax is not None:
        fcut = np.where(fcut > sig.fmax, sig.fmax, fcut)
    return int(np.amax(np.floor(1 + (fcut - sig.fmin) / df)))


def _check_fisher_options(res, df, spacing, use_m1m2, use_chi1chi2, computeDerivFinDiff):
    if res is None and df is None:
        raise ValueError('Provide either resolution in frequency or step size.')
    if res is not None and np.ndim(res) == 0 and int(res) < 2:
        raise ValueError('res must be at least 2')
    if spacing not in ('geom', 'lin'):
        raise ValueError("spacing must be 'geom' or 'lin'")
    if computeDerivFinDiff:
        raise NotImplementedError('finite-difference derivatives are not provided (derivatives are exact)')


# ------------------------------------------------------------ device helpers
def _dets(signals, lin=False):
    arr = (N.GwfDetector * len(signals))(*[s.c_struct(lin) for s in signals])
    return arr


def _res_arg(res, n, device):
    """int -> (int, None); per-event array -> (0, device int32 tensor)."""
    if np.ndim(res) == 0:
        return int(res), None
    t = E.torch()
    a = np.asarray(res)
    if a.size and (a.min() < 2 or a.max() > 1 << 30):
        raise ValueError('every grid length must be between 2 and 2**30')
    r = t.as_tensor(a, dtype=t.int32, device=device).contiguous()
    if r.numel() != n:
        raise ValueError('res must be an int or have one entry per event')
    return 0, r


def snr_sq_device(signals, rows, res):
    """Per-arm trapz((Ap^2+Ac^2)/S, f) for a list of GWSignal sharing one model,
    fmin and fmax: (n_arms, N_ev) device tensor.  res: int, or one grid length
    per event (ragged)."""
    t = E.torch()
    wf = signals[0].wf_model
    E.check_rows(rows, wf)
    n = rows.shape[1]
    narms = sum(s.n_arms for s in signals)
    out = t.empty((narms, n), dtype=t.float64, device=rows.device)
    ws, nb = E.workspace(n)
    r, r_ev = _res_arg(res, n, rows.device)
    if r_ev is None:
        E.call('gwf_snr', wf.c_struct(), _dets(signals), len(signals), N.ptr(rows), rows.stride(0), n, r, N.ptr(out),
               N.ptr(ws), nb, N.stream_ptr())
    else:
        E.call('gwf_snr_ragged', wf.c_struct(), _dets(signals), len(signals), N.ptr(rows), rows.stride(0), n,
               N.ptr(r_ev), N.ptr(out), N.ptr(ws), nb, N.stream_ptr())
    return out


def fisher_device(signals, wf, rows, res, arm=-1, out=None, lin=False):
    """Fused network Fisher (P, P, N_ev) device tensor; arm >= 0 selects one
    arm; lin: linearly spaced grids (FisherMatr spacing='lin')."""
    t = E.torch()
    E.check_rows(rows, wf)
    n = rows.shape[1]
    P = wf.nParams
    if out is None:
        out = t.empty((P, P, n), dtype=t.float64, device=rows.device)
    ws, nb = E.workspace(n)
    r, r_ev = _res_arg(res, n, rows.device)
    if r_ev is None:
        E.call('gwf_fisher', wf.c_struct(), _dets(signals, lin), len(signals), N.ptr(rows), rows.stride(0), n, r, arm,
               N.ptr(out), N.ptr(ws), nb, N.stream_ptr())
    else:
        E.call('gwf_fisher_ragged', wf.c_struct(), _dets(signals, lin), len(signals), N.ptr(rows), rows.stride(0), n,
               N.ptr(r_ev), arm, N.ptr(out), N.ptr(ws), nb, N.stream_ptr())
    return out


def fisher_arms_device(signals, wf, rows, res, lin=False):
    """The Fisher matrix of every physical arm of signals (in network arm
    order) from one gwf_fisher_arms launch: (n_arms, P, P, N_ev) device
    tensor; res: int, or one grid length per event (ragged)."""
    t = E.torch()
    E.check_rows(rows, wf)
    n = rows.shape[1]
    P = wf.nParams
    n_arms = sum(s.n_arms for s in signals)
    out = t.empty((n_arms, P, P, n), dtype=t.float64, device=rows.device)
    ws, nb = E.workspace(n)
    r, r_ev = _res_arg(res, n, rows.device)
    E.call('gwf_fisher_arms', wf.c_struct(), _dets(signals, lin), len(signals), N.ptr(rows), rows.stride(0), n, r,
           N.ptr(r_ev), N.ptr(out), N.ptr(ws), nb, N.stream_ptr())
    return out


ARMS_PER_LAUNCH = 8  # gwf_fisher_arms' limit


def mask_words(masks_per_signal):
    """Per-event arm bit masks for gwf_fisher_masked from each signal's
    (n_arms, N) 0/1 masks (None: every arm on), in network arm order; a
    (N,) int32 device tensor holding the uint32 bits."""
    words = None
    off = 0
    for m in masks_per_signal:
        m = np.asarray(m)
        if words is None:
            words = np.zeros(m.shape[1], dtype=np.int64)
        for a in range(m.shape[0]):
            words |= (m[a] != 0).astype(np.int64) << (off + a)
        off += m.shape[0]
    return E.to_dev(words.astype(np.int32), E.torch().int32)


def fisher_masked_device(signals, wf, rows, res, words, lin=False):
    """Network Fisher (P, P, N_ev) device tensor summing, per event, the arms
    whose bit is set in words (one fused launch)."""
    t = E.torch()
    E.check_rows(rows, wf)
    n = rows.shape[1]
    if np.ndim(res) != 0:
        raise NotImplementedError('duty-cycle masks need one grid length for all events')
    if words.dtype != t.int32 or words.shape != (n,) or words.device != rows.device:
        raise TypeError('arm masks must be an int32 device tensor of one word per event')
    out = t.empty((wf.nParams, wf.nParams, n), dtype=t.float64, device=rows.device)
    ws, nb = E.workspace(n)
    E.call('gwf_fisher_masked', wf.c_struct(), _dets(signals, lin), len(signals), N.ptr(rows), rows.stride(0), n, int(res),
           N.ptr(words), N.ptr(out), N.ptr(ws), nb, N.stream_ptr())
    return out


def grid_device(sig, rows, res, lin=False):
    t = E.torch()
    E.check_rows(rows, sig.wf_model)
    n = rows.shape[1]
    f = t.empty((res, n), dtype=t.float64, device=rows.device)
    S = t.empty((res, n), dtype=t.float64, device=rows.device)
    ws, nb = E.workspace(n)
    E.call('gwf_grid', sig.wf_model.c_struct(), _dets([sig], lin), N.ptr(rows), rows.stride(0), n, res, N.ptr(f), N.ptr(S),
           N.ptr(ws), nb, N.stream_ptr())
    return f, S


def rot_detector(sig, rot):
    """The detector block of the single arm at rotation rot (degrees) from
    the x-arm, for any rot: the pattern functions depend on the arm only
    through xax + rot (signal.py:360-376), so this is arm 0 of the same site
    with its x-arm turned by rot and no null stream."""
    det = sig.c_struct()
    det.compute2arms = 0
    det.xax_deg = float(sig.det_xax) + float(rot)
    return det


def _grid_dev(f, n):
    f2 = np.asarray(f, dtype=np.float64)
    if f2.ndim == 1:
        f2 = f2[:, None] * np.ones((1, n))
    return E.to_dev(f2)


def _complete_for(wf, evParams):
    """The input completion of WFOverlap for one model (signal.py:1776-1812)."""
    ev = {k: np.atleast_1d(np.asarray(v, dtype=np.float64)) for k, v in evParams.items()}
    if 'chi1z' not in ev:
        try:
            ev['chi1z'] = ev['chiS'] + ev['chiA']
            ev['chi2z'] = ev['chiS'] - ev['chiA']
        except KeyError:
            raise ValueError('Two among chi1z, chi2z and chiS, chiA have to be provided.')
    if wf.is_tidal and 'LambdaTilde' not in ev and 'Lambda1' not in ev:
        raise ValueError('Two among Lambda1, Lambda2 and LambdaTilde and deltaLambda have to be provided.')
    return ev


def strain_device(sig, rows, f, arm, want_h=True, want_J=True, null_stream=True, wf=None, det=None):
    """h (N_f, N_ev, 2) and J (P, N_ev, N_f, 2) device tensors for one arm
    (of the model wf, the signal's own by default; det: a detector block
    other than the signal's own, e.g. from rot_detector)."""
    t = E.torch()
    wf = sig.wf_model if wf is None else wf
    E.check_rows(rows, wf)
    n = rows.shape[1]
    nf = f.shape[0]
    P = wf.nParams
    h = t.empty((nf, n, 2), dtype=t.float64, device=rows.device) if want_h else None
    J = t.empty((P, n, nf, 2), dtype=t.float64, device=rows.device) if want_J else None
    ws, nb = E.workspace(n)
    if det is None:
        det = sig.c_struct()
        if not null_stream:
            det.compute2arms = 0
    E.call('gwf_strain', wf.c_struct(), ctypes_byref(det), arm, N.ptr(rows), rows.stride(0), n, N.ptr(f), nf,
           N.ptr(h), N.ptr(J), N.ptr(ws), nb, N.stream_ptr())
    return h, J


def strain_components_device(sig, rows, f, arm, det=None):
    """(Ap, Ac, Psi) (3, N_f, N_ev) device tensor for one physical arm
    (det: a detector block other than the signal's own)."""
    t = E.torch()
    E.check_rows(rows, sig.wf_model)
    n = rows.shape[1]
    nf = f.shape[0]
    out = t.empty((3, nf, n), dtype=t.float64, device=rows.device)
    ws, nb = E.workspace(n)
    if det is None:
        det = sig.c_struct()
        det.compute2arms = 0
    E.call('gwf_strain_components', sig.wf_model.c_struct(), ctypes_byref(det), arm, N.ptr(rows), rows.stride(0), n,
           N.ptr(f), nf, N.ptr(out), N.ptr(ws), nb, N.stream_ptr())
    return out


def strain_device_into(sig, rows, f, arm, J):
    """Derivative rows of one arm into a caller buffer J (P, N_ev, N_f, 2)."""
    E.check_rows(rows, sig.wf_model)
    n = rows.shape[1]
    nf = f.shape[0]
    ws, nb = E.workspace(n)
    E.call('gwf_strain', sig.wf_model.c_struct(), ctypes_byref(sig.c_struct()), arm, N.ptr(rows), rows.stride(0), n,
           N.ptr(f), nf, None, N.ptr(J), N.ptr(ws), nb, N.stream_ptr())
    return J


def gram_device(J, f, S, P, out=None, accumulate=False, mode='fp64'):
    """Fisher integral of materialised derivatives J (P, n_ev, n_f, 2):
    mode 'fp64' (FP64 VALU) or 'split_fp32' (FP32 MFMA on hi/lo split
    operands with FP64 reduction)."""
    t = E.torch()
    n = J.shape[1]
    nf = J.shape[2]
    if mode not in ('fp64', 'split_fp32'):
        raise ValueError("mode must be 'fp64' or 'split_fp32'")
    if out is None:
        out = t.empty((P, P, n), dtype=t.float64, device=J.device)
    E.call('gwf_gram' if mode == 'fp64' else 'gwf_gram_split', N.ptr(J), N.ptr(f), N.ptr(S), P, n, nf,
           int(accumulate), N.ptr(out), N.stream_ptr())
    return out


def ctypes_byref(x):
    import ctypes
    return ctypes.byref(x)
