"""Waveform models of the event-batched path, with gwfast's WaveFormModel API.

``TaylorF2_RestrictedPN``, ``IMRPhenomD`` and ``IMRPhenomD_NRTidalv2`` keep
the reference's constructors, flags, ``ParNums``/``nParams`` and methods
(gwfast/waveforms.py:49-199, 697-953, 959-1333, 1339-1837).  ``Phi``, ``Ampl``,
``tau_star`` and ``fcut`` are evaluated by the HIP library
(gwf_waveform_eval / gwf_fcut); inside SNR and Fisher calls the same device
code runs fused with the projection and the integrals.
"""
import os
from abc import ABC

import numpy as np

from . import _engine as E
from . import _native as N
from . import gwfastGlobals as glob


class WaveFormModel(ABC):
    """gwfast/waveforms.py:49-199 (aligned-spin subset: no precession, no
    higher modes; eccentricity for TaylorF2)."""

    def __init__(self, objType, fcutPar, is_newtonian=False, is_tidal=False, is_HigherModes=False,
                 is_chi1chi2=True, is_Precessing=False, is_LAL=False, is_prec_ang=False, is_eccentric=False,
                 is_holomorphic=False, apply_fcut=True):
        if is_newtonian or is_HigherModes or is_Precessing or is_LAL:
            raise NotImplementedError('only aligned-spin, dominant-mode models are on this path')
        self.objType = objType
        self.fcutPar = fcutPar
        self.is_newtonian = is_newtonian
        self.is_tidal = is_tidal
        self.is_HigherModes = is_HigherModes
        self.is_chi1chi2 = is_chi1chi2
        self.is_Precessing = is_Precessing
        self.is_LAL = is_LAL
        self.is_prec_ang = is_prec_ang
        self.is_eccentric = is_eccentric
        self.is_holomorphic = is_holomorphic
        self.apply_fcut = apply_fcut
        # Fisher row order (waveforms.py:78, 106-111, 125-129)
        self.ParNums = {'Mc': 0, 'eta': 1, 'dL': 2, 'theta': 3, 'phi': 4, 'iota': 5, 'psi': 6, 'tcoal': 7,
                        'Phicoal': 8, 'chiS': 9, 'chiA': 10}
        self.nParams = 11
        if is_tidal:
            self.ParNums['LambdaTilde'] = 11
            self.ParNums['deltaLambda'] = 12
            self.nParams = 13
        if is_eccentric:  # after the tidal rows when both (waveforms.py:106-115, 124-126)
            self.ParNums['ecc'] = self.nParams
            self.nParams += 1
        if is_chi1chi2:
            self.ParNums['chi1z'] = self.ParNums.pop('chiS')
            self.ParNums['chi2z'] = self.ParNums.pop('chiA')
        self.ParNums = dict(sorted(self.ParNums.items(), key=lambda item: item[1]))
        self._cstruct = None

    # ------------------------------------------------------------------ ABI
    def _desc(self):
        raise NotImplementedError

    def c_struct(self):
        if self._cstruct is None:
            self._cstruct = self._desc()
        return self._cstruct

    def _eval(self, f, kwargs, what):
        ev = dict(kwargs)
        f = np.asarray(f, dtype=np.float64)
        n = len(np.atleast_1d(ev['Mc']))
        for k in ('dL', 'theta', 'phi', 'iota', 'psi', 'tcoal', 'Phicoal'):
            ev.setdefault(k, np.ones(n) if k == 'dL' else np.zeros(n))
        if 'chi1z' not in ev:
            ev['chi1z'], ev['chi2z'] = np.zeros(n), np.zeros(n)
        scalar_ev = np.ndim(ev['Mc']) == 0
        if f.ndim == 0:
            f2 = np.full((1, n), float(f))
        elif f.ndim == 1:
            f2 = np.repeat(f[:, None], n, axis=1) if (scalar_ev or f.shape[0] != n) else f[None, :]
        else:
            f2 = f
        rows = E.param_rows({k: np.atleast_1d(v) for k, v in ev.items()}, self.is_tidal, self.is_eccentric) \
            if not self.is_tidal else self._tidal_rows(ev)
        t = E.torch()
        fd = E.to_dev(f2)
        pd = E.to_dev(rows)
        out = t.empty_like(fd)
        ws, nb = E.workspace(n)
        ptrs = {'A': None, 'P': None, 'T': None}
        ptrs[what] = N.ptr(out)
        E.call('gwf_waveform_eval', self.c_struct(), N.ptr(pd), n, n, N.ptr(fd), f2.shape[0], ptrs['A'], ptrs['P'],
               ptrs['T'], N.ptr(ws), nb, N.stream_ptr())
        res = out.cpu().numpy()
        if f.ndim == 0:
            return res[0] if not scalar_ev else res[0, 0]
        if f.ndim == 1 and not (scalar_ev or f.shape[0] != n):
            return res[0]
        if scalar_ev:
            return res[:, 0]
        return res

    def _tidal_rows(self, ev):
        ev = {k: np.atleast_1d(v) for k, v in ev.items()}
        return E.param_rows(ev, True, self.is_eccentric)

    def Phi(self, f, **kwargs):
        """GW phase on the grid f (waveforms.py:148-160)."""
        return self._eval(f, kwargs, 'P')

    def Ampl(self, f, **kwargs):
        """GW amplitude on the grid f (waveforms.py:162-174)."""
        return self._eval(f, kwargs, 'A')

    def tau_star(self, f, **kwargs):
        """3.5PN time to coalescence in seconds (waveforms.py:878-901)."""
        return self._eval(f, kwargs, 'T')

    def fcut(self, **kwargs):
        """Cut frequency in Hz (waveforms.py:190-199, 903-953, 1324-1333)."""
        ev = {k: np.atleast_1d(np.asarray(v, dtype=np.float64)) for k, v in kwargs.items()}
        n = len(ev['Mc'])
        for k in ('dL', 'theta', 'phi', 'iota', 'psi', 'tcoal', 'Phicoal', 'chi1z', 'chi2z'):
            ev.setdefault(k, np.zeros(n))
        if self.is_tidal and 'Lambda1' not in ev and 'LambdaTilde' not in ev:
            ev['Lambda1'] = ev['Lambda2'] = np.zeros(n)
        if self.is_eccentric:
            ev.setdefault('ecc', np.zeros(n))  # the cut does not depend on it
        rows = E.param_rows(ev, self.is_tidal, self.is_eccentric)
        t = E.torch()
        pd = E.to_dev(rows)
        out = t.empty(n, dtype=t.float64, device=E.device())
        E.call('gwf_fcut', self.c_struct(), N.ptr(pd), n, n, 0.0, N.ptr(out), N.stream_ptr())
        res = out.cpu().numpy()
        return res[0] if np.ndim(kwargs['Mc']) == 0 else res


class TaylorF2_RestrictedPN(WaveFormModel):
    """TaylorF2 3.5PN restricted waveform (waveforms.py:697-953), optionally
    with 5 and 6PN tidal terms or with the low-eccentricity 3PN phase
    (is_eccentric, fRef_ecc; waveforms.py:814-845), or both (14 parameters)."""

    def __init__(self, fHigh=None, is_tidal=False, use_3p5PN_SpinHO=False, phiref_vlso=False, is_eccentric=False,
                 fRef_ecc=None, which_ISCO='Schw', use_QuadMonTid=False, **kwargs):
        if fHigh is None:
            fHigh = 1. / (6. * np.pi * np.sqrt(6.) * glob.GMsun_over_c3)
        if which_ISCO not in ('Schw', 'Kerr'):
            raise ValueError("which_ISCO must be 'Schw' or 'Kerr'")
        self.use_3p5PN_SpinHO = use_3p5PN_SpinHO
        self.phiref_vlso = phiref_vlso
        self.fRef_ecc = fRef_ecc
        self.which_ISCO = which_ISCO
        self.use_QuadMonTid = use_QuadMonTid
        super().__init__('BNS' if is_tidal else 'BBH', fHigh, is_tidal=is_tidal, is_eccentric=is_eccentric,
                         is_holomorphic=True, **kwargs)

    def _desc(self):
        w = N.GwfWaveform()
        w.model = N.MODEL_TAYLORF2
        w.is_tidal = int(self.is_tidal)
        w.use_3p5PN_SpinHO = int(self.use_3p5PN_SpinHO)
        w.phiref_vlso = int(self.phiref_vlso)
        w.kerr_isco = int(self.which_ISCO == 'Kerr')
        w.use_quadmon = int(self.use_QuadMonTid)
        w.apply_fcut = int(self.apply_fcut)
        w.fcut_par = float(self.fcutPar)
        w.is_eccentric = int(bool(self.is_eccentric))
        w.fref_ecc = float(self.fRef_ecc) if self.fRef_ecc is not None else 0.0
        return w


class IMRPhenomD(WaveFormModel):
    """IMRPhenomD (waveforms.py:959-1333).  fRef=None uses the minimum of the
    frequency grid as reference frequency, as the reference does."""

    def __init__(self, fRef=None, **kwargs):
        self.AMP_fJoin_INS = 0.014
        self.PHI_fJoin_INS = 0.018
        self.fRef = fRef
        super().__init__('BBH', 0.2, **kwargs)
        self.QNMgrid_a = np.loadtxt(os.path.join(glob.WFfilesPath, 'QNMData_a.txt'))
        self.QNMgrid_fring = np.loadtxt(os.path.join(glob.WFfilesPath, 'QNMData_fring.txt'))
        self.QNMgrid_fdamp = np.loadtxt(os.path.join(glob.WFfilesPath, 'QNMData_fdamp.txt'))
        self._qnm_dev = None

    def _desc(self):
        if self._qnm_dev is None:
            self._qnm_dev = [E.to_dev(self.QNMgrid_a), E.to_dev(self.QNMgrid_fring), E.to_dev(self.QNMgrid_fdamp)]
        w = N.GwfWaveform()
        w.model = N.MODEL_IMRPHENOMD
        w.apply_fcut = int(self.apply_fcut)
        w.fcut_par = float(self.fcutPar)
        w.fref = float(self.fRef) if self.fRef is not None else 0.0
        w.n_qnm = len(self.QNMgrid_a)
        w.qnm_a = self._qnm_dev[0].data_ptr()
        w.qnm_fring = self._qnm_dev[1].data_ptr()
        w.qnm_fdamp = self._qnm_dev[2].data_ptr()
        return w


class IMRPhenomD_NRTidalv2(IMRPhenomD):
    """IMRPhenomD_NRTidalv2 (waveforms.py:1339-1837): IMRPhenomD with the
    NRTidalv2 tidal phase and amplitude, the spin-induced quadrupole (and
    octupole) moments from Lambda1, Lambda2, and a Planck taper that ends the
    waveform at 1.2 f_merger; 13 parameters (LambdaTilde, deltaLambda).  The
    taper's end point itself is cut (value 0): the reference's expression
    gives 1 there, at a frequency the grid's last point hits or misses by
    ulp-level rounding."""

    def __init__(self, fRef=None, **kwargs):
        super().__init__(fRef=fRef, **kwargs)
        self.objType = 'BNS'
        self.is_tidal = True
        self.ParNums = {'Mc': 0, 'eta': 1, 'dL': 2, 'theta': 3, 'phi': 4, 'iota': 5, 'psi': 6, 'tcoal': 7,
                        'Phicoal': 8, 'chi1z': 9, 'chi2z': 10, 'LambdaTilde': 11, 'deltaLambda': 12}
        if not self.is_chi1chi2:
            self.ParNums['chiS'] = self.ParNums.pop('chi1z')
            self.ParNums['chiA'] = self.ParNums.pop('chi2z')
            self.ParNums = dict(sorted(self.ParNums.items(), key=lambda item: item[1]))
        self.nParams = 13

    def _desc(self):
        w = super()._desc()
        w.model = N.MODEL_IMRPHENOMD_NRTIDALV2
        w.is_tidal = 1
        return w
