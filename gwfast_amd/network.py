"""Detector networks with gwfast's DetNet API (gwfast/network.py:12-152).

When every detector of the network shares the waveform model, fmin and fmax
(the usual case) the whole network is evaluated by ONE fused launch per
quantity: each event's waveform and tangents are computed once per frequency
and projected onto every arm of every detector before the Gram update.
"""
import numpy as np

from . import signal as S
from .gwfastUtils import check_evparams


class DetNet(object):
    """gwfast/network.py:12-30."""

    def __init__(self, signals, verbose=True):
        self.signals = signals
        self.verbose = verbose

    def _clear_cache(self):
        """network.py:31-33."""
        for sig in self.signals.values():
            sig._clear_cache()

    def _update_all_seeds(self, seeds=[], verbose=True):
        """network.py:36-51: reseed the duty-cycle draw of each detector
        (None draws a fresh seed)."""
        if seeds == []:
            seeds = [None] * len(self.signals)
        for i, (name, sig) in enumerate(self.signals.items()):
            sig._update_seed(seed=seeds[i])
            if verbose:
                print('\nSeed for detector %s is %s' % (name, sig.seedUse))

    def _groups(self):
        """Detectors that can share one launch (same model object, fmin, fmax)."""
        groups = []
        for name, sig in self.signals.items():
            for g in groups:
                s0 = g[0][1]
                if (sig.wf_model is s0.wf_model and sig.fmin == s0.fmin and sig.fmax == s0.fmax
                        and len(g) < 8):
                    g.append((name, sig))
                    break
            else:
                groups.append([(name, sig)])
        return groups

    def _prepare(self, evParams):
        check_evparams(evParams)
        sig0 = next(iter(self.signals.values()))
        sig0._complete(evParams)

    def SNR(self, evParams, res=1000, return_all=False):
        """network.py:53-81: sqrt(sum_d SNR_d^2)."""
        self._prepare(evParams)
        snrs = {}
        for g in self._groups():
            sigs = [s for _, s in g]
            rows = sigs[0]._rows(evParams)
            sq = S.snr_sq_device(sigs, rows, res).cpu().numpy()
            k = 0
            for name, s in g:
                part = sq[k:k + s.n_arms]
                k += s.n_arms
                masks = s._duty_masks(part.shape[1])
                if masks is not None:
                    part = part * masks
                if s.detector_shape == 'T' and return_all:
                    for i in range(3):
                        snrs[name + '_%s' % i] = 2 * np.sqrt(part[i])
                else:
                    snrs[name] = 2 * np.sqrt(part.sum(axis=0))
        net_snr = np.sqrt(np.array([snrs[k] ** 2 for k in snrs.keys()]).sum(axis=0))
        if return_all:
            snrs['net'] = net_snr
            return snrs
        return net_snr

    def FisherMatr(self, evParams, return_all=False, return_derivatives=False, return_SNR_derivatives=False,
                   **kwargs):
        """network.py:84-152: sum of the detectors' Fisher matrices."""
        self._prepare(evParams)
        res = kwargs.pop('res', 1000)
        res = None if res is None else (int(res) if np.ndim(res) == 0 else np.asarray(res))
        use_m1m2 = kwargs.get('use_m1m2', False)
        use_chi1chi2 = kwargs.get('use_chi1chi2', True)
        df, spacing = kwargs.get('df'), kwargs.get('spacing', 'geom')
        S._check_fisher_options(res, df, spacing, use_m1m2, use_chi1chi2, kwargs.get('computeDerivFinDiff', False))
        lin = spacing == 'lin'
        opts = {'use_m1m2': use_m1m2, 'use_chi1chi2': use_chi1chi2, 'df': df, 'spacing': spacing}
        if return_derivatives or return_SNR_derivatives:
            allF, allD = {}, {}
            for name, sig in self.signals.items():
                F_, D_ = sig.FisherMatr(evParams, res=res, return_derivatives=return_derivatives,
                                        return_SNR_derivatives=return_SNR_derivatives, **opts)
                if sig.detector_shape == 'T':
                    for i in range(3):
                        allF[name + '_%s' % i] = F_[i]
                        allD[name + '_%s' % i] = D_[i]
                else:
                    allF[name] = F_[0]
                    allD[name] = D_[0]
            if return_SNR_derivatives:
                # the reference's network SNR here is at SNR's default resolution (network.py:139-143)
                SNRs_net = kwargs['SNRs'] if 'SNRs' in kwargs else self.SNR(evParams, return_all=False)
                allD['net'] = np.array([allD[k] for k in allD.keys()]).sum(axis=0) / SNRs_net
            allF['net'] = np.array([allF[k] for k in allF.keys()]).sum(axis=0)
            return allF, allD
        if not return_all and any(sig.DutyFactor is not None for sig in self.signals.values()):
            total, rows = None, None
            n = None
            for g in self._groups():
                sigs = [s for _, s in g]
                rows = sigs[0]._rows(evParams)
                n = rows.shape[1]
                masks = [s._duty_masks(n) for s in sigs]
                masks = [m if m is not None else np.ones((s.n_arms, n)) for m, s in zip(masks, sigs)]
                F = S.fisher_masked_device(sigs, sigs[0].wf_model, rows, S.resolve_res(sigs[0], evParams, res, df),
                                           S.mask_words(masks), lin=lin)
                total = F if total is None else total + F
            S.reparam_device(total, rows, use_m1m2, use_chi1chi2)
            return total.cpu().numpy()
        if return_all:
            # per-arm matrices, one gwf_fisher_arms launch per group of detectors; the
            # duty-cycle masks are drawn first, detector by detector, as the per-detector
            # calls of the reference draw them
            n = next(iter(self.signals.values()))._rows(evParams).shape[1]
            masks = {name: sig._duty_masks(n) for name, sig in self.signals.items()}
            per_det = {}
            for g in self._groups():
                sigs = [s for _, s in g]
                r = S.resolve_res(sigs[0], evParams, res, df)
                rows = sigs[0]._rows(evParams)
                Fa = None
                if sum(s.n_arms for s in sigs) <= S.ARMS_PER_LAUNCH:
                    Fa = S.fisher_arms_device(sigs, sigs[0].wf_model, rows, r, lin=lin)
                a = 0
                for name, sig in g:
                    part = Fa[a:a + sig.n_arms] if Fa is not None else None
                    per_det[name] = sig._fisher_per_arm(rows, r, lin, use_m1m2, use_chi1chi2, masks[name], Fa=part)
                    a += sig.n_arms
            allF = {}
            for name, sig in self.signals.items():
                F_ = per_det[name]
                if sig.detector_shape == 'T':
                    for i in range(3):
                        allF[name + '_%s' % i] = F_[i]
                else:
                    allF[name] = F_[0]
            allF['net'] = np.array([allF[k] for k in allF.keys()]).sum(axis=0)
            return allF if return_all else allF['net']
        total, rows = None, None
        for g in self._groups():
            sigs = [s for _, s in g]
            rows = sigs[0]._rows(evParams)
            F = S.fisher_device(sigs, sigs[0].wf_model, rows, S.resolve_res(sigs[0], evParams, res, df), lin=lin)
            total = F if total is None else total + F
        S.reparam_device(total, rows, use_m1m2, use_chi1chi2)
        return total.cpu().numpy()

    def WFOverlap(self, WF1, WF2, evParams1, evParams2, res=1000, **kwargs):
        """network.py:198-232: overlap summed over the detectors, normalised by
        the network SNRs of the two waveforms."""
        overlap_all = 0.
        SNR1_all = 0.
        SNR2_all = 0.
        for sig in self.signals.values():
            overlap_int, SNR1, SNR2 = sig.WFOverlap(WF1, WF2, evParams1, evParams2, res=res, return_separate=True)
            overlap_all = overlap_all + overlap_int
            SNR1_all = SNR1_all + SNR1 ** 2
            SNR2_all = SNR2_all + SNR2 ** 2
        return overlap_all / np.sqrt(SNR1_all * SNR2_all)

    # ------------------------------------------------------- device-resident
    def fisher_device(self, rows, res=1000, out=None):
        """Network Fisher of device-resident parameter rows (P_rows, N_ev):
        returns a (P, P, N_ev) device tensor.  Single fused launch when the
        network forms one group."""
        groups = self._groups()
        total = out
        for gi, g in enumerate(groups):
            sigs = [s for _, s in g]
            if gi == 0:
                total = S.fisher_device(sigs, sigs[0].wf_model, rows, res, out=total)
            else:
                total += S.fisher_device(sigs, sigs[0].wf_model, rows, res)
        return total

    def snr_device(self, rows, res=1000):
        """Network SNR of device-resident parameter rows: (N_ev,) tensor."""
        total = None
        for g in self._groups():
            sigs = [s for _, s in g]
            sq = S.snr_sq_device(sigs, rows, res).sum(dim=0)
            total = sq if total is None else total + sq
        return 2.0 * total.sqrt()
