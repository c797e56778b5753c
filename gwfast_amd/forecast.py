"""Catalogue forecasts: SNR -> threshold -> Fisher -> fixParams -> CheckFisher
-> CovMatr -> sky area, with the reference driver's output files.

Follows run/calculate_forecasts_from_catalog.py: ``compute_errs`` (421-630)
and ``to_file`` (222-247), the per-chunk suffix ``_<i>_to_<j>`` and
``--resume_run`` (738-741: a chunk whose snrs file exists is skipped).  The
whole chain stays in device memory; each chunk's results are copied to the
host once, when written.  Differences from the driver: events come as an
in-memory dict (its h5 loader needs h5py), chunks run one after the other on
one GPU instead of a process pool, and an optional learned SNR screen
(gwfast_amd.screen) can replace the exact SNR pass.
"""
import os
import time

import numpy as np

from . import _engine as E
from . import fisherTools as FT
from .gwfastUtils import check_evparams


def _subset(events, sl):
    return {k: np.asarray(v)[sl] for k, v in events.items()}


def duty_masks(net, seeds, duty_factor, n):
    """calculate_forecasts_from_catalog.py:423-447: per detector, numpy's
    global generator reseeded with seeds[i], then one Bernoulli(duty_factor)
    draw of n per arm (arm order); keys as DetNet.SNR(return_all=True)."""
    masks = {}
    for i, (name, sig) in enumerate(net.signals.items()):
        np.random.seed(seeds[i])
        if sig.detector_shape == 'L':
            masks[name] = np.random.choice([0, 1], n, p=[1. - duty_factor, duty_factor])
        else:
            for a in range(3):
                masks[name + '_%s' % a] = np.random.choice([0, 1], n, p=[1. - duty_factor, duty_factor])
    return masks


def draw_seeds(net):
    """calculate_forecasts_from_catalog.py:856-860: one fresh seed per detector."""
    np.random.seed(None)
    return [np.random.randint(2 ** 32 - 1, size=1)[0] for _ in range(len(net.signals))]


def _duty_snr_fisher(net, rows, res, snr_th, masks):
    """Duty-cycled SNR (network SNR from the masked per-arm SNRs) and the
    masked network Fisher of the detected events, one fused launch per
    detector group (gwf_fisher_masked)."""
    from . import screen as SC
    from . import signal as S
    n = rows.shape[1]
    groups = net._groups()
    net_sq = np.zeros(n)
    group_masks = []
    for g in groups:
        sigs = [s for _, s in g]
        sq = S.snr_sq_device(sigs, rows, res).cpu().numpy()
        k = 0
        gm = []
        for name, s in g:
            keys = [name] if s.detector_shape == 'L' else [name + '_%s' % a for a in range(3)]
            m = np.array([masks[key] for key in keys])
            net_sq += (((2 * np.sqrt(sq[k:k + s.n_arms])) * m) ** 2).sum(axis=0)
            gm.append(m)
            k += s.n_arms
        group_masks.append(gm)
    snr = np.sqrt(net_sq)
    det = np.nonzero(snr > snr_th)[0]
    idx = E.to_dev(det.astype(np.int64), E.torch().int64)
    F = None
    if det.size:
        rd = SC.gather(rows, idx)
        for g, gm in zip(groups, group_masks):
            sigs = [s for _, s in g]
            Fg = S.fisher_masked_device(sigs, sigs[0].wf_model, rd, res, S.mask_words([m[:, det] for m in gm]))
            F = Fg if F is None else F + Fg
    return E.to_dev(snr), idx, F


def _arm_keys(name, sig):
    return [name] if sig.detector_shape == 'L' else [name + '_%s' % a for a in range(3)]


def _all_arms(net, rows, res, snr_th, masks=None):
    """--return_all (calculate_forecasts_from_catalog.py:421, 507-513): the
    per-arm SNRs of every event and the per-arm Fisher matrices of the
    detected events (gwf_fisher_arms, one launch per detector group), keyed as
    DetNet.SNR / FisherMatr(return_all=True), times the duty-cycle masks when
    given; 'net' holds the network SNR and the summed Fisher."""
    from . import screen as SC
    from . import signal as S
    n = rows.shape[1]
    groups = net._groups()
    snrs = {}
    net_sq = np.zeros(n)
    for g in groups:
        sq = S.snr_sq_device([s for _, s in g], rows, res).cpu().numpy()
        k = 0
        for name, s in g:
            for key in _arm_keys(name, s):
                snrs[key] = 2 * np.sqrt(sq[k]) * (masks[key] if masks is not None else 1.0)
                net_sq += snrs[key] ** 2
                k += 1
    snrs['net'] = np.sqrt(net_sq)
    det = np.nonzero(snrs['net'] > snr_th)[0]
    idx = E.to_dev(det.astype(np.int64), E.torch().int64)
    if not det.size:
        return snrs, idx, None
    rd = SC.gather(rows, idx)
    fishers, total = {}, None
    for g in groups:
        sigs = [s for _, s in g]
        n_arms = sum(s.n_arms for s in sigs)
        if n_arms <= S.ARMS_PER_LAUNCH:
            Fa = S.fisher_arms_device(sigs, sigs[0].wf_model, rd, res)
            arms = [Fa[a] for a in range(n_arms)]
        else:
            arms = [S.fisher_device(sigs, sigs[0].wf_model, rd, res, arm=a) for a in range(n_arms)]
        a = 0
        for name, s in g:
            for key in _arm_keys(name, s):
                F = arms[a] if masks is None else arms[a] * E.to_dev(masks[key][det])
                fishers[key] = F
                total = F.clone() if total is None else total + F
                a += 1
    fishers['net'] = total
    return snrs, idx, fishers


def compute_errs(events, net, snr_th=12.0, params_fix=(), res=1000, screen=None, band=(8.0, 16.0),
                 duty_factor=None, seeds=None, return_all=False):
    """One chunk: returns a dict with 'snrs' (network SNR of every event;
    with a screen, NaN where the exact SNR was not evaluated), 'idxs_det',
    'fisher' (P', P', n_det), 'cov', 'eps', 'sky_area', 'errors' (P', n_det),
    'cond_numbers', 'parNums'.  Device-resident until the end.  duty_factor
    < 1 applies the driver's per-arm duty-cycle masks (seeds: one per
    detector) to the SNRs and Fisher matrices.  return_all adds 'all_snrs'
    and 'all_fishers': per-arm dicts as DetNet's return_all (with params_fix
    applied to every matrix), the network Fisher being their sum."""
    t = E.torch()
    from . import screen as SC
    check_evparams(events)
    sig0 = next(iter(net.signals.values()))
    sig0._complete(events)
    wf = sig0.wf_model
    rows = sig0._rows(events)
    n = rows.shape[1]
    all_snrs = all_fishers = None
    if return_all:
        if screen is not None:
            raise NotImplementedError('return_all needs the exact per-arm SNRs; use it without the screen')
        masks = None
        if duty_factor is not None and duty_factor < 1:
            if seeds is None or len(seeds) != len(net.signals):
                raise ValueError('duty_factor needs one seed per detector')
            masks = duty_masks(net, seeds, duty_factor, n)
        all_snrs, idx, all_fishers = _all_arms(net, rows, res, snr_th, masks)
        snrs = E.to_dev(all_snrs['net'])
        F = all_fishers['net'] if all_fishers is not None else None
    elif duty_factor is not None and duty_factor < 1:
        if screen is not None:
            raise NotImplementedError('the learned screen predicts duty-free SNRs; use it without duty_factor')
        if seeds is None or len(seeds) != len(net.signals):
            raise ValueError('duty_factor needs one seed per detector')
        snrs, idx, F = _duty_snr_fisher(net, rows, res, snr_th, duty_masks(net, seeds, duty_factor, n))
    elif screen is None:
        snr = net.snr_device(rows, res)
        idx = (snr > snr_th).nonzero().flatten()
        F = net.fisher_device(SC.gather(rows, idx), res) if idx.numel() else None
        snrs = snr
    else:
        r = SC.screened_fisher(net, screen, rows, snr_th, band, res)
        idx, F = r['idx'], (r['F'] if r['idx'].numel() else None)
        snrs = t.full((n,), float('nan'), dtype=t.float64, device=rows.device)
        snrs[idx] = r['snr']
    parNums = dict(wf.ParNums)
    out = {'snrs': snrs.cpu().numpy(), 'idxs_det': idx.cpu().numpy()}
    if all_snrs is not None:
        out['all_snrs'] = all_snrs
    if F is None:
        out.update(fisher=None, parNums=parNums)
        return out
    if len(params_fix):
        fix_idx = sorted(parNums[p] for p in params_fix)
        F = FT.fix_params_device(F, fix_idx)
        if all_fishers is not None:
            all_fishers = {k: FT.fix_params_device(v, fix_idx) for k, v in all_fishers.items()}
        for pm in params_fix:
            for k in parNums.keys():
                if parNums[k] > parNums[pm]:
                    parNums[k] -= 1
            parNums.pop(pm, None)
    P, nd = F.shape[0], F.shape[2]
    ev_ = t.empty((nd, P), dtype=t.float64, device=F.device)
    cond = t.empty(nd, dtype=t.float64, device=F.device)
    from . import _native as N
    E.call('gwf_check_fisher', N.ptr(F), P, nd, N.ptr(ev_), None, N.ptr(cond), N.stream_ptr())
    C, eps, _, _ = FT.cov_device(F)
    theta = rows[wf.ParNums['theta']].index_select(0, idx).contiguous()
    sig, area = FT.sigma_sky_area_device(C, theta, parNums['theta'], parNums['phi'], 90, 'SqDeg')
    out.update(fisher=F.cpu().numpy(), cov=C.cpu().numpy(), eps=eps.cpu().numpy(), sky_area=area.cpu().numpy(),
               errors=sig.cpu().numpy(), cond_numbers=cond.cpu().numpy(), parNums=parNums)
    if all_fishers is not None:
        out['all_fishers'] = {k: v.cpu().numpy() for k, v in all_fishers.items()}
    return out


def to_file(res, out_path, suff=''):
    """calculate_forecasts_from_catalog.py:222-247 (file names and layouts)."""
    np.savetxt(os.path.join(out_path, 'snrs' + suff + '.txt'), res['snrs'])
    # --return_all (to_file_all, 337-380): the per-arm SNRs and Fisher matrices,
    # as .npz archives keyed like the driver's hdf5 groups (h5py is absent)
    if res.get('all_snrs') is not None:
        np.savez(os.path.join(out_path, 'all_snrs' + suff + '.npz'), **res['all_snrs'])
    if res.get('all_fishers') is not None:
        np.savez(os.path.join(out_path, 'all_fishers' + suff + '.npz'), **res['all_fishers'])
    if res.get('fisher') is None:
        return
    np.save(os.path.join(out_path, 'fishers' + suff + '.npy'), res['fisher'])
    np.save(os.path.join(out_path, 'covs' + suff + '.npy'), res['cov'])
    np.savetxt(os.path.join(out_path, 'sky_area' + suff + '.txt'), res['sky_area'])
    np.savetxt(os.path.join(out_path, 'errors' + suff + '.txt'), res['errors'])
    np.savetxt(os.path.join(out_path, 'inversion_errors' + suff + '.txt'), res['eps'])
    np.savetxt(os.path.join(out_path, 'cond_numbers' + suff + '.txt'), res['cond_numbers'])
    np.savetxt(os.path.join(out_path, 'idxs_det' + suff + '.txt'), res['idxs_det'][:, None])


def run_catalog(events, net, out_path, batch_size=10000, snr_th=12.0, params_fix=(), res=1000, resume=False,
                screen=None, band=(8.0, 16.0), idx_in=0, verbose=True, duty_factor=None, seeds=None,
                return_all=False):
    """Chunked forecast of a catalogue with per-chunk files (suffix
    _<i>_to_<j>, global event indices from idx_in).  With resume=True a chunk
    whose snrs file exists is skipped.  duty_factor (< 1) and seeds (one per
    detector; drawn when None, as the driver does) apply the duty cycle to
    every chunk with the same seeds; return_all also writes the per-arm
    SNRs and Fisher matrices (all_snrs / all_fishers<suffix>.npz).  Returns
    the list of suffixes."""
    os.makedirs(out_path, exist_ok=True)
    n = len(np.atleast_1d(events['dL']))
    if duty_factor is not None and duty_factor >= 1:
        duty_factor = None
    if duty_factor is not None and seeds is None:
        seeds = draw_seeds(net)
    suffs = []
    for i0 in range(0, n, batch_size):
        i1 = min(n, i0 + batch_size)
        suff = '_%d_to_%d' % (idx_in + i0, idx_in + i1)
        suffs.append(suff)
        if resume and os.path.exists(os.path.join(out_path, 'snrs' + suff + '.txt')):
            if verbose:
                print('Batch already present, skipping...', flush=True)
            continue
        t0 = time.perf_counter()
        r = compute_errs(_subset(events, slice(i0, i1)), net, snr_th, params_fix, res, screen, band, duty_factor, seeds,
                         return_all)
        r['idxs_det'] = r['idxs_det'] + idx_in + i0
        to_file(r, out_path, suff)
        if verbose:
            print('chunk %s: %d events, %d detected, %.2f s' % (suff, i1 - i0, len(r['idxs_det']),
                                                                 time.perf_counter() - t0), flush=True)
    return suffs
