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


def compute_errs(events, net, snr_th=12.0, params_fix=(), res=1000, screen=None, band=(8.0, 16.0)):
    """One chunk: returns a dict with 'snrs' (network SNR of every event;
    with a screen, NaN where the exact SNR was not evaluated), 'idxs_det',
    'fisher' (P', P', n_det), 'cov', 'eps', 'sky_area', 'errors' (P', n_det),
    'cond_numbers', 'parNums'.  Device-resident until the end."""
    t = E.torch()
    from . import screen as SC
    check_evparams(events)
    sig0 = next(iter(net.signals.values()))
    sig0._complete(events)
    wf = sig0.wf_model
    rows = E.to_dev(E.param_rows(events, wf.is_tidal))
    n = rows.shape[1]
    if screen is None:
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
    if F is None:
        out.update(fisher=None, parNums=parNums)
        return out
    if len(params_fix):
        fix_idx = sorted(parNums[p] for p in params_fix)
        F = FT.fix_params_device(F, fix_idx)
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
    return out


def to_file(res, out_path, suff=''):
    """calculate_forecasts_from_catalog.py:222-247 (file names and layouts)."""
    np.savetxt(os.path.join(out_path, 'snrs' + suff + '.txt'), res['snrs'])
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
                screen=None, band=(8.0, 16.0), idx_in=0, verbose=True):
    """Chunked forecast of a catalogue with per-chunk files (suffix
    _<i>_to_<j>, global event indices from idx_in).  With resume=True a chunk
    whose snrs file exists is skipped.  Returns the list of suffixes."""
    os.makedirs(out_path, exist_ok=True)
    n = len(np.atleast_1d(events['dL']))
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
        r = compute_errs(_subset(events, slice(i0, i1)), net, snr_th, params_fix, res, screen, band)
        r['idxs_det'] = r['idxs_det'] + idx_in + i0
        to_file(r, out_path, suff)
        if verbose:
            print('chunk %s: %d events, %d detected, %.2f s' % (suff, i1 - i0, len(r['idxs_det']),
                                                                 time.perf_counter() - t0), flush=True)
    return suffs
