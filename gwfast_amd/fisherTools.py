"""Covariance inversion and Fisher checks with gwfast's fisherTools API.

Mirrors gwfast/fisherTools.py:32-196 (CovMatr), 199-210
(compute_inversion_error) and 216-279 (CheckFisher).  The reference runs a
per-event mpmath loop at 53-bit precision and stores float128; here each
event is inverted by one GPU thread in FP64 (gwf_cov_invert /
gwf_check_fisher) and float64 is returned.
"""
import numpy as np

from . import _engine as E
from . import _native as N

METHOD_NAMES = {0: 'cho', 1: 'svd', -1: 'nan'}


def _dev(F):
    t = E.torch()
    if isinstance(F, t.Tensor):
        return F.to(device=E.device(), dtype=t.float64).contiguous()
    return E.to_dev(np.asarray(F, dtype=np.float64))


def cov_device(Fd, want_eps=True):
    """(Cov, eps, cond, method) device tensors for a (P, P, N) device Fisher."""
    t = E.torch()
    P, n = Fd.shape[0], Fd.shape[2]
    C = t.empty_like(Fd)
    eps = t.empty(n, dtype=t.float64, device=Fd.device) if want_eps else None
    cond = t.empty(n, dtype=t.float64, device=Fd.device)
    meth = t.empty(n, dtype=t.int32, device=Fd.device)
    E.call('gwf_cov_invert', N.ptr(Fd), P, n, N.ptr(C), N.ptr(eps), N.ptr(cond), N.ptr(meth), N.stream_ptr())
    return C, eps, cond, meth


def CovMatr(FisherMatrix, invMethodIn='cho', condNumbMax=1e50, truncate=False, svals_thresh=1e-15, verbose=False,
            alt_method='svd', return_details=False):
    """fisherTools.py:32-196 for invMethodIn='cho', alt_method='svd'
    (the catalogue driver's settings, calculate_forecasts_from_catalog.py:613-619).
    Returns (Cov (P,P,N), eps (N,)); return_details adds (cond, method)."""
    if invMethodIn != 'cho' or alt_method != 'svd':
        raise NotImplementedError("only invMethodIn='cho' with alt_method='svd' is on this path")
    if truncate:
        raise NotImplementedError('singular-value truncation is not supported')
    Fd = _dev(FisherMatrix)
    C, eps, cond, meth = cov_device(Fd)
    out = (C.cpu().numpy(), eps.cpu().numpy())
    if verbose:
        m = meth.cpu().numpy()
        print('Inversion error: min=%s, max=%s; cholesky on %d, svd on %d, failed %d' % (
            np.nanmin(out[1]), np.nanmax(out[1]), (m == 0).sum(), (m == 1).sum(), (m == -1).sum()))
    if return_details:
        return out + (cond.cpu().numpy(), meth.cpu().numpy())
    return out


def compute_inversion_error(Fisher, Cov):
    """fisherTools.py:199-210: max |Cov F - I| per event."""
    Fd, Cd = _dev(Fisher), _dev(Cov)
    t = E.torch()
    P, n = Fd.shape[0], Fd.shape[2]
    eps = t.empty(n, dtype=t.float64, device=Fd.device)
    E.call('gwf_inversion_error', N.ptr(Fd), N.ptr(Cd), P, n, N.ptr(eps), N.stream_ptr())
    return eps.cpu().numpy()


def CheckFisher(FisherM, condNumbMax=1.0e15, use_mpmath=True, verbose=False):
    """fisherTools.py:216-279: (evals (N,P) ascending, evecs (N,P,P) with
    eigenvectors as columns, condition number (N,))."""
    Fd = _dev(FisherM)
    t = E.torch()
    P, n = Fd.shape[0], Fd.shape[2]
    ev = t.empty((n, P), dtype=t.float64, device=Fd.device)
    vec = t.empty((n, P, P), dtype=t.float64, device=Fd.device)
    cond = t.empty(n, dtype=t.float64, device=Fd.device)
    E.call('gwf_check_fisher', N.ptr(Fd), P, n, N.ptr(ev), N.ptr(vec), N.ptr(cond), N.stream_ptr())
    evals, evecs, cond = ev.cpu().numpy(), vec.cpu().numpy(), cond.cpu().numpy()
    if np.any(evals <= 0.):
        print('WARNING: one or more eigenvalues are negative at position(s) %s' % str(np.unique(np.where(evals < 0)[0])))
    if verbose:
        if np.any(cond > condNumbMax):
            print('WARNING: the condition number is too large (%s>%s)' % (cond, condNumbMax))
        else:
            print('Condition number= %s . Ok. ' % cond)
    return evals, evecs, cond


# ------------------------------------------------------------ post-Fisher
def fix_params_device(Fd, fix_idx, prior_diag=None):
    """(Pn, Pn, N) device tensor: Fd with the rows / columns fix_idx removed
    (and prior_diag (P,) added to the diagonal first)."""
    t = E.torch()
    P, n = Fd.shape[0], Fd.shape[2]
    mask = 0
    for i in fix_idx:
        mask |= 1 << int(i)
    Pn = P - bin(mask).count('1')
    out = t.empty((Pn, Pn, n), dtype=t.float64, device=Fd.device)
    pd = None if prior_diag is None else E.to_dev(np.asarray(prior_diag, dtype=np.float64))
    E.call('gwf_fix_params', N.ptr(Fd), P, n, mask, N.ptr(pd), N.ptr(out), N.stream_ptr())
    return out


def fixParams(MatrIn, ParNums_inp, ParMarg):
    """fisherTools.py:340-375: remove the rows and columns of the parameters
    ParMarg; returns (Fisher (Pn, Pn, N), ParNums renumbered)."""
    ParNums = dict(ParNums_inp)
    idx = sorted(ParNums[p] for p in ParMarg)
    out = fix_params_device(_dev(MatrIn), idx).cpu().numpy()
    for pm in ParMarg:
        for k in ParNums.keys():
            if ParNums[k] > ParNums[pm]:
                ParNums[k] -= 1
        ParNums.pop(pm, None)
    return out, ParNums


def addPrior(Matr, vals, ParNums, ParAdd):
    """fisherTools.py:378-403: Gaussian priors vals added on the diagonal
    entries of the parameters ParAdd."""
    M = np.asarray(Matr, dtype=np.float64)
    two_d = M.ndim == 2
    Fd = _dev(M[:, :, None] if two_d else M)
    diag = np.zeros(Fd.shape[0])
    diag[np.sort(np.array([ParNums[p] for p in ParAdd]))] = vals
    out = fix_params_device(Fd, [], prior_diag=diag).cpu().numpy()
    return out[:, :, 0] if two_d else out


def sigma_sky_area_device(Cd, thetas, i_theta=3, i_phi=4, perc_level=90, units='SqDeg'):
    """(sigma (P, N), sky area (N,)) device tensors from a device covariance."""
    t = E.torch()
    P, n = Cd.shape[0], Cd.shape[2]
    sig = t.empty((P, n), dtype=t.float64, device=Cd.device)
    area = t.empty(n, dtype=t.float64, device=Cd.device)
    th = thetas if isinstance(thetas, t.Tensor) else E.to_dev(np.asarray(thetas, dtype=np.float64))
    th = th.to(dtype=t.float64).contiguous()
    E.call('gwf_sigma_sky_area', N.ptr(Cd), P, n, N.ptr(th), int(i_theta), int(i_phi), float(perc_level),
           int(units == 'SqDeg'), N.ptr(sig), N.ptr(area), N.stream_ptr())
    return sig, area


def compute_localization_region(Cov, parNum, thFid, perc_level=90, units='SqDeg'):
    """fisherTools.py:830-858 (Barack & Cutler 2004): the perc_level sky
    area, in deg^2 ('SqDeg') or sr ('Sterad')."""
    if units not in ('SqDeg', 'Sterad'):
        return None
    Cd = _dev(Cov)
    th = np.broadcast_to(np.asarray(thFid, dtype=np.float64), (Cd.shape[2],))
    _, area = sigma_sky_area_device(Cd, th, parNum['theta'], parNum['phi'], perc_level, units)
    return area.cpu().numpy()
