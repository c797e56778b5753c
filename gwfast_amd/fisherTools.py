"""Covariance inversion and Fisher checks with gwfast's fisherTools API.

Mirrors gwfast/fisherTools.py:32-196 (CovMatr, every inversion method),
199-210 (compute_inversion_error), 216-279 (CheckFisher) and the helpers
after them.  The reference runs a
per-event mpmath loop at 53-bit precision and stores float128; here each
event is inverted by one GPU wavefront in FP64, its matrices in LDS
(gwf_cov_invert / gwf_check_fisher), and float64 is returned.
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


INV_METHODS = {'cho': 0, 'svd': 1, 'inv': 2, 'svd_reg': 3, 'lu': 4}


def cov_device(Fd, want_eps=True, invMethodIn='cho', alt_method='svd', truncate=False, svals_thresh=1e-15,
               condNumbMax=1e50):
    """(Cov, eps, cond, method) device tensors for a (P, P, N) device Fisher;
    method holds the INV_METHODS id used per event (-1: NaN)."""
    t = E.torch()
    for name in (invMethodIn, alt_method):
        if name not in INV_METHODS:
            raise ValueError("Inversion method must be one of 'inv', 'cho', 'svd', 'svd_reg', 'lu'")
    if invMethodIn == 'cho' and alt_method == 'cho':
        raise ValueError('alt_method has to be different from invMethodIn')
    P, n = Fd.shape[0], Fd.shape[2]
    C = t.empty_like(Fd)
    eps = t.empty(n, dtype=t.float64, device=Fd.device) if want_eps else None
    cond = t.empty(n, dtype=t.float64, device=Fd.device)
    meth = t.empty(n, dtype=t.int32, device=Fd.device)
    E.call('gwf_cov_invert_ex', N.ptr(Fd), P, n, INV_METHODS[invMethodIn], INV_METHODS[alt_method], int(bool(truncate)),
           float(svals_thresh), float(condNumbMax), N.ptr(C), N.ptr(eps), N.ptr(cond), N.ptr(meth), N.stream_ptr())
    return C, eps, cond, meth


def CovMatr(FisherMatrix, invMethodIn='cho', condNumbMax=1e50, truncate=False, svals_thresh=1e-15, verbose=False,
            alt_method='svd', return_details=False):
    """fisherTools.py:32-196 with every inversion method ('cho', 'inv',
    'svd' with optional truncation, 'svd_reg', 'lu') and alt_method.
    Returns (Cov (P,P,N), eps (N,)); return_details adds (cond, method)."""
    Fd = _dev(FisherMatrix)
    C, eps, cond, meth = cov_device(Fd, True, invMethodIn, alt_method, truncate, svals_thresh, condNumbMax)
    out = (C.cpu().numpy(), eps.cpu().numpy())
    if verbose:
        m = meth.cpu().numpy()
        e = out[1]
        print(' Inversion error with method %s: min=%s, max=%s, mean=%s, std=%s ' % (
            invMethodIn, np.nanmin(e), np.nanmax(e), np.nanmean(e), np.nanstd(e)))
        if invMethodIn == 'cho':
            print('Method %s not possible on %s non-positive definite matrices, %s was used in those cases. ' % (
                invMethodIn, int((m == INV_METHODS[alt_method]).sum()), alt_method))
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


# ------------------------------------------------------- variable changes
def _block_transform(M, i0, i1, L, R):
    """A M B per event on the device (gwf_fisher_block_transform); M is
    (P, P) for one event or (P, P, N); L, R are the blocks (2, 2[, N]) of A
    and B on rows / columns (i0, i1), or scalars with i1 = -1."""
    M = np.asarray(M, dtype=np.float64)
    two_d = M.ndim == 2
    Fd = _dev(M[:, :, None] if two_d else M)
    P, n = Fd.shape[0], Fd.shape[2]

    def soa(B):
        B = np.asarray(B, dtype=np.float64)
        if i1 < 0:
            B = np.broadcast_to(B, (n,))
            return np.stack([B, np.zeros(n), np.zeros(n), np.ones(n)])
        return np.stack([np.broadcast_to(B[a][b], (n,)) for a in (0, 1) for b in (0, 1)])

    Ld, Rd = E.to_dev(soa(L)), E.to_dev(soa(R))
    out = E.torch().empty_like(Fd)
    E.call('gwf_fisher_block_transform', N.ptr(Fd), N.ptr(out), P, n, int(i0), int(i1), N.ptr(Ld), N.ptr(Rd),
           N.stream_ptr())
    out = out.cpu().numpy()
    return out[:, :, 0] if two_d else out


def _check_nparams(M, ParNums):
    if len(list(ParNums.keys())) != np.shape(M)[0]:
        raise ValueError('ParNums has %d entries, the matrix %d rows' % (len(list(ParNums.keys())), np.shape(M)[0]))


def log_dL_to_dL_derivative_cov(or_matrix, ParNums, evParams):
    """fisherTools.py:410-432: row and column dL of a (P, P, N) covariance
    in log(dL) multiplied by dL (float64 here, float128 in the reference)."""
    i = ParNums['dL']
    dL = np.asarray(evParams['dL'], dtype=np.float64)
    return _block_transform(or_matrix, i, -1, dL, dL)


def log_dL_to_dL_derivative_fish(or_matrix, ParNums, evParams):
    """fisherTools.py:435-458: row and column dL of a (P, P, N) Fisher in
    log(dL) divided by dL."""
    i = ParNums['dL']
    inv = 1. / np.asarray(evParams['dL'], dtype=np.float64)
    return _block_transform(or_matrix, i, -1, inv, inv)


def dm1_dMc(eta):
    """fisherTools.py:461-470."""
    return (1 + np.sqrt(1 - 4 * eta)) * eta ** (-3. / 5.) / 2


def dm2_dMc(eta):
    """fisherTools.py:473-482."""
    return (1 - np.sqrt(1 - 4 * eta)) * eta ** (-3. / 5.) / 2


def dm1_deta(Mc, eta):
    """fisherTools.py:485-496."""
    return -Mc * (3 - 2 * eta + 3 * np.sqrt(1 - 4 * eta)) / (10 * np.sqrt(1 - 4 * eta) * eta ** (8. / 5.))


def dm2_deta(Mc, eta):
    """fisherTools.py:499-510."""
    return -Mc * (-3 + 2 * eta + 3 * np.sqrt(1 - 4 * eta)) / (10 * np.sqrt(1 - 4 * eta) * eta ** (8. / 5.))


def dMc_dm1(m1, m2):
    """fisherTools.py:513-523."""
    return m2 * (2 * m1 + 3 * m2) / (5 * (m1 * m2) ** (2 / 5) * (m1 + m2) ** (6 / 5))


def dMc_dm2(m1, m2):
    """fisherTools.py:526-536."""
    return dMc_dm1(m2, m1)


def deta_dm1(m1, m2):
    """fisherTools.py:539-549."""
    return m2 * (m2 - m1) / (m1 + m2) ** 3


def deta_dm2(m1, m2):
    """fisherTools.py:552-562."""
    return deta_dm1(m2, m1)


def J_m1m2_Mceta(Mc, eta):
    """fisherTools.py:566-576: d(m1, m2) / d(Mc, eta)."""
    return np.array([[dm1_dMc(eta), dm1_deta(Mc, eta)], [dm2_dMc(eta), dm2_deta(Mc, eta)]])


def J_Mceta_m1m2(m1, m2):
    """fisherTools.py:579-589: d(Mc, eta) / d(m1, m2)."""
    return np.array([[dMc_dm1(m1, m2), dMc_dm2(m1, m2)], [deta_dm1(m1, m2), deta_dm2(m1, m2)]])


def m1m2_from_Mceta(Mc, eta):
    """fisherTools.py:592-603 (no guard on 1 - 4 eta, unlike
    gwfastUtils.m1m2_from_Mceta)."""
    delta = 1 - 4 * eta
    return (1 + np.sqrt(delta)) / 2 * Mc / eta ** (3. / 5.), (1 - np.sqrt(delta)) / 2 * Mc / eta ** (3. / 5.)


def Mceta_from_m1m2(m1, m2):
    """fisherTools.py:606-618, evaluated as written there: with Python's
    operator precedence the first value is (m1 m2)^3 / 5 / (m1 + m2) / 5,
    not the chirp mass (gwfastUtils.Mceta_from_m1m2 gives the chirp mass)."""
    Mc = (m1 * m2) ** 3 / 5 / (m1 + m2) ** 1 / 5
    eta = (m1 * m2) / (m1 + m2) ** 2
    return Mc, eta


def m1m2_to_Mceta_fish(or_matrix, ParNums, evParams):
    """fisherTools.py:620-640: R F R with R the identity except
    J_m1m2_Mceta on the (Mc, eta) block (as in the reference, R is not
    transposed on the left)."""
    _check_nparams(or_matrix, ParNums)
    J = J_m1m2_Mceta(evParams['Mc'], evParams['eta'])
    return _block_transform(or_matrix, ParNums['Mc'], ParNums['eta'], J, J)


def m1m2_to_Mceta_cov(or_matrix, ParNums, evParams):
    """fisherTools.py:643-663: R C R with J_Mceta_m1m2 at the component
    masses of (Mc, eta)."""
    _check_nparams(or_matrix, ParNums)
    J = J_Mceta_m1m2(*m1m2_from_Mceta(evParams['Mc'], evParams['eta']))
    return _block_transform(or_matrix, ParNums['Mc'], ParNums['eta'], J, J)


def Mceta_to_m1m2_fish(or_matrix, ParNums, evParams):
    """fisherTools.py:666-687: R^T F R with R = J_Mceta_m1m2 on the
    (Mc, eta) block: the Fisher matrix in (m1, m2)."""
    m1, m2 = m1m2_from_Mceta(evParams['Mc'], evParams['eta'])
    J = J_Mceta_m1m2(m1, m2)
    return _block_transform(or_matrix, ParNums['Mc'], ParNums['eta'], np.swapaxes(J, 0, 1), J)


def Mceta_to_m1m2_cov(or_matrix, ParNums, evParams):
    """fisherTools.py:690-709: R^T C R with R = J_m1m2_Mceta."""
    J = J_m1m2_Mceta(evParams['Mc'], evParams['eta'])
    return _block_transform(or_matrix, ParNums['Mc'], ParNums['eta'], np.swapaxes(J, 0, 1), J)


def dchi1_dchieff(m1, m2):
    """fisherTools.py:712-722."""
    return 1.


def dchi2_dchieff(m1, m2):
    """fisherTools.py:725-735."""
    return 1.


def dchi1_dDelchi(m1, m2):
    """fisherTools.py:738-748."""
    return m2 / (m1 + m2)


def dchi2_dDelchi(m1, m2):
    """fisherTools.py:751-761."""
    return -m1 / (m1 + m2)


def J_chi1chi2_chieffDeltachi(m1, m2):
    """fisherTools.py:764-774: d(chi1z, chi2z) / d(chi_eff, Delta chi)."""
    one = np.ones_like(np.asarray(m1 + m2, dtype=np.float64))
    return np.array([[one, dchi1_dDelchi(m1, m2)], [one, dchi2_dDelchi(m1, m2)]])


def chi1chi2_to_chieffDeltachi_fish(or_matrix, ParNums, evParams):
    """fisherTools.py:777-797: R^T F R with R = J_chi1chi2_chieffDeltachi
    on the (chi1z, chi2z) block."""
    _check_nparams(or_matrix, ParNums)
    J = J_chi1chi2_chieffDeltachi(*m1m2_from_Mceta(evParams['Mc'], evParams['eta']))
    return _block_transform(or_matrix, ParNums['chi1z'], ParNums['chi2z'], np.swapaxes(J, 0, 1), J)


def chiSchiA_to_chi1chi2_fish(or_matrix, ParNums, evParams):
    """fisherTools.py:800-822: R F R with R = [[1/2, 1/2], [1/2, -1/2]] on
    the (chiS, chiA) block of ParNums."""
    _check_nparams(or_matrix, ParNums)
    J = np.array([[.5, .5], [.5, -0.5]])
    return _block_transform(or_matrix, ParNums['chiS'], ParNums['chiA'], J, J)


def perturb_Fisher(totF, eps=1e-10, **kwargs):
    """fisherTools.py:282-299: invert F and F + eps U[0, 1) (numpy's global
    generator) and print the inf-norm relative change of each covariance."""
    Cov_base, _ = CovMatr(totF, **kwargs)
    totF_random = totF + np.random.rand(*totF.shape) * eps
    Cov, _ = CovMatr(totF_random, **kwargs)
    epsErr = [np.linalg.norm(Cov_base[i] / Cov[i] - 1, ord=np.inf) for i in range(Cov.shape[-1])]
    print('Relative errors when perturbing at the %s level: %s' % (eps, epsErr))


def check_covariance(FisherM, Cov, tol=1e-10):
    """fisherTools.py:302-333: Cov F per event (on the device), printed
    diagnostics; returns the list of products."""
    t = E.torch()
    Fd, Cd = _dev(FisherM), _dev(Cov)
    prod = t.matmul(Cd.permute(2, 0, 1).contiguous(), Fd.permute(2, 0, 1).contiguous()).cpu().numpy()
    recovered_Ids = [prod[i] for i in range(prod.shape[0])]
    epsErr = compute_inversion_error(FisherM, Cov)
    print('Inversion errors: %s' % epsErr)
    diag_diff = [r.diagonal() - 1 for r in recovered_Ids]
    print('diagonal-1 = %s' % str(diag_diff))
    offDiag = [r[~np.eye(r.shape[0], dtype=bool)] for r in recovered_Ids]
    print('Max off diagonal: %s' % str([max(o) for o in offDiag]))
    print('\nmask: where F*S(off-diagonal)>%s (--> problematic if True off diagonal)' % tol)
    print([r > tol for r in recovered_Ids])
    return recovered_Ids
