"""Oracle restatement of gwfast's covariance inversion and Fisher checks
(test infrastructure only).

Follows gwfast/fisherTools.py:32-196 (CovMatr, every inversion method and
the alternative method), 199-210 (compute_inversion_error), 216-279
(CheckFisher) and 340-858 (the helpers after them).
Like the reference it works per event in mpmath at the default 53-bit
precision; the output is float64 here (the reference stores float128 where the
platform has it, fisherTools.py:22-27).
"""
import mpmath
import numpy as np


METHOD_IDS = {'cho': 0, 'svd': 1, 'inv': 2, 'svd_reg': 3, 'lu': 4}


def invert_scaled(Fp, method, cond=None, truncate=False, svals_thresh=1e-15, cond_max=1e50, c=None):
    """The inverse of one scaled matrix Fp by method (fisherTools.py:113-161),
    before symmetrisation: returns (cc, truncated matrix or None)."""
    if method == 'cho':
        return c.T * c, None
    if method == 'inv':
        return Fp ** -1, None
    if method == 'svd':
        U, Sm, V = mpmath.svd_r(Fp)
        S = np.array([float(x) for x in Sm])
        if truncate and abs(cond) > cond_max:
            maxev = np.max(np.abs(S))
            Sinv = [1 / s if abs(s) / maxev > svals_thresh else 1 / (maxev * svals_thresh) for s in S]
            St = [s if abs(s) / maxev > svals_thresh else maxev * svals_thresh for s in S]
            tF = U * mpmath.diag(St) * V
            return V.T * mpmath.diag(Sinv) * U.T, (tF + tF.T) / 2
        return V.T * mpmath.diag([1 / s for s in Sm]) * U.T, None
    if method == 'svd_reg':
        U, Sm, V = mpmath.svd_r(Fp)
        S = np.array([float(x) for x in Sm])
        Um, Vm = np.array(U.tolist(), dtype=float), np.array(V.tolist(), dtype=float)
        kv = int(np.sum(S > svals_thresh))
        return mpmath.matrix(Um[:, :kv] @ np.diag(1. / S[:kv]) @ Vm[:kv, :]), None
    if method == 'lu':
        Pm, L, U = mpmath.lu(Fp)
        return U ** -1 * (Pm * L) ** -1, None
    raise ValueError(method)


def cov_matr(F, inv_method='cho', alt_method='svd', truncate=False, svals_thresh=1e-15, cond_max=1e50):
    """Returns (Cov (P,P,N), eps (N,), method (N,) with the METHOD_IDS id used,
    -1 = NaN input/failure).  eps is taken against the truncated scaled
    matrix where 'svd' truncates (fisherTools.py:128-136, 189)."""
    P, _, N = F.shape
    cov = np.zeros(F.shape)
    Fref = np.array(F, dtype=float, copy=True)
    method = np.zeros(N, dtype=int)
    for k in range(N):
        Fk = F[:, :, k]
        if np.all(np.isnan(Fk)):
            cov[:, :, k] = np.nan
            method[k] = -1
            continue
        ff = mpmath.matrix(Fk.tolist())
        try:
            E, _ = mpmath.eigh(ff)
            E = np.array([float(e) for e in E])
            reweighted = False
            try:
                ws = mpmath.diag([1 / mpmath.sqrt(ff[i, i]) for i in range(P)])
                Fp = ws * ff * ws
                EE, _ = mpmath.eigh(Fp)
                E = np.array([float(e) for e in EE])
                reweighted = True
            except ZeroDivisionError:
                Fp = ff
            cond = np.max(np.abs(E)) / np.min(np.abs(E))
            m = inv_method
            c = None
            if np.any(E < 0):
                if inv_method == 'cho':
                    m = alt_method
            elif m == 'cho':
                try:
                    c = (mpmath.cholesky(Fp)) ** -1
                except Exception:
                    m = alt_method
            cc, tF = invert_scaled(Fp, m, cond, truncate, svals_thresh, cond_max, c)
            if tF is not None:
                Fref[:, :, k] = np.array(tF.tolist(), dtype=float)
            method[k] = METHOD_IDS[m]
            cc = (cc + cc.T) / 2
            C = ws * cc * ws if reweighted else cc
            cov[:, :, k] = np.array(C.tolist(), dtype=float)
        except Exception:
            cov[:, :, k] = np.nan
            method[k] = -1
    return cov, compute_inversion_error(Fref, cov), method


def compute_inversion_error(F, C):
    """fisherTools.py:199-210: max |C F - I| per event."""
    P = F.shape[0]
    return np.array([np.max(np.abs(C[:, :, i] @ F[:, :, i] - np.eye(P))) for i in range(F.shape[-1])])


def check_fisher(F):
    """fisherTools.py:216-279 with use_mpmath=True: eigenvalues (N,P),
    eigenvectors (N,P,P), condition number max|l|/min|l| (N,)."""
    P, _, N = F.shape
    evals = np.zeros((N, P))
    evecs = np.zeros((N, P, P))
    for k in range(N):
        if np.all(np.isnan(F[:, :, k])):
            evals[k] = np.nan
            evecs[k] = np.nan
            continue
        E, ER = mpmath.eigh(mpmath.matrix(F[:, :, k].tolist()))
        evals[k] = np.array([float(e) for e in E])
        evecs[k] = np.array(ER.tolist(), dtype=float)
    cond = np.abs(evals).max(axis=1) / np.abs(evals).min(axis=1)
    return evals, evecs, cond


def fix_params(F, par_nums, par_marg):
    """fisherTools.py:340-375: delete the rows/columns of par_marg, renumber."""
    idx = np.sort(np.array([par_nums[p] for p in par_marg], dtype=int))
    out = np.delete(np.delete(np.asarray(F), idx, axis=0), idx, axis=1)
    new = dict(par_nums)
    for pm in par_marg:
        for k in new.keys():
            if new[k] > new[pm]:
                new[k] -= 1
        new.pop(pm, None)
    return out, new


def add_prior(F, vals, par_nums, par_add):
    """fisherTools.py:378-403."""
    d = np.zeros(F.shape[0])
    d[np.sort(np.array([par_nums[p] for p in par_add]))] = vals
    pp = np.diag(d)
    return pp + F if F.ndim == 2 else pp[:, :, None] + F


def localization_region(C, par_nums, theta, perc_level=90, units='SqDeg'):
    """fisherTools.py:830-858 (Barack & Cutler, PRD 69, 082005)."""
    tt = C[par_nums['theta'], par_nums['theta']]
    pp = C[par_nums['phi'], par_nums['phi']]
    tp = C[par_nums['phi'], par_nums['theta']]
    base = 2 * np.pi * np.sqrt(tt * pp - tp ** 2) * np.abs(np.sin(theta))
    sr = -base * np.log(1 - perc_level / 100)
    return sr if units == 'Sterad' else (180 / np.pi) ** 2 * sr


# ------------------------------------------------ matrix variable changes
def jac_m1m2_mceta(Mc, eta):
    """fisherTools.py:461-510, 566-576: d(m1, m2)/d(Mc, eta), (2, 2[, N])."""
    s = np.sqrt(1 - 4 * eta)
    return np.array([[(1 + s) * eta ** (-0.6) / 2, -Mc * (3 - 2 * eta + 3 * s) / (10 * s * eta ** 1.6)],
                     [(1 - s) * eta ** (-0.6) / 2, -Mc * (-3 + 2 * eta + 3 * s) / (10 * s * eta ** 1.6)]])


def jac_mceta_m1m2(m1, m2):
    """fisherTools.py:513-562, 579-589: d(Mc, eta)/d(m1, m2), (2, 2[, N])."""
    M = m1 + m2
    dmc1 = m2 * (2 * m1 + 3 * m2) / (5 * (m1 * m2) ** 0.4 * M ** 1.2)
    dmc2 = m1 * (2 * m2 + 3 * m1) / (5 * (m1 * m2) ** 0.4 * M ** 1.2)
    return np.array([[dmc1, dmc2], [m2 * (m2 - m1) / M ** 3, m1 * (m1 - m2) / M ** 3]])


def m1m2_from_mceta_unguarded(Mc, eta):
    """fisherTools.py:592-603."""
    d = np.sqrt(1 - 4 * eta)
    return (1 + d) / 2 * Mc / eta ** 0.6, (1 - d) / 2 * Mc / eta ** 0.6


def jac_chi1chi2_chieff_dchi(m1, m2):
    """fisherTools.py:712-774: d(chi1z, chi2z)/d(chi_eff, Delta chi)."""
    one = np.ones_like(np.asarray(m1 + m2, dtype=float))
    return np.array([[one, m2 / (m1 + m2)], [one, -m1 / (m1 + m2)]])


def sandwich(M, i0, i1, left, right):
    """A M B for each event, A and B the identity with the 2x2 blocks left /
    right (2, 2[, N]) on rows and columns (i0, i1): the matrix products of
    fisherTools.py:620-822, written as explicit per-event matrices."""
    M = np.asarray(M, dtype=float)
    two_d = M.ndim == 2
    M3 = M[:, :, None] if two_d else M
    P, _, n = M3.shape
    out = np.empty_like(M3)
    for e in range(n):
        A, B = np.eye(P), np.eye(P)
        for (a, ia) in ((0, i0), (1, i1)):
            for (b, ib) in ((0, i0), (1, i1)):
                A[ia, ib] = np.broadcast_to(left[a][b], (n,))[e]
                B[ia, ib] = np.broadcast_to(right[a][b], (n,))[e]
        out[:, :, e] = A @ M3[:, :, e] @ B
    return out[:, :, 0] if two_d else out


def scale_row_col(M, i, s):
    """fisherTools.py:410-458: row and column i times s (per event)."""
    out = np.array(M, dtype=float, copy=True)
    out[:, i] *= s
    out[i, :] *= s
    return out
