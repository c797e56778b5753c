"""Oracle restatement of gwfast's covariance inversion and Fisher checks
(test infrastructure only).

Follows gwfast/fisherTools.py:32-196 (CovMatr, default 'cho' method with the
'svd' fallback), 199-210 (compute_inversion_error) and 216-279 (CheckFisher).
Like the reference it works per event in mpmath at the default 53-bit
precision; the output is float64 here (the reference stores float128 where the
platform has it, fisherTools.py:22-27).
"""
import mpmath
import numpy as np


def cov_matr(F, alt_method='svd'):
    """Returns (Cov (P,P,N), eps (N,), method (N,) with 0 = cholesky,
    1 = svd fallback, -1 = NaN input/failure)."""
    P, _, N = F.shape
    cov = np.zeros(F.shape)
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
            inv_method = 'cho'
            c = None
            if np.any(E < 0):
                inv_method = alt_method
            else:
                try:
                    c = (mpmath.cholesky(Fp)) ** -1
                except Exception:
                    inv_method = alt_method
            if inv_method == 'cho':
                cc = c.T * c
                method[k] = 0
            else:
                U, Sm, V = mpmath.svd_r(Fp)
                cc = V.T * mpmath.diag([1 / s for s in Sm]) * U.T
                method[k] = 1
            cc = (cc + cc.T) / 2
            C = ws * cc * ws if reweighted else cc
            cov[:, :, k] = np.array(C.tolist(), dtype=float)
        except Exception:
            cov[:, :, k] = np.nan
            method[k] = -1
    return cov, compute_inversion_error(F, cov), method


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
