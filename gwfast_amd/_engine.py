"""Device-side plumbing between the gwfast-compatible classes and the C ABI.

Owns the device copies of PSD and QNM tables, the per-call workspace and the
conversion of gwfast ``evParams`` dictionaries to the FP64 SoA parameter rows
the kernels read (row order = Fisher ParNums order, waveforms.py:78-129).
"""
import ctypes

import numpy as np

from . import _native as N

PARAM_ROWS = ['Mc', 'eta', 'dL', 'theta', 'phi', 'iota', 'psi', 'tcoal', 'Phicoal', 'chi1z', 'chi2z',
              'LambdaTilde', 'deltaLambda']


def torch():
    import torch as _t
    return _t


def device():
    return torch().device('cuda', torch().cuda.current_device())


def to_dev(a, dtype=None):
    """Device copy of a host array: float64 unless dtype is given; integer
    and boolean arrays keep their own dtype."""
    t = torch()
    a = np.ascontiguousarray(a)
    if not a.flags.writeable:      # broadcast views: torch wants memory it may own
        a = a.copy()
    if dtype is None and a.dtype.kind not in 'iub':
        dtype = t.float64
    return t.as_tensor(a, dtype=dtype, device=device())


class Workspace:
    """Grow-only device scratch for the per-event records."""

    def __init__(self):
        self.buf = None

    def get(self, n_ev):
        L = N.lib()
        need = int(L.gwf_workspace_bytes(int(n_ev)))
        if self.buf is None or self.buf.numel() < need:
            self.buf = torch().empty(max(need, 8), dtype=torch().uint8, device=device())
        return self.buf, need


_WS = Workspace()


def workspace(n_ev):
    return _WS.get(n_ev)


def param_rows(evParams, is_tidal, is_eccentric=False):
    """evParams dict -> (P_rows, N) float64 with LambdaTilde/deltaLambda for
    tidal models (signal.py:858-866 computes them from Lambda1/Lambda2) and
    the eccentricity for eccentric ones (signal.py:876-882), in Fisher row
    order."""
    from .gwfastUtils import Lamt_delLam_from_Lam12
    n = len(np.atleast_1d(evParams['Mc']))
    rows = [np.broadcast_to(np.asarray(evParams[k], dtype=np.float64), (n,)) for k in PARAM_ROWS[:11]]
    if is_tidal:
        if 'LambdaTilde' in evParams and 'deltaLambda' in evParams and 'Lambda1' not in evParams:
            Lt, dL = evParams['LambdaTilde'], evParams['deltaLambda']
        else:
            try:
                Lt, dL = Lamt_delLam_from_Lam12(np.asarray(evParams['Lambda1'], dtype=np.float64),
                                                np.asarray(evParams['Lambda2'], dtype=np.float64),
                                                np.asarray(evParams['eta'], dtype=np.float64))
            except KeyError:
                raise ValueError('Two among Lambda1, Lambda2 and LambdaTilde and deltaLambda have to be provided.')
        rows += [np.broadcast_to(np.asarray(Lt, dtype=np.float64), (n,)),
                 np.broadcast_to(np.asarray(dL, dtype=np.float64), (n,))]
    if is_eccentric:
        try:
            rows.append(np.broadcast_to(np.asarray(evParams['ecc'], dtype=np.float64), (n,)))
        except KeyError:
            raise ValueError('Eccentricity has to be provided.')
    return np.ascontiguousarray(np.array(rows, dtype=np.float64))


def n_rows(wf):
    """Parameter rows a model's kernels read: 11, + 2 tidal, + 1 eccentric."""
    return 11 + 2 * int(bool(wf.is_tidal)) + int(bool(getattr(wf, 'is_eccentric', False)))


def check_rows(rows, wf=None, need=None):
    """Rejects parameter rows the kernels would read out of bounds: a 2-D
    float64 tensor with unit stride along events, at least n_rows(wf) (or
    need) rows, on the GPU.  The C ABI takes the row stride but not the row
    count, so this is the check."""
    t = torch()
    if not isinstance(rows, t.Tensor) or rows.dim() != 2:
        raise ValueError('parameter rows must be a 2-D (P_rows, N_ev) tensor')
    if rows.dtype != t.float64:
        raise ValueError('parameter rows must be float64, got %s' % rows.dtype)
    need = n_rows(wf) if need is None else need
    if rows.shape[0] < need:
        raise ValueError('the model needs %d parameter rows, got %d' % (need, rows.shape[0]))
    if rows.shape[1] > 1 and rows.stride(1) != 1:
        raise ValueError('parameter rows must be contiguous along events')
    if rows.device.type != 'cuda':
        raise ValueError('parameter rows must be on the GPU')


def call(fn, *args):
    L = N.lib()
    N.check(getattr(L, fn)(*args), L)
