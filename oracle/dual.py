"""Forward-mode dual numbers over NumPy arrays (test infrastructure only).

A ``Dual`` carries a value array ``v`` and a tangent array ``d`` of shape
``(K,) + v.shape``: the K directional derivatives of ``v``.  The elementary
rules below are the textbook ones and reproduce what JAX's AD applies to the
same jnp expressions (gwfast/signal.py:1174-1189), including:

* ``sqrt`` at 0 with zero tangent gives NaN (0 * inf), as JAX's sqrt rule does
  for the ``sqrt(where(eta < 0.25, 1 - 4 eta, 0))`` guard at eta == 0.25;
* ``interp`` differentiates as the slope of the bracketing segment found with
  ``searchsorted(side='right')`` clipped to [1, n-1], which is how jnp.interp
  (JAX 0.4.30) locates its segment;
* ``amin`` sends the tangent of the minimal element (the reference's
  ``fRef = amin(fgrid)`` at waveforms.py:1139).

Values may be real or complex; complex arithmetic is holomorphic so
``exp(1j * Psi)`` carries tangents through the strain.
"""
import numpy as np


def _expand(d, vnd, ndim):
    # d has shape (K,) + vshape; insert broadcast axes after K so the trailing
    # axes line up with a value of ndim dimensions.
    if vnd == ndim:
        return d
    return d.reshape((d.shape[0],) + (1,) * (ndim - vnd) + d.shape[1:])


class Dual:
    __array_priority__ = 1000

    def __init__(self, v, d):
        self.v = np.asarray(v)
        d = np.asarray(d)
        if d.shape[1:] != self.v.shape:
            d = np.broadcast_to(d, (d.shape[0],) + self.v.shape).copy()
        self.d = d

    # -- construction -------------------------------------------------------
    @staticmethod
    def seed(values, k, K):
        """Dual whose tangent is the k-th unit direction of K."""
        v = np.asarray(values, dtype=float)
        d = np.zeros((K,) + v.shape, dtype=float)
        d[k] = 1.0
        return Dual(v, d)

    @property
    def K(self):
        return self.d.shape[0]

    @property
    def shape(self):
        return self.v.shape

    @property
    def real(self):
        return Dual(np.real(self.v), np.real(self.d))

    @property
    def imag(self):
        return Dual(np.imag(self.v), np.imag(self.d))

    def __len__(self):
        return len(self.v)

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        return Dual(self.v[idx], self.d[(slice(None),) + idx])

    # -- arithmetic ---------------------------------------------------------
    def _bin(self, other):
        if isinstance(other, Dual):
            nd = np.broadcast_shapes(self.v.shape, other.v.shape)
            return other.v, _expand(self.d, self.v.ndim, len(nd)), _expand(other.d, other.v.ndim, len(nd))
        o = np.asarray(other)
        nd = np.broadcast_shapes(self.v.shape, o.shape)
        return o, _expand(self.d, self.v.ndim, len(nd)), None

    def __add__(self, other):
        ov, sd, od = self._bin(other)
        return Dual(self.v + ov, sd if od is None else sd + od)

    __radd__ = __add__

    def __sub__(self, other):
        ov, sd, od = self._bin(other)
        return Dual(self.v - ov, sd if od is None else sd - od)

    def __rsub__(self, other):
        ov, sd, od = self._bin(other)
        return Dual(ov - self.v, -sd)

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __pos__(self):
        return self

    def __mul__(self, other):
        ov, sd, od = self._bin(other)
        d = sd * ov
        if od is not None:
            d = d + od * self.v
        return Dual(self.v * ov, d)

    __rmul__ = __mul__

    def __truediv__(self, other):
        ov, sd, od = self._bin(other)
        v = self.v / ov
        if od is None:
            return Dual(v, sd / ov)
        return Dual(v, (sd - od * v) / ov)

    def __rtruediv__(self, other):
        ov, sd, od = self._bin(other)
        v = ov / self.v
        return Dual(v, -sd * v / self.v)

    def __pow__(self, c):
        if isinstance(c, Dual):
            raise TypeError('Dual ** Dual is not used on this path')
        c = np.asarray(c)
        v = self.v ** c
        nd = np.broadcast_shapes(self.v.shape, c.shape)
        return Dual(v, _expand(self.d, self.v.ndim, len(nd)) * (c * self.v ** (c - 1.0)))

    def __rpow__(self, base):
        v = np.asarray(base) ** self.v
        return Dual(v, self.d * v * np.log(base))

    # comparisons act on values (used inside where())
    def __lt__(self, o):
        return self.v < (o.v if isinstance(o, Dual) else o)

    def __le__(self, o):
        return self.v <= (o.v if isinstance(o, Dual) else o)

    def __gt__(self, o):
        return self.v > (o.v if isinstance(o, Dual) else o)

    def __ge__(self, o):
        return self.v >= (o.v if isinstance(o, Dual) else o)


def value(x):
    return x.v if isinstance(x, Dual) else x


def _un(x, f, df):
    if isinstance(x, Dual):
        v = f(x.v)
        return Dual(v, x.d * df(x.v, v))
    return f(x)


def sqrt(x):
    return _un(x, np.sqrt, lambda a, v: 0.5 / v)


def log(x):
    return _un(x, np.log, lambda a, v: 1.0 / a)


def log10(x):
    return _un(x, np.log10, lambda a, v: 1.0 / (a * np.log(10.0)))


def exp(x):
    return _un(x, np.exp, lambda a, v: v)


def sin(x):
    return _un(x, np.sin, lambda a, v: np.cos(a))


def cos(x):
    return _un(x, np.cos, lambda a, v: -np.sin(a))


def arctan(x):
    return _un(x, np.arctan, lambda a, v: 1.0 / (1.0 + a * a))


def fabs(x):
    return _un(x, np.abs, lambda a, v: np.sign(a))


def real(x):
    return x.real if isinstance(x, Dual) else np.real(x)


def where(cond, a, b):
    cond = np.asarray(cond)
    if not isinstance(a, Dual) and not isinstance(b, Dual):
        return np.where(cond, a, b)
    av, bv = value(a), value(b)
    v = np.where(cond, av, bv)
    nd = v.ndim
    K = a.K if isinstance(a, Dual) else b.K
    ad = _expand(a.d, a.v.ndim, nd) if isinstance(a, Dual) else np.zeros((K,) + (1,) * nd)
    bd = _expand(b.d, b.v.ndim, nd) if isinstance(b, Dual) else np.zeros((K,) + (1,) * nd)
    d = np.where(cond[None, ...], ad, bd)
    return Dual(v, d)


def interp(x, xp, fp, left=None, right=None):
    """jnp.interp semantics (JAX 0.4.30): segment from searchsorted(xp, x,
    side='right') clipped to [1, len(xp)-1]; constant outside the table."""
    xp = np.asarray(xp, dtype=float)
    fp = np.asarray(fp, dtype=float)
    xv = np.real(value(x))
    i = np.clip(np.searchsorted(xp, xv, side='right'), 1, len(xp) - 1)
    dfp = fp[i] - fp[i - 1]
    dx = xp[i] - xp[i - 1]
    delta = xv - xp[i - 1]
    eps = np.spacing(np.finfo(xp.dtype).eps)
    dx0 = np.abs(dx) <= eps
    slope = np.where(dx0, 0.0, dfp / np.where(dx0, 1.0, dx))
    f = np.where(dx0, fp[i - 1], fp[i - 1] + (delta / np.where(dx0, 1.0, dx)) * dfp)
    lv = fp[0] if left is None else left
    rv = fp[-1] if right is None else right
    outside = (xv < xp[0]) | (xv > xp[-1])
    f = np.where(xv < xp[0], lv, f)
    f = np.where(xv > xp[-1], rv, f)
    slope = np.where(outside, 0.0, slope)
    if isinstance(x, Dual):
        return Dual(f, np.real(x.d) * slope)
    return f


def amin0(x):
    """np.amin(x, axis=0), sending the tangent of the arg-min element."""
    if not isinstance(x, Dual):
        return np.amin(x, axis=0)
    idx = np.argmin(x.v, axis=0)
    cols = np.arange(x.v.shape[1])
    return Dual(x.v[idx, cols], x.d[:, idx, cols])
