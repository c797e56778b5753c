"""CPU restatement of the learned SNR screen (TEST INFRASTRUCTURE ONLY).

The screen has no counterpart in gwfast (which evaluates every SNR with
GWSignal.SNRInteg, gwfast/signal.py:658-777); this module is the checker of
the HIP training and inference kernels: an FP64 NumPy MLP
11 (padded to 16) -> 256 -> 256 -> 256 -> 1 with SiLU activations.

``emulate=True`` applies the same bf16 roundings as the kernels (weights of
the three hidden layers, activations H after SiLU, and the dZ operands of
the weight/input-gradient products), so GPU-vs-oracle differences reduce to
fp32 accumulation order; ``emulate=False`` is the exact FP64 network.

Pinned by tests/test_oracle_pins.py: backward vs central finite differences
of forward (FP64, emulate=False), SiLU' vs finite differences, bf16 rounding
vs torch's float32 -> bfloat16 cast (round to nearest even), Adam's first
step = lr * sign(g) (Kingma & Ba 2015, bias-corrected moments), features vs
the pattern-function invariance phi -> phi + 2 pi d, tcoal -> tcoal + d.
"""
import numpy as np

NFEAT = 11
KFEAT = 16
HID = 256

# flat parameter layout (fp32 master weights), shared with gwfast_amd.screen
LAYOUT = (('W1', (HID, KFEAT)), ('W2', (HID, HID)), ('W3', (HID, HID)), ('b1', (HID,)), ('b2', (HID,)),
          ('b3', (HID,)), ('w4', (HID,)), ('b4', (1,)))


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def features(rows):
    """(n, 11) features of parameter rows (Fisher order, >= 11 rows)."""
    r = np.asarray(rows, dtype=np.float64)
    ang = r[4] - 2.0 * np.pi * r[7]
    return np.stack([np.log(r[0]), r[1], np.log(r[2]), np.cos(r[3]), np.cos(ang), np.sin(ang), np.cos(r[5]),
                     np.cos(2.0 * r[6]), np.sin(2.0 * r[6]), r[9], r[10]], axis=1)


def standardise(f, mean, istd, emulate=True):
    X = np.zeros((f.shape[0], KFEAT))
    z = (np.asarray(f, np.float32) - np.asarray(mean, np.float32)) * np.asarray(istd, np.float32)
    X[:, :NFEAT] = bf16(z) if emulate else z
    return X


def unflatten(theta):
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def flatten(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def forward(p, X, emulate=True):
    """y (n,) and the cache for backward.  X (n, 16)."""
    rnd = bf16 if emulate else (lambda a: a)
    H = [np.asarray(X, dtype=np.float64)]
    Z = []
    for l in (1, 2, 3):
        W = rnd(p['W%d' % l])
        z = H[-1] @ W.T + p['b%d' % l]
        Z.append(z)
        H.append(rnd(silu(z)))
    y = H[-1] @ p['w4'] + p['b4'][0]
    return y, (H, Z)


def mse(y, t):
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(p, cache, g, emulate=True):
    """Gradients (dict like p) of sum(g * y)."""
    rnd = bf16 if emulate else (lambda a: a)
    H, Z = cache
    grads = {'w4': H[3].T @ g, 'b4': np.array([g.sum()])}
    dH = np.outer(g, p['w4'])
    for l in (3, 2, 1):
        dZ = dH * dsilu(Z[l - 1])
        dZr = rnd(dZ)
        grads['b%d' % l] = dZ.sum(axis=0)
        grads['W%d' % l] = dZr.T @ H[l - 1]
        if l > 1:
            dH = dZr @ rnd(p['W%d' % l])
    return grads


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
