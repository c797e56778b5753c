"""Pins of the SNR-screen oracle (oracle/mlp.py) on the CPU.

Each check is independent of the restated formulas: finite differences of
the forward pass, torch's own bf16 cast, a property of Adam's first step
and an invariance of the features.
"""
import numpy as np
import pytest

from oracle import mlp as M


def _params(seed=1, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in M.LAYOUT:
        fan = shape[1] if len(shape) == 2 else M.HID
        p[name] = rng.normal(0.0, scale / np.sqrt(fan), shape)
    p['W1'][:, M.NFEAT:] = 0.0
    return p


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip('torch')
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-20, 1000), rng.normal(0, 1e20, 1000),
                        [0.0, -0.0, 1.0, -1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)


def test_silu_derivative_vs_finite_differences():
    z = np.linspace(-8, 8, 401)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.dsilu(z))) < 1e-8


def test_backward_vs_finite_differences():
    """Exact FP64 network: every gradient tensor against central differences
    of the loss along random coordinates."""
    rng = np.random.default_rng(3)
    p = _params(scale=1.5)
    X = np.zeros((37, M.KFEAT))
    X[:, :M.NFEAT] = rng.normal(0, 1, (37, M.NFEAT))
    t = rng.normal(0, 1, 37)

    def loss(q):
        y, _ = M.forward(q, X, emulate=False)
        return M.mse(y, t)[0]

    y, cache = M.forward(p, X, emulate=False)
    _, g = M.mse(y, t)
    grads = M.backward(p, cache, g, emulate=False)
    h = 1e-6
    for name, shape in M.LAYOUT:
        flat = p[name].ravel()
        picks = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        if name == 'W1':  # pad columns multiply zero features: zero gradient
            picks = np.array([r * M.KFEAT + c for r, c in zip(rng.integers(0, M.HID, 6), rng.integers(0, M.NFEAT, 6))])
        for i in picks:
            qp = {k: v.copy() for k, v in p.items()}
            qm = {k: v.copy() for k, v in p.items()}
            qp[name].ravel()[i] += h
            qm[name].ravel()[i] -= h
            fd = (loss(qp) - loss(qm)) / (2 * h)
            an = grads[name].ravel()[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, an)
    assert np.all(grads['W1'][:, M.NFEAT:] == 0.0)


def test_emulated_backward_close_to_exact():
    """bf16 emulation perturbs gradients at the bf16 level, not beyond."""
    rng = np.random.default_rng(4)
    p = _params()
    X = np.zeros((64, M.KFEAT))
    X[:, :M.NFEAT] = M.bf16(rng.normal(0, 1, (64, M.NFEAT)))
    t = rng.normal(0, 1, 64)
    ye, ce = M.forward(p, X, emulate=True)
    yx, cx = M.forward(p, X, emulate=False)
    assert np.max(np.abs(ye - yx)) < 0.05 * np.max(np.abs(yx))
    ge = M.backward(p, ce, M.mse(ye, t)[1], emulate=True)
    gx = M.backward(p, cx, M.mse(yx, t)[1], emulate=False)
    for name, _ in M.LAYOUT:
        err = np.linalg.norm(ge[name] - gx[name]) / np.linalg.norm(gx[name])
        assert err < 0.05, (name, err)


def test_adam_first_step_is_lr_sign_g():
    """With bias correction, Adam's first step moves each weight by lr*sign(g)
    (up to eps)."""
    rng = np.random.default_rng(5)
    w = rng.normal(0, 1, 1000)
    g = rng.normal(0, 1, 1000)
    w1, m, v = M.adam(w, g, np.zeros_like(w), np.zeros_like(w), 1, lr=1e-3, eps=0.0)
    assert np.allclose(w - w1, 1e-3 * np.sign(g), rtol=1e-12, atol=0)
    # second step with the same gradient: again lr*sign(g)
    w2, _, _ = M.adam(w1, g, m, v, 2, lr=1e-3, eps=0.0)
    assert np.allclose(w1 - w2, 1e-3 * np.sign(g), rtol=1e-12, atol=0)


def test_features_invariant_under_sidereal_shift():
    """The pattern functions depend on phi - 2 pi tcoal only: shifting both
    by the same fraction of a day leaves the features unchanged."""
    from gwfast_amd.catalog import draw_bbh
    from gwfast_amd._engine import param_rows
    ev = draw_bbh(50, seed=2)
    r = param_rows(ev, False)
    r2 = r.copy()
    d = 0.137
    r2[4] += 2 * np.pi * d
    r2[7] += d
    f1, f2 = M.features(r), M.features(r2)
    assert np.max(np.abs(f1 - f2)) < 1e-12
    assert f1.shape == (50, 11)
    assert np.allclose(f1[:, 0], np.log(ev['Mc']))
    assert np.allclose(f1[:, 2], np.log(ev['dL']))


def test_flatten_roundtrip_and_layout_size():
    p = _params()
    theta = M.flatten(p)
    assert theta.size == 136193
    q = M.unflatten(theta)
    for name, _ in M.LAYOUT:
        assert np.array_equal(p[name], q[name])
