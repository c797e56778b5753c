"""GPU tests of the learned SNR screen (gwf_mlp.hip through the C ABI).

Kernels are checked against a float64 matmul of the same bf16 values (GEMM),
torch's bf16 cast (rounding) and the CPU oracle oracle/mlp.py (forward,
backward, Adam, features).  The trained screen is audited against the exact
FP64 network SNR of every event, and the screened Fisher pipeline against
the same pipeline without the screen.
"""
import ctypes
import time

import numpy as np
import pytest

from tests._common import NET_C1, hip_net
from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _call(name, *args):
    from gwfast_amd import _native as N
    L = N.lib()
    N.check(getattr(L, name)(*args), L)


def _s():
    from gwfast_amd import _native as N
    return N.stream_ptr()


def _bf(torch, a):
    """bf16 device tensor and its exact float64 values on the host."""
    t = torch.as_tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.cuda(), t.to(torch.float64).numpy()


@pytest.mark.parametrize('M,Nn,K', [(64, 64, 16), (70, 33, 45), (130, 256, 256), (256, 16, 1000)])
@pytest.mark.parametrize('a_kcontig', [True, False])
@pytest.mark.parametrize('b_kcontig', [True, False])
@pytest.mark.parametrize('split', [1, 4])
def test_gemm_bf16_matches_fp64_product(gpu, M, Nn, K, a_kcontig, b_kcontig, split):
    import torch
    rng = np.random.default_rng(M * 7 + Nn * 3 + K)
    A = rng.normal(0, 1, (M, K))
    B = rng.normal(0, 1, (K, Nn))
    # A(m, k) stored row-major (k contiguous) or as its transpose (m contiguous)
    if a_kcontig:
        Ad, Ah = _bf(torch, A)
        sam, sak = K, 1
    else:
        Ad, AhT = _bf(torch, A.T.copy())
        Ah = AhT.T
        sam, sak = 1, M
    # B(k, n) stored as (n, k) row-major (k contiguous) or (k, n) row-major
    if b_kcontig:
        Bd, BhT = _bf(torch, B.T.copy())
        Bh = BhT.T
        sbk, sbn = 1, K
    else:
        Bd, Bh = _bf(torch, B)
        sbk, sbn = Nn, 1
    C0 = rng.normal(0, 1, (M, Nn)).astype(np.float32)
    C = torch.as_tensor(C0).cuda()
    _call('gwf_gemm_bf16', M, Nn, K, 1.5, _p(Ad), sam, sak, _p(Bd), sbk, sbn, 0.5, _p(C), Nn, split, _s())
    ref = 1.5 * (Ah @ Bh) + 0.5 * C0.astype(np.float64)
    bound = 2 * K * 2.0 ** -24 * (1.5 * np.abs(Ah) @ np.abs(Bh) + np.abs(C0)) + 1e-30
    err = np.abs(C.cpu().numpy() - ref)
    assert np.all(err <= bound + 2.0 ** -23 * np.abs(ref)), float(np.max(err / bound))


def test_to_bf16_matches_torch_cast(gpu):
    import torch
    x = torch.randn(100003, dtype=torch.float32, device='cuda') * 3
    y = torch.empty(x.numel(), dtype=torch.bfloat16, device='cuda')
    _call('gwf_to_bf16', _p(x), x.numel(), _p(y), _s())
    assert torch.equal(y.view(torch.int16), x.to(torch.bfloat16).view(torch.int16))


def _random_screen(torch, seed=0, scale=1.0):
    from gwfast_amd.screen import SNRScreen, N_PARAMS
    s = SNRScreen(seed=seed)
    rng = np.random.default_rng(seed + 1)
    p = {}
    for name, shape in OM.LAYOUT:
        fan = shape[1] if len(shape) == 2 else OM.HID
        p[name] = rng.normal(0.0, scale / np.sqrt(fan), shape)
    p['W1'][:, OM.NFEAT:] = 0.0
    theta = OM.flatten(p).astype(np.float32)
    assert theta.size == N_PARAMS
    s.set_params(theta)
    return s, OM.unflatten(theta.astype(np.float64))


def test_forward_backward_match_oracle(gpu):
    """One training step's prediction and every gradient tensor against the
    oracle run with the same bf16 roundings (batch not a multiple of 64)."""
    import torch
    s, p = _random_screen(torch, seed=3, scale=1.5)
    B = 300
    rng = np.random.default_rng(9)
    Xh = np.zeros((B, OM.KFEAT))
    Xh[:, :OM.NFEAT] = rng.normal(0, 1, (B, OM.NFEAT))
    Xd, Xh = _bf(torch, Xh)
    t = rng.normal(0, 1, B)
    td = torch.as_tensor(t.astype(np.float32)).cuda()
    y, bt = s.forward_train(Xd)
    loss = torch.zeros(1, dtype=torch.float32, device='cuda')
    _call('gwf_mse_grad', _p(y), _p(td), B, _p(bt.g), _p(loss), _s())
    s.backward(Xd, bt, bt.g)
    yo, cache = OM.forward(p, Xh, emulate=True)
    lo, go = OM.mse(yo, t)
    grads = OM.backward(p, cache, go, emulate=True)
    yg = y.cpu().numpy()
    assert np.max(np.abs(yg - yo)) <= 2e-3 * max(1.0, np.max(np.abs(yo)))
    assert abs(float(loss.item()) - lo) <= 1e-3 * lo
    g_gpu = OM.unflatten(s.grad.cpu().numpy().astype(np.float64))
    for name, _ in OM.LAYOUT:
        ref = grads[name]
        err = np.linalg.norm(g_gpu[name] - ref) / np.linalg.norm(ref)
        assert err < 5e-3, (name, err)
    assert np.all(g_gpu['W1'][:, OM.NFEAT:] == 0.0)


def test_adam_matches_oracle(gpu):
    import torch
    rng = np.random.default_rng(11)
    n = 5000
    w = rng.normal(0, 1, n).astype(np.float32)
    wd = torch.as_tensor(w).cuda()
    md = torch.zeros(n, dtype=torch.float32, device='cuda')
    vd = torch.zeros(n, dtype=torch.float32, device='cuda')
    wb = torch.empty(n, dtype=torch.bfloat16, device='cuda')
    wo, mo, vo = w.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in range(1, 6):
        g = rng.normal(0, 1, n).astype(np.float32)
        gd = torch.as_tensor(g).cuda()
        _call('gwf_adam', _p(wd), _p(gd), _p(md), _p(vd), n, 1e-2, 0.9, 0.999, 1e-8, step, _p(wb), _s())
        wo, mo, vo = OM.adam(wo, g.astype(np.float64), mo, vo, step, lr=1e-2)
    assert np.max(np.abs(wd.cpu().numpy() - wo)) < 1e-6
    assert torch.equal(wb.view(torch.int16), wd.to(torch.bfloat16).view(torch.int16))


def test_features_match_oracle(gpu):
    import torch
    from gwfast_amd.catalog import draw_bbh
    from gwfast_amd._engine import param_rows, to_dev
    from gwfast_amd.screen import SNRScreen
    r = param_rows(draw_bbh(1000, seed=21), False)
    s = SNRScreen()
    raw = s.raw_features(to_dev(r)).cpu().numpy()
    ref = OM.features(r)
    assert np.max(np.abs(raw - ref)) < 1e-5 * np.max(np.abs(ref))
    s.mean.copy_(torch.as_tensor(ref.mean(0).astype(np.float32)))
    s.istd.copy_(torch.as_tensor((1 / ref.std(0)).astype(np.float32)))
    X = s.features(to_dev(r)).to(torch.float64).cpu().numpy()
    Xo = OM.standardise(raw, ref.mean(0), 1 / ref.std(0))
    assert np.array_equal(X, Xo)


def test_fused_screen_matches_oracle_and_unfused(gpu):
    """gwf_snr_screen (MFMA 16x16x32, activations in LDS) against the
    oracle forward and the unfused training forward on the same weights."""
    import torch
    from gwfast_amd.catalog import draw_bbh
    from gwfast_amd._engine import param_rows, to_dev
    s, p = _random_screen(torch, seed=5, scale=1.5)
    n = 1000 + 37
    r = param_rows(draw_bbh(n, seed=22), False)
    f = OM.features(r)
    mean, istd = f.mean(0), 1 / f.std(0)
    s.mean.copy_(torch.as_tensor(mean.astype(np.float32)))
    s.istd.copy_(torch.as_tensor(istd.astype(np.float32)))
    s.y_shift, s.y_scale = 2.5, 0.75
    rows = to_dev(r)
    ls = s.predict_log_snr(rows).cpu().numpy()
    y_unf, _ = s.forward_train(s.features(rows))
    ls_unf = y_unf.cpu().numpy() * 0.75 + 2.5
    Xo = OM.standardise(s.raw_features(rows).cpu().numpy(), mean.astype(np.float32), istd.astype(np.float32))
    yo, _ = OM.forward(p, Xo, emulate=True)
    ls_o = yo * 0.75 + 2.5
    scale = max(1.0, np.max(np.abs(yo)))
    assert np.max(np.abs(ls - ls_o)) <= 3e-3 * scale, np.max(np.abs(ls - ls_o))
    assert np.max(np.abs(ls - ls_unf)) <= 3e-3 * scale
    cls = s.classify(rows, 8.0, 16.0).cpu().numpy()
    exp = np.where(ls < np.log(np.float32(8.0)), 0, np.where(ls > np.log(np.float32(16.0)), 2, 1))
    assert np.array_equal(cls, exp)


def test_select_and_gather(gpu):
    import torch
    from gwfast_amd.screen import select, gather
    rng = np.random.default_rng(31)
    for n in (0, 1, 255, 256, 100003):
        c = rng.integers(0, 3, n).astype(np.int8)
        cd = torch.as_tensor(c).cuda()
        for want in (0, 1, 2):
            idx = select(cd, want).cpu().numpy()
            assert np.array_equal(idx, np.nonzero(c == want)[0])
    rows = torch.as_tensor(rng.normal(0, 1, (11, 777))).cuda()
    idx = torch.as_tensor(np.sort(rng.choice(777, 100, replace=False))).cuda()
    assert torch.equal(gather(rows, idx), rows[:, idx])


def test_snr_from_fisher_is_exact_snr(gpu):
    """F[dL, dL] = SNR^2 / dL^2: the Fisher step re-checks the threshold for free."""
    from gwfast_amd.waveforms import IMRPhenomD
    from gwfast_amd.screen import snr_from_fisher, training_set
    net = hip_net(IMRPhenomD(), NET_C1, True)
    rows, snr = training_set(net, 200, seed=41)
    F = net.fisher_device(rows, 2000)
    rel = ((snr_from_fisher(F, rows) - snr).abs() / snr).max().item()
    assert rel < 1e-10


# ------------------------------------------------------ trained screen
N_TRAIN, N_VAL, N_TEST = 200_000, 20_000, 100_000


@pytest.fixture(scope='module')
def trained(gpu):
    from gwfast_amd.waveforms import IMRPhenomD
    from gwfast_amd.screen import SNRScreen, training_set
    net = hip_net(IMRPhenomD(), NET_C1, True)
    rows, snr = training_set(net, N_TRAIN, seed=1001)
    vrows, vsnr = training_set(net, N_VAL, seed=1002)
    s = SNRScreen(seed=7)
    hist = s.train(rows, snr, epochs=12, batch=2048, lr=2e-3, seed=3, val=(vrows, vsnr))
    print('\nscreen training: %d events, %d steps, %.1f s, loss %s, val rmse %s' %
          (N_TRAIN, hist['steps'], hist['time_s'], ['%.2e' % x for x in hist['loss'][::3]],
           ['%.4f' % x for x in hist['val_rmse'][::3]]), flush=True)
    return net, s, hist


def test_training_converges(trained):
    net, s, hist = trained
    assert hist['loss'][-1] < 0.2 * hist['loss'][0]
    assert hist['val_rmse'][-1] < 0.1, hist['val_rmse']


def test_screen_audit_zero_false_negatives(trained):
    from gwfast_amd.screen import audit, training_set
    net, s, _ = trained
    rows, snr = training_set(net, N_TEST, seed=1003)
    a = audit(s, rows, snr, threshold=12.0, band=(8.0, 16.0))
    print('\naudit:', a, flush=True)
    assert a['false_negatives'] == 0
    assert a['rmse_log_snr'] < 0.1


def test_screened_fisher_matches_plain_and_gains(trained):
    """Same confirmed events and bit-identical Fisher matrices as the plain
    pipeline (exact SNR of every event, then Fisher above threshold), with
    fewer exact SNR evaluations; both wall times are printed."""
    import torch
    from gwfast_amd.screen import screened_fisher, plain_fisher, training_set
    net, s, _ = trained
    rows, _ = training_set(net, N_TEST, seed=1004)
    for _ in range(2):  # second pass timed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        a = screened_fisher(net, s, rows, 12.0, (8.0, 16.0), 2000)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        b = plain_fisher(net, rows, 12.0, 2000)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    print('\nscreened %.1f ms (%d exact SNR, %d Fisher), plain %.1f ms (%d exact SNR, %d Fisher)' %
          (1e3 * (t1 - t0), a['n_exact_snr'], a['n_fisher'], 1e3 * (t2 - t1), b['n_exact_snr'], b['n_fisher']),
          flush=True)
    assert torch.equal(a['idx'], b['idx'])
    assert torch.equal(a['F'], b['F'])
    assert a['n_exact_snr'] < b['n_exact_snr']
