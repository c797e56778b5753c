"""Differentiation checks on the GPU: the tangent kernel's Jacobian against
central finite differences of the strain kernel itself, the
torch.autograd.Function gradient against finite differences of a scalar
loss, and a 1000-event IMRPhenomD network Fisher against the oracle."""
import numpy as np
import pytest

from tests._common import NET_C1, bbh, hip_net, norm_err, oracle_net

pytestmark = pytest.mark.gpu


def _setup(model='PD', n=4, res=400, seed=40):
    import torch
    from gwfast_amd import _engine as E
    from gwfast_amd import waveforms as W
    wf = W.IMRPhenomD() if model == 'PD' else W.TaylorF2_RestrictedPN()
    net = hip_net(wf, NET_C1, True)
    sig = net.signals['ETS']
    ev = bbh(n, seed)
    # tcoal ~ 1e-6 day: the strain's 2 pi f tcoal phase term (tcoal in
    # seconds, up to 1.6e9 rad at 3 kHz for tcoal ~ 1 day) carries an FP64
    # rounding of ~1e-7 rad, which a central difference with a 1e-3 rad
    # phase step would read as a 1e-4 relative error; the tangent itself is
    # exact at any tcoal (the Fisher, where e^{i Psi} cancels, is checked at
    # generic tcoal below).
    ev['tcoal'] = np.random.default_rng(seed).uniform(0.5e-6, 1.5e-6, n)
    sig._complete(ev)
    rows = E.to_dev(E.param_rows(ev, False))
    from gwfast_amd import signal as S
    f, _ = S.grid_device(sig, rows, res)
    # keep the last point clear of IMRPhenomD's cut at Mf = 0.2, where the
    # amplitude jumps to 0 (waveforms.py:1250) and a finite difference would
    # straddle the jump
    f = f * (1.0 - 1e-5)
    return torch, sig, rows, f


def _steps(rows, J, h):
    """Per (parameter, event) step: relative 1e-7, reduced so the largest
    phase change over the grid stays below 1e-3 rad."""
    Jm = (J[..., 0] ** 2 + J[..., 1] ** 2).sqrt()           # (P, N, Nf)
    hm = (h[..., 0] ** 2 + h[..., 1] ** 2).sqrt().T          # (N, Nf)
    sens = (Jm / hm[None].clamp_min(1e-300)).amax(dim=2)     # (P, N) per unit
    sens[7] *= 86400.0                                       # J's tcoal row is per second
    rel = 1e-7 * rows.abs().clamp_min(1e-2)
    return torch_min(rel, 1e-3 / sens.clamp_min(1e-300))


def torch_min(a, b):
    import torch
    return torch.minimum(a, b)


@pytest.mark.parametrize('model', ['PD', 'TF2'])
def test_tangent_jacobian_vs_central_fd(gpu, model):
    torch, sig, rows, f = _setup(model)
    from gwfast_amd import signal as S
    for arm in (0, 2):
        h, J = S.strain_device(sig, rows, f, arm)
        steps = _steps(rows, J, h)
        P = J.shape[0]
        for p in range(P):
            rp, rm = rows.clone(), rows.clone()
            rp[p] += steps[p]
            rm[p] -= steps[p]
            hp, _ = S.strain_device(sig, rp, f, arm, want_J=False)
            hm, _ = S.strain_device(sig, rm, f, arm, want_J=False)
            fd = (hp - hm) / (2 * steps[p][None, :, None])     # (Nf, N, 2)
            if p == 7:
                fd = fd / 86400.0
            Jp = J[p].permute(1, 0, 2)                          # (Nf, N, 2)
            scale = Jp.abs().amax(dim=(0, 2)).clamp_min(1e-300)
            err = ((fd - Jp).abs().amax(dim=(0, 2)) / scale).max().item()
            assert err <= 1e-5, (model, arm, p, err)


def test_autograd_function_gradient(gpu):
    torch, sig, rows, f = _setup('PD', n=3, res=300, seed=41)
    from gwfast_amd import autodiff as AD
    g = torch.Generator(device='cuda').manual_seed(0)
    R = torch.randn((f.shape[0], rows.shape[1], 2), dtype=torch.float64, device='cuda', generator=g)
    x = rows.clone().requires_grad_(True)
    h = AD.strain(sig, x, f, arm=1)
    # weight by 1/|h| so every frequency contributes O(1)
    wgt = 1.0 / (h.detach() ** 2).sum(-1, keepdim=True).sqrt().clamp_min(1e-300)
    loss = (h * R * wgt).sum()
    loss.backward()
    from gwfast_amd import signal as S
    hh, J = S.strain_device(sig, rows, f, 1)
    steps = _steps(rows, J, hh)
    for p in range(rows.shape[0]):
        for e in range(rows.shape[1]):
            rp, rm = rows.clone(), rows.clone()
            rp[p, e] += steps[p, e]
            rm[p, e] -= steps[p, e]
            lp = (AD.strain(sig, rp, f, arm=1) * R * wgt).sum()
            lm = (AD.strain(sig, rm, f, arm=1) * R * wgt).sum()
            fd = ((lp - lm) / (2 * steps[p, e])).item()
            gr = x.grad[p, e].item()
            assert abs(gr - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(x.grad[p].cpu().numpy()).max()), (p, e, gr, fd)


def test_fisher_1000_events_vs_oracle(gpu):
    """The 1000-event validation subset of the 10^5-event workload: IMRPhenomD,
    ET + 2 CE with Earth rotation, 2000 frequencies."""
    from gwfast_amd import waveforms as W
    from oracle import detector as OD, waveforms as OW
    n, res = 1000, 2000
    ev = bbh(n, seed=42)
    net = hip_net(W.IMRPhenomD(), NET_C1, True)
    F = net.FisherMatr(dict(ev), res=res)
    od = oracle_net(NET_C1, True)
    owf = OW.IMRPhenomD()
    worst_d = worst_f = 0.0
    for s in range(0, n, 125):
        sub = {k: v[s:s + 125] for k, v in ev.items()}
        Fd = OD.network_fisher(od, owf, sub, res, mode='dual')
        Ff = OD.network_fisher(od, owf, sub, res, mode='faithful')
        worst_d = max(worst_d, norm_err(F[:, :, s:s + 125], Fd))
        worst_f = max(worst_f, norm_err(F[:, :, s:s + 125], Ff))
    print('1000-event Fisher: max normalised err vs dual %.3e, vs faithful %.3e' % (worst_d, worst_f), flush=True)
    # measured 1.0e-10: GPU vs NumPy transcendental rounding, amplified by
    # the cancellations inside the PN phase derivatives
    assert worst_d <= 1e-9
    assert worst_f <= 1e-6
