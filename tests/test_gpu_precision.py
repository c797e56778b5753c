"""Split-precision Fisher integral: FP32 MFMA on hi/lo split operands with
FP64 reduction (gwf_gram_split) against the FP64 integral (gwf_gram) on the
same materialised derivatives, and against the oracle.  Errors are stated
per Fisher element (normalised by sqrt(F_aa F_bb)) and per marginalised
sigma = sqrt(Cov_aa)."""
import numpy as np
import pytest

from tests._common import NET_C1, bbh, hip_net, norm_err, oracle_net

pytestmark = pytest.mark.gpu


def _plain(net, rows, res, mode):
    import torch
    from gwfast_amd import signal as S
    n = rows.shape[1]
    F = torch.zeros((11, 11, n), dtype=torch.float64, device=rows.device)
    J = torch.empty((11, n, res, 2), dtype=torch.float64, device=rows.device)
    for sig in net.signals.values():
        f, Sg = S.grid_device(sig, rows, res)
        for k in range(sig.n_arms):
            S.strain_device_into(sig, rows, f, k, J)
            S.gram_device(J, f, Sg, 11, out=F, accumulate=True, mode=mode)
    return F


def test_split_gram_element_and_sigma_error(gpu):
    import torch
    from gwfast_amd import _engine as E
    from gwfast_amd.waveforms import IMRPhenomD
    from gwfast_amd.fisherTools import CovMatr
    n, res = 256, 2000
    ev = bbh(n, seed=61)
    net = hip_net(IMRPhenomD(), NET_C1, True)
    net._prepare(ev)
    rows = E.to_dev(E.param_rows(ev, False))
    F64 = _plain(net, rows, res, 'fp64').cpu().numpy()
    F32 = _plain(net, rows, res, 'split_fp32').cpu().numpy()
    Ffused = net.fisher_device(rows, res).cpu().numpy()
    e_split = norm_err(F32, F64)
    e_fused = norm_err(Ffused, F64)
    C64, _ = CovMatr(F64)
    C32, _ = CovMatr(F32)
    s64 = np.sqrt(np.abs(np.einsum('iin->in', C64)))
    s32 = np.sqrt(np.abs(np.einsum('iin->in', C32)))
    rel = np.abs(s32 - s64) / s64
    print('\nsplit-FP32 Gram: max normalised element err %.2e (fused FP64 vs plain FP64 %.2e); '
          'sigma rel err median %.2e, 90%% %.2e, max %.2e' %
          (e_split, e_fused, np.median(rel), np.quantile(rel, 0.9), np.max(rel)), flush=True)
    assert e_split <= 1e-6
    assert e_fused <= 1e-12
    # F_{dL,dL} = SNR^2 / dL^2 holds in both
    snr = net.snr_device(rows, res).cpu().numpy()
    d = ev['dL']
    assert np.max(np.abs(np.sqrt(F32[2, 2]) * d / snr - 1)) < 1e-6


def test_split_gram_vs_oracle(gpu):
    from gwfast_amd import _engine as E
    from gwfast_amd.waveforms import IMRPhenomD
    from oracle import detector as OD, waveforms as OW
    n, res = 24, 2000
    ev = bbh(n, seed=62)
    net = hip_net(IMRPhenomD(), NET_C1, True)
    net._prepare(ev)
    rows = E.to_dev(E.param_rows(ev, False))
    F32 = _plain(net, rows, res, 'split_fp32').cpu().numpy()
    Fo = OD.network_fisher(oracle_net(NET_C1, True), OW.IMRPhenomD(), ev, res, mode='faithful')
    assert norm_err(F32, Fo) <= 1e-6
