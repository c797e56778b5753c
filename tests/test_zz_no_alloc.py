"""gwf_fisher_noalloc: the workspace-free variant of gwf_fisher (event
records and queue counters in the library's per-device scratch array),
called through the C ABI and checked against gwf_fisher with a caller
workspace on the same inputs: two calls in a row, the second smaller, so
the second reads the array the first left behind."""
import numpy as np
import pytest

from tests._common import NET_C1, bbh, bns, hip_net

pytestmark = pytest.mark.gpu


def _abi_fisher(net, wf, rows, n, res, arm, noalloc):
    import torch
    from gwfast_amd import _engine as E, _native as N
    from gwfast_amd import signal as S
    L = N.lib()
    sigs = list(net.signals.values())
    P = wf.nParams
    out = torch.full((P, P, n), float('nan'), dtype=torch.float64, device=rows.device)
    if noalloc:
        rc = L.gwf_fisher_noalloc(wf.c_struct(), S._dets(sigs), len(sigs), N.ptr(rows), rows.stride(0), n, res, arm,
                                  N.ptr(out), N.stream_ptr())
    else:
        ws = torch.empty(int(L.gwf_workspace_bytes(n)), dtype=torch.uint8, device=rows.device)
        rc = L.gwf_fisher(wf.c_struct(), S._dets(sigs), len(sigs), N.ptr(rows), rows.stride(0), n, res, arm,
                          N.ptr(out), N.ptr(ws), ws.numel(), N.stream_ptr())
    N.check(rc, L)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize('model', ['PD', 'TF2-tidal'])
def test_fisher_noalloc_equals_fisher(gpu, model):
    from gwfast_amd import _engine as E
    from gwfast_amd.waveforms import IMRPhenomD, TaylorF2_RestrictedPN
    wf = IMRPhenomD() if model == 'PD' else TaylorF2_RestrictedPN(is_tidal=True)
    draw = bbh if model == 'PD' else bns
    net = hip_net(wf, NET_C1, True)
    for n, res, seed, arm in ((3000, 600, 1, -1), (700, 400, 2, -1), (500, 300, 3, 2)):
        ev = draw(n, seed=seed)
        rows = E.to_dev(E.param_rows(ev, wf.is_tidal))
        a = _abi_fisher(net, wf, rows, n, res, arm, True)
        b = _abi_fisher(net, wf, rows, n, res, arm, False)
        print(model, 'n %d res %d arm %d: finite %s, identical %s' % (n, res, arm, np.all(np.isfinite(a)),
                                                                     np.array_equal(a, b)), flush=True)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)


def test_fisher_noalloc_rejects_oversize(gpu):
    import torch
    from gwfast_amd import _engine as E, _native as N
    from gwfast_amd import signal as S
    from gwfast_amd.waveforms import IMRPhenomD
    L = N.lib()
    wf = IMRPhenomD()
    net = hip_net(wf, NET_C1, True)
    sigs = list(net.signals.values())
    n = 8193
    rows = E.to_dev(E.param_rows(bbh(n, seed=4), False))
    out = torch.empty((11, 11, n), dtype=torch.float64, device=rows.device)
    rc = L.gwf_fisher_noalloc(wf.c_struct(), S._dets(sigs), len(sigs), N.ptr(rows), rows.stride(0), n, 300, -1,
                              N.ptr(out), N.stream_ptr())
    assert rc == -22 and b'at most 8192 events' in L.gwf_last_error()
