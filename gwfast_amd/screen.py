"""Learned SNR screen in front of the Fisher step.

No counterpart in gwfast: the reference evaluates every event's SNR with
``GWSignal.SNRInteg`` (signal.py:658-777, summed over detectors by
``DetNet.SNR``, network.py:53-81) before a catalogue's Fisher matrices are
computed (calculate_forecasts_from_catalog.py applies the SNR threshold to
the Fisher results).  Here an MLP 11 -> 256 -> 256 -> 256 -> 1 (SiLU, bf16
weights and activations, fp32 accumulation) predicts the network log SNR;
events whose prediction falls inside a band around the threshold are
recomputed with the exact FP64 SNR kernel, events above the band go to the
Fisher step directly, and events below it are dropped.

Every network SNR is also available exactly from its Fisher matrix:
dh/d dL = -h/dL, so F[dL, dL] = SNR^2 / dL^2 and SNR = dL sqrt(F[dL, dL]).
Events the screen sends to the Fisher step above the band are therefore
re-checked against the threshold at no extra cost; the only error the
screen can make is a false negative (a true SNR > threshold predicted
below the band), which ``audit`` counts against the exact SNR of every
event.

Training, inference, stream compaction and gathers all run in this
package's HIP kernels (gwf_mlp.hip) through the C ABI; torch provides
device memory, index permutations and the boolean masks of small arrays.
"""
import ctypes
import time

import numpy as np

from . import _native as N
from ._engine import torch, device, to_dev, call, check_rows

NFEAT = 11
KFEAT = 16
HID = 256
# flat fp32 parameter vector: W (out, in) row-major, then biases
LAYOUT = (('W1', (HID, KFEAT)), ('W2', (HID, HID)), ('W3', (HID, HID)), ('b1', (HID,)), ('b2', (HID,)),
          ('b3', (HID,)), ('w4', (HID,)), ('b4', (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)
N_EFFECTIVE_PARAMS = NFEAT * HID + HID + 2 * (HID * HID + HID) + HID + 1   # 134,913 (W1's pad columns are inert)
FLOP_PER_INFERENCE = 2 * (NFEAT * HID + 2 * HID * HID + HID)                # 268,288


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return N.stream_ptr()


def _offsets():
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = (o, n, shape)
        o += n
    return out


OFFSETS = _offsets()


def init_params(seed=0):
    """He-uniform hidden layers (zero pad columns), small output layer, zero biases."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(N_PARAMS, dtype=np.float32)

    def put(name, a):
        o, n, _ = OFFSETS[name]
        theta[o:o + n] = np.asarray(a, dtype=np.float32).ravel()

    W1 = np.zeros((HID, KFEAT))
    W1[:, :NFEAT] = rng.uniform(-1, 1, (HID, NFEAT)) * np.sqrt(6.0 / NFEAT)
    put('W1', W1)
    put('W2', rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6.0 / HID))
    put('W3', rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6.0 / HID))
    put('w4', rng.uniform(-1, 1, HID) * np.sqrt(3.0 / HID))
    return theta


class _Batch:
    """Activation and gradient buffers of one training batch size."""

    def __init__(self, B):
        T = torch()
        dev = device()
        f32, bf = T.float32, T.bfloat16
        self.B = B
        self.C = T.empty((B, HID), dtype=f32, device=dev)
        self.Z = [T.empty((B, HID), dtype=f32, device=dev) for _ in range(3)]
        self.H = [T.empty((B, HID), dtype=bf, device=dev) for _ in range(3)]
        self.dH = T.empty((B, HID), dtype=f32, device=dev)
        self.dZ = T.empty((B, HID), dtype=f32, device=dev)
        self.dZb = T.empty((B, HID), dtype=bf, device=dev)
        self.y = T.empty(B, dtype=f32, device=dev)
        self.g = T.empty(B, dtype=f32, device=dev)


class SNRScreen(object):
    """The MLP, its feature standardisation and its Adam state, on the GPU."""

    def __init__(self, seed=0):
        T = torch()
        N.lib()
        self.theta = to_dev(init_params(seed), T.float32)
        self.theta_bf = T.empty(N_PARAMS, dtype=T.bfloat16, device=device())
        self.grad = T.zeros(N_PARAMS, dtype=T.float32, device=device())
        self.m = T.zeros(N_PARAMS, dtype=T.float32, device=device())
        self.v = T.zeros(N_PARAMS, dtype=T.float32, device=device())
        self.step = 0
        self.mean = T.zeros(NFEAT, dtype=T.float32, device=device())
        self.istd = T.ones(NFEAT, dtype=T.float32, device=device())
        self.y_shift = 0.0
        self.y_scale = 1.0
        self._sync_bf16()
        self._batch = None
        self._loss = T.zeros(1, dtype=T.float32, device=device())

    # ------------------------------------------------------------ views
    def _view(self, flat, name):
        o, n, shape = OFFSETS[name]
        return flat[o:o + n].view(*shape)

    def _sync_bf16(self):
        call('gwf_to_bf16', _ptr(self.theta), N_PARAMS, _ptr(self.theta_bf), _stream())

    def set_params(self, theta):
        """Load a flat fp32 parameter vector (N_PARAMS,) and reset Adam."""
        T = torch()
        self.theta.copy_(T.as_tensor(np.asarray(theta, dtype=np.float32)))
        self.m.zero_()
        self.v.zero_()
        self.step = 0
        self._sync_bf16()

    def params(self):
        return self.theta.cpu().numpy().copy()

    # --------------------------------------------------------- features
    def raw_features(self, rows):
        T = torch()
        rows = rows.contiguous()
        check_rows(rows, need=11)
        n = rows.shape[1]
        raw = T.empty((n, NFEAT), dtype=T.float32, device=device())
        call('gwf_snr_features', _ptr(rows), rows.stride(0), n, None, None, _ptr(raw), None, _stream())
        return raw

    def features(self, rows):
        """(n, 16) bf16 standardised features of device parameter rows."""
        T = torch()
        rows = rows.contiguous()
        check_rows(rows, need=11)
        n = rows.shape[1]
        X = T.empty((n, KFEAT), dtype=T.bfloat16, device=device())
        call('gwf_snr_features', _ptr(rows), rows.stride(0), n, _ptr(self.mean), _ptr(self.istd), None, _ptr(X),
             _stream())
        return X

    def fit_normalisation(self, rows, log_snr):
        """Feature mean / 1/std and target shift / scale from a training set."""
        T = torch()
        raw = self.raw_features(rows).cpu().numpy().astype(np.float64)
        mu = raw.mean(axis=0)
        sd = raw.std(axis=0)
        sd[sd == 0] = 1.0
        self.mean.copy_(T.as_tensor(mu.astype(np.float32)))
        self.istd.copy_(T.as_tensor((1.0 / sd).astype(np.float32)))
        ls = np.asarray(log_snr.cpu().numpy() if hasattr(log_snr, 'cpu') else log_snr, dtype=np.float64)
        self.y_shift = float(ls.mean())
        self.y_scale = float(ls.std()) or 1.0

    # ----------------------------------------------------- forward/back
    def _gemm(self, M, Nn, K, A, sam, sak, B, sbk, sbn, C, ldc, beta=0.0, split=1):
        call('gwf_gemm_bf16', int(M), int(Nn), int(K), 1.0, _ptr(A), int(sam), int(sak), _ptr(B), int(sbk), int(sbn),
             float(beta), _ptr(C), int(ldc), int(split), _stream())

    def forward_train(self, X, bt=None):
        """Standardised prediction y (B,) keeping the activations for backward."""
        B = X.shape[0]
        if bt is None:
            if self._batch is None or self._batch.B != B:
                self._batch = _Batch(B)
            bt = self._batch
        s = _stream()
        inp, k_in = X, KFEAT
        for l in range(3):
            W = self._view(self.theta_bf, 'W%d' % (l + 1))
            self._gemm(B, HID, k_in, inp, k_in, 1, W, 1, k_in, bt.C, HID)
            call('gwf_bias_silu', _ptr(bt.C), _ptr(self._view(self.theta, 'b%d' % (l + 1))), B, HID, _ptr(bt.Z[l]),
                 _ptr(bt.H[l]), s)
            inp, k_in = bt.H[l], HID
        call('gwf_head_forward', _ptr(bt.H[2]), _ptr(self._view(self.theta, 'w4')), _ptr(self._view(self.theta, 'b4')),
             B, _ptr(bt.y), s)
        return bt.y, bt

    def backward(self, X, bt, g):
        """Gradient of sum(g * y) into self.grad (overwritten)."""
        B = X.shape[0]
        s = _stream()
        split = max(1, min(64, B // 256))
        call('gwf_head_backward', _ptr(bt.H[2]), _ptr(self._view(self.theta, 'w4')), _ptr(g), B, _ptr(bt.dH),
             _ptr(self._view(self.grad, 'w4')), _ptr(self._view(self.grad, 'b4')), s)
        for l in (2, 1, 0):
            call('gwf_silu_backward', _ptr(bt.dH), _ptr(bt.Z[l]), B * HID, _ptr(bt.dZ), _ptr(bt.dZb), s)
            call('gwf_colsum', _ptr(bt.dZ), None, None, B, HID, _ptr(self._view(self.grad, 'b%d' % (l + 1))), 0, s)
            inp, k_in = (bt.H[l - 1], HID) if l > 0 else (X, KFEAT)
            # dW (out, in) = dZ^T H_prev
            self._gemm(HID, k_in, B, bt.dZb, 1, HID, inp, k_in, 1, self._view(self.grad, 'W%d' % (l + 1)), k_in,
                       split=split)
            if l > 0:
                # dH_prev (B, in) = dZ W
                W = self._view(self.theta_bf, 'W%d' % (l + 1))
                self._gemm(B, HID, HID, bt.dZb, HID, 1, W, HID, 1, bt.dH, HID)

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        call('gwf_adam', _ptr(self.theta), _ptr(self.grad), _ptr(self.m), _ptr(self.v), N_PARAMS, float(lr),
             float(beta1), float(beta2), float(eps), int(self.step), _ptr(self.theta_bf), _stream())

    # ---------------------------------------------------------- training
    def train(self, rows, snr, epochs=20, batch=4096, lr=2e-3, lr_min=2e-5, seed=0, val=None, verbose=False):
        """Fit log SNR of device parameter rows (P_rows, n) to exact SNRs (n,).

        Mean-squared error on the standardised log SNR, Adam with a cosine
        learning-rate schedule; the normalisation is fitted first.  Returns
        a history dict (loss per epoch, validation RMSE in log SNR if
        ``val=(rows_val, snr_val)`` is given, wall time)."""
        T = torch()
        t0 = time.perf_counter()
        log_snr = snr.to(T.float64).log()
        self.fit_normalisation(rows, log_snr)
        X = self.features(rows)
        t = ((log_snr - self.y_shift) / self.y_scale).to(T.float32)
        n = X.shape[0]
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        gen = T.Generator(device=device())
        gen.manual_seed(int(seed))
        bt = _Batch(batch)
        hist = {'loss': [], 'val_rmse': [], 'steps': total}
        k = 0
        for ep in range(epochs):
            perm = T.randperm(n, device=device(), generator=gen)
            self._loss.zero_()
            for i in range(steps_per_epoch):
                idx = perm[i * batch:(i + 1) * batch]
                Xb = X.index_select(0, idx)
                tb = t.index_select(0, idx)
                y, _ = self.forward_train(Xb, bt)
                call('gwf_mse_grad', _ptr(y), _ptr(tb), batch, _ptr(bt.g), _ptr(self._loss), _stream())
                self.backward(Xb, bt, bt.g)
                lr_k = lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * k / max(1, total - 1)))
                self.adam_step(lr_k)
                k += 1
            hist['loss'].append(float(self._loss.item()) / steps_per_epoch)
            if val is not None:
                hist['val_rmse'].append(self.rmse(*val))
            if verbose:
                print('epoch %d loss %.3e%s' % (ep, hist['loss'][-1],
                                                 (' val rmse %.4f' % hist['val_rmse'][-1]) if val is not None else ''),
                      flush=True)
        T.cuda.synchronize()
        hist['time_s'] = time.perf_counter() - t0
        return hist

    # --------------------------------------------------------- inference
    def _screen_call(self, rows, snr_lo, snr_hi, want_log, want_cls):
        T = torch()
        rows = rows.contiguous()
        check_rows(rows, need=11)
        n = rows.shape[1]
        ls = T.empty(n, dtype=T.float32, device=device()) if want_log else None
        cl = T.empty(n, dtype=T.int8, device=device()) if want_cls else None
        v = lambda name, flat: _ptr(self._view(flat, name))
        call('gwf_snr_screen', _ptr(rows), rows.stride(0), n, _ptr(self.mean), _ptr(self.istd),
             v('W1', self.theta_bf), v('b1', self.theta), v('W2', self.theta_bf), v('b2', self.theta),
             v('W3', self.theta_bf), v('b3', self.theta), v('w4', self.theta), v('b4', self.theta),
             float(self.y_scale), float(self.y_shift), float(snr_lo), float(snr_hi), _ptr(ls), _ptr(cl), _stream())
        return ls, cl

    def predict_log_snr(self, rows):
        """Predicted network log SNR (n,) float32 of device parameter rows."""
        return self._screen_call(rows, 1.0, 1.0, True, False)[0]

    def classify(self, rows, snr_lo=8.0, snr_hi=16.0):
        """int8 (n,): 0 predicted SNR < snr_lo, 1 inside [snr_lo, snr_hi], 2 above."""
        return self._screen_call(rows, snr_lo, snr_hi, False, True)[1]

    def rmse(self, rows, snr):
        T = torch()
        d = self.predict_log_snr(rows).to(T.float64) - snr.to(T.float64).log()
        return float(d.square().mean().sqrt().item())

    # -------------------------------------------------------- persistence
    def state_dict(self):
        return {'theta': self.params(), 'mean': self.mean.cpu().numpy(), 'istd': self.istd.cpu().numpy(),
                'y_shift': np.float64(self.y_shift), 'y_scale': np.float64(self.y_scale)}

    def load_state_dict(self, d):
        T = torch()
        self.set_params(d['theta'])
        self.mean.copy_(T.as_tensor(np.asarray(d['mean'], dtype=np.float32)))
        self.istd.copy_(T.as_tensor(np.asarray(d['istd'], dtype=np.float32)))
        self.y_shift = float(d['y_shift'])
        self.y_scale = float(d['y_scale'])

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        s = cls()
        s.load_state_dict({k: d[k] for k in d.files})
        return s


# ------------------------------------------------------------ pipeline
def select(cls, want):
    """Ascending indices (int64 device tensor) of the entries of cls == want."""
    T = torch()
    n = cls.numel()
    idx = T.empty(max(n, 1), dtype=T.int64, device=device())
    cnt = T.empty(1, dtype=T.int64, device=device())
    L = N.lib()
    wsb = int(L.gwf_select_workspace_bytes(n))
    ws = T.empty(max(wsb, 8), dtype=T.uint8, device=device())
    call('gwf_select', _ptr(cls), n, int(want), _ptr(idx), _ptr(cnt), _ptr(ws), wsb, _stream())
    return idx[:int(cnt.item())]


def gather(rows, idx):
    """Columns idx of device parameter rows, as a new (P_rows, len(idx)) tensor."""
    T = torch()
    rows = rows.contiguous()
    if idx.dtype != T.int64 or idx.device != rows.device:
        raise TypeError('gather needs int64 event indices on the rows\' device')
    m = idx.numel()
    out = T.empty((rows.shape[0], m), dtype=T.float64, device=device())
    call('gwf_gather_events', _ptr(rows), rows.stride(0), rows.shape[0], _ptr(idx), m, _ptr(out), m, _stream())
    return out


def snr_from_fisher(F, rows):
    """Exact network SNR of each Fisher matrix: dL sqrt(F[dL, dL])."""
    return rows[2] * F[2, 2].sqrt()


def screened_fisher(net, screen, rows, threshold=12.0, band=(8.0, 16.0), res=2000):
    """Fisher matrices of the events with network SNR > threshold (the
    catalogue driver's cut, calculate_forecasts_from_catalog.py:474), screened.

    Returns a dict: ``idx`` (ascending event indices), ``F`` (P, P, len(idx)),
    ``snr`` (exact, from F), class counts, the number of exact SNR
    evaluations (``n_exact_snr``) and Fisher evaluations (``n_fisher``)."""
    T = torch()
    cls = screen.classify(rows, band[0], band[1])
    band_idx = select(cls, 1)
    above_idx = select(cls, 2)
    snr_band = net.snr_device(gather(rows, band_idx), res) if band_idx.numel() else band_idx.to(T.float64)
    cand = T.cat([above_idx, band_idx[snr_band > threshold]]).sort().values
    n_above, n_band = int(above_idx.numel()), int(band_idx.numel())
    counts = {'n_below': int(cls.numel()) - n_above - n_band, 'n_band': n_band, 'n_above': n_above,
              'n_exact_snr': n_band, 'n_fisher': int(cand.numel())}
    if cand.numel() == 0:
        P = next(iter(net.signals.values())).wf_model.nParams
        return dict(counts, idx=cand, F=T.empty((P, P, 0), dtype=T.float64, device=device()),
                    snr=T.empty(0, dtype=T.float64, device=device()), n_dropped_after_fisher=0)
    cand_rows = gather(rows, cand)
    F = net.fisher_device(cand_rows, res)
    snr = snr_from_fisher(F, cand_rows)
    keep = snr > threshold
    return dict(counts, idx=cand[keep], F=F[:, :, keep], snr=snr[keep],
                n_dropped_after_fisher=int((~keep).sum().item()))


def plain_fisher(net, rows, threshold=12.0, res=2000):
    """The same result without the screen: exact SNR of every event, then
    the Fisher matrices of those above the threshold."""
    snr = net.snr_device(rows, res)
    idx = (snr > threshold).nonzero().flatten()
    F = net.fisher_device(gather(rows, idx), res)
    return {'idx': idx, 'F': F, 'snr': snr[idx], 'snr_all': snr, 'n_exact_snr': int(snr.numel()),
            'n_fisher': int(idx.numel())}


def audit(screen, rows, snr_exact, threshold=12.0, band=(8.0, 16.0)):
    """False negatives (exact SNR > threshold, predicted below the band),
    false positives sent to Fisher (exact < threshold, predicted above the
    band), and the RMSE of the predicted log SNR."""
    T = torch()
    ls = screen.predict_log_snr(rows).to(T.float64)
    cls = screen.classify(rows, band[0], band[1])
    above = snr_exact > threshold
    fn = int((above & (cls == 0)).sum().item())
    fp = int((~above & (cls == 2)).sum().item())
    d = ls - snr_exact.to(T.float64).log()
    return {'false_negatives': fn, 'false_positives_to_fisher': fp, 'rmse_log_snr': float(d.square().mean().sqrt()),
            'max_abs_err_log_snr': float(d.abs().max()), 'n_true_above': int(above.sum().item()),
            'n_below_band': int((cls == 0).sum().item()), 'n_band': int((cls == 1).sum().item()),
            'n_above_band': int((cls == 2).sum().item())}


def training_set(net, n, seed, res=2000):
    """Parameter rows of a BBH catalogue (independent seed) and their exact
    network SNRs from the plain FP64 SNR kernel."""
    from .catalog import draw_bbh
    from ._engine import param_rows
    ev = draw_bbh(int(n), seed=int(seed))
    rows = to_dev(param_rows(ev, False))
    return rows, net.snr_device(rows, res)
