"""Benchmark: network Fisher matrices per second on one MI355X.

Workload (the metric's configuration): IMRPhenomD, 11 parameters, a
5-interferometer network -- the ET triangle at the ETS site (ET-0000A-18 ASD,
null-stream third arm) plus Cosmic Explorer L detectors at the CE1Id and
CE2NM sites (CE 40 km ASD) -- Earth rotation on, fmin 2 Hz, 2000-point
geometric grids to the waveform cut, 10^4 synthetic BBH events (seed 0),
FP64 throughout.

A step is one DetNet.FisherMatr over the whole batch with the parameters
already resident in HBM: the per-event coefficient kernel plus the fused
waveform/projection/tangent/Gram kernel.  Also reported: the roofline of the
dominant (Fisher) kernel, timed with HIP events on the stream it runs on,
and a CPU baseline (the oracle, a NumPy restatement of the reference's
algorithm) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_EV = 10000
RES = 2000
SITES = [('ETS', 'ET'), ('CE1Id', 'CE'), ('CE2NM', 'CE')]
GRAM_FLOP_PER_FISHER = 4 * 11 * 11 * RES * 5   # 4 P^2 N_f N_ifo = 4.84 MFLOP
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per k_fisher_pq launch at 10^4 events, from rocprofv3 --pmc FETCH_SIZE
# (26,213.6 KiB, x2.0: a 256 MiB device copy reads 131,083.5 KiB by that counter)
# and WRITE_SIZE (9,457.4 KiB, x1.0) runs of this script; summary in
# profiles/k_fisher_imrphenomd_net_10k_2000f.md
K_FISHER_TRAFFIC_10K = (26213.6 * 2.0 + 9457.4) * 1024


def build_net():
    from gwfast_amd import gwfastGlobals as glob
    from gwfast_amd.network import DetNet
    from gwfast_amd.signal import GWSignal
    from gwfast_amd.waveforms import IMRPhenomD
    wf = IMRPhenomD()
    psd = {'ET': glob.PSD_ET, 'CE': glob.PSD_CE}
    sigs = {}
    for name, p in SITES:
        d = glob.detectors[name]
        sigs[name] = GWSignal(wf, psd_path=psd[p], detector_shape=d['shape'], det_lat=d['lat'], det_long=d['long'],
                              det_xax=d['xax'], useEarthMotion=True, fmin=2., verbose=False)
    return wf, DetNet(sigs, verbose=False)


def _cpu_chunk(args):
    """Worker of cpu_baseline: the oracle's network Fisher of one chunk of
    events on one core."""
    ev, res = args
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        from gwfast_amd import gwfastGlobals as glob
        from oracle import detector as OD, waveforms as OW
        psd = {'ET': glob.PSD_ET, 'CE': glob.PSD_CE}
        dets = []
        for name, p in SITES:
            d = glob.detectors[name]
            dets.append(OD.Detector(psd[p], d['shape'], d['lat'], d['long'], d['xax'], useEarthMotion=True, fmin=2.))
        OD.network_fisher(dets, OW.IMRPhenomD(), ev, res, mode='faithful')
    return len(ev['Mc'])


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(per_core=512, res=RES, seed=0, max_cores=16):
    """The oracle (reference-faithful construction: AD-equivalent rows through
    exp(i Psi), closed-form rows, trapz Gram) on per_core events per core of
    the same workload, one single-threaded process per core (the reference's
    catalogue driver parallelises over a process pool the same way,
    run/calculate_forecasts_from_catalog.py:152-186), two chunks per core.
    Runs before the GPU is initialised, so the pool forks a clean process."""
    import multiprocessing as mp
    from gwfast_amd import catalog
    avail = len(os.sched_getaffinity(0))
    cores = max(1, min(avail, max_cores))
    n_sample = per_core * cores
    ev = catalog.draw_bbh(n_sample, seed=seed)
    chunk = max(1, min(128, per_core // 2))  # bounds host memory of the broadcast derivative arrays
    jobs = [({k: v[s:s + chunk] for k, v in ev.items()}, res) for s in range(0, n_sample, chunk)]
    t = time.perf_counter()
    with mp.get_context('fork').Pool(cores) as pool:
        done = sum(pool.map(_cpu_chunk, jobs, chunksize=1))
    dt = time.perf_counter() - t
    return {'value': done / dt, 'unit': 'Fisher matr/s', 'cores': cores, 'kind': 'port',
            'cores_available': avail, 'cores_used': cores, 'cpu_model': _cpu_model(),
            'sample': '%d events of the same workload (IMRPhenomD, ETS+CE1Id+CE2NM, Earth rotation, %d freqs): '
                      'the oracle (CPU restatement of gwfast\'s algorithm in NumPy FP64, not gwfast), one process '
                      'per core, %.1f s wall' % (done, res, dt)}


def plain_path(net, rows, n, steps=3, mode='fp64'):
    """The plain FP64 GPU path on the same workload: per detector the grid and
    PSD (gwf_grid), per arm the materialised strain derivatives J (gwf_strain,
    (11, N, N_f) complex128 in HBM) and the trapezoid Gram over J (gwf_gram;
    mode 'split_fp32': gwf_gram_split, FP32 MFMA on hi/lo operands), summed
    over arms.  Returns (Fisher/s, Fisher tensor)."""
    import torch
    from gwfast_amd import signal as S
    sigs = list(net.signals.values())
    F = torch.zeros((11, 11, n), dtype=torch.float64, device=rows.device)
    J = torch.empty((11, n, RES, 2), dtype=torch.float64, device=rows.device)

    def once():
        F.zero_()
        for sig in sigs:
            f, Sg = S.grid_device(sig, rows, RES)
            for k in range(sig.n_arms):
                S.strain_device_into(sig, rows, f, k, J)
                S.gram_device(J, f, Sg, 11, out=F, accumulate=True, mode=mode)
    once()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        once()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    return n / dt, F


def pcie_leg(net, ev, n, steps=5):
    """The same step with the parameters starting in (pinned) host memory and
    the Fisher matrices copied back to it: Fisher/s including both PCIe
    transfers."""
    import torch
    from gwfast_amd import _engine as E
    rows_h = torch.from_numpy(E.param_rows(ev, False)).pin_memory()
    F_h = torch.empty((11, 11, n), dtype=torch.float64).pin_memory()
    rows_d = torch.empty_like(rows_h, device='cuda')

    def once():
        rows_d.copy_(rows_h, non_blocking=True)
        F = net.fisher_device(rows_d, RES)
        F_h.copy_(F, non_blocking=True)
    once()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        once()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    return {'value': n / dt, 'unit': 'Fisher matr/s', 'ms_per_step': 1e3 * dt,
            'bytes_per_step': int(rows_h.numel() * 8 + F_h.numel() * 8),
            'path': 'pinned host rows -> HBM, DetNet.fisher_device, (11, 11, N) -> pinned host'}


def per_arm_leg(net, rows, n, steps=5):
    """DetNet.FisherMatr(return_all=True), the call of the reference's
    catalogue driver: the Fisher matrix of each of the 5 arms of the same
    workload from one gwf_fisher_arms launch (setup included), against one
    single-arm gwf_fisher launch per arm."""
    import torch
    from gwfast_amd import signal as S
    sigs = list(net.signals.values())
    wf = sigs[0].wf_model

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    one = timed(lambda: S.fisher_arms_device(sigs, wf, rows, RES))
    per = timed(lambda: [S.fisher_device(sigs, wf, rows, RES, arm=k) for k in range(5)])
    return {'value': n / one, 'unit': 'per-arm Fisher sets/s (5 matrices each)', 'ms_per_step': 1e3 * one,
            'single_arm_launches_ms': 1e3 * per, 'speedup_vs_single_arm_launches': per / one,
            'path': 'gwf_fisher_arms (k_fisher_pq, matrix-core Gram per arm)'}


def ragged_leg(n=100_000, steps=3, seed=11):
    """Config C3: TaylorF2 with tidal terms (13 parameters, cut at the Kerr
    ISCO of the remnant: 1.3-4.2 kHz for this population), the same network
    with Earth rotation, 10^5 BNS events each on its own grid length drawn
    uniformly in [500, 10^4] (gwf_fisher_ragged, one persistent launch)."""
    import torch
    from gwfast_amd import _engine as E, catalog, gwfastGlobals as glob
    from gwfast_amd.network import DetNet
    from gwfast_amd.signal import GWSignal
    from gwfast_amd.waveforms import TaylorF2_RestrictedPN
    wf = TaylorF2_RestrictedPN(is_tidal=True, which_ISCO='Kerr')
    psd = {'ET': glob.PSD_ET, 'CE': glob.PSD_CE}
    sigs = {}
    for name, p in SITES:
        d = glob.detectors[name]
        sigs[name] = GWSignal(wf, psd_path=psd[p], detector_shape=d['shape'], det_lat=d['lat'], det_long=d['long'],
                              det_xax=d['xax'], useEarthMotion=True, fmin=2., verbose=False)
    net = DetNet(sigs, verbose=False)
    ev = catalog.draw_bns(n, seed=seed)
    res = np.random.default_rng(seed).integers(500, 10_001, n).astype(np.int32)
    rows = E.to_dev(E.param_rows(ev, True))
    F = net.fisher_device(rows, res)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        F = net.fisher_device(rows, res)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / steps
    Fh = F.cpu().numpy()
    fc = np.asarray(wf.fcut(**{k: ev[k] for k in ('Mc', 'eta', 'chi1z', 'chi2z')}))
    return {'value': n / dt, 'unit': 'Fisher matr/s', 'ms_per_step': 1e3 * dt, 'n_events': n, 'n_params': 13,
            'n_freqs': '500..10000 per event (uniform), %d points in total' % int(res.sum()),
            'points_per_s': float(res.sum()) / dt, 'finite': bool(np.all(np.isfinite(Fh))),
            'cutoffs_hz': [float(fc.min()), float(fc.max())],
            'path': 'TaylorF2 tidal (Kerr ISCO cut), ETS+CE1Id+CE2NM, Earth rotation, gwf_fisher_ragged'}


def cov_leg(F, steps=10):
    """CovMatr (default 'cho' with the 'svd' fallback, inversion error and
    condition number) and CheckFisher (eigenvalues, eigenvectors) on the
    step's Fisher matrices, timed with HIP events on the library's stream."""
    import torch
    from gwfast_amd import _native as N
    L = N.lib()
    P, n = F.shape[0], F.shape[2]
    C = torch.empty_like(F)
    eps = torch.empty(n, dtype=torch.float64, device=F.device)
    cond = torch.empty(n, dtype=torch.float64, device=F.device)
    meth = torch.empty(n, dtype=torch.int32, device=F.device)
    ev = torch.empty((n, P), dtype=torch.float64, device=F.device)
    vec = torch.empty((n, P, P), dtype=torch.float64, device=F.device)
    st = N.stream_ptr()

    def cov():
        N.check(L.gwf_cov_invert(N.ptr(F), P, n, N.ptr(C), N.ptr(eps), N.ptr(cond), N.ptr(meth), st), L)

    def check():
        N.check(L.gwf_check_fisher(N.ptr(F), P, n, N.ptr(ev), N.ptr(vec), N.ptr(cond), st), L)

    out = {}
    for name, fn in (('cov_invert', cov), ('check_fisher', check)):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        out[name + '_ms'] = a.elapsed_time(b) / steps
    cov()
    torch.cuda.synchronize()
    m = meth.cpu().numpy()
    e = eps.cpu().numpy()
    out.update({'n_events': n, 'n_params': P, 'events_per_s': n / (out['cov_invert_ms'] * 1e-3),
                'methods': {'cho': int((m == 0).sum()), 'svd': int((m == 1).sum()), 'nan': int((m == -1).sum())},
                'median_inversion_error': float(np.nanmedian(e)),
                'path': 'gwf_cov_invert and gwf_check_fisher: one event per lane with the matrices in registers '
                         '(k_cov_lane, k_check_lane), the events they leave by one wavefront per event '
                         '(k_cov, k_check)'})
    return out


def screen_leg(net, n_cat=1_000_000, n_train=1_000_000, epochs=8, res=RES, threshold=12.0, band=(8.0, 16.0)):
    """The learned SNR screen on a 10^6-event catalogue of the same network:
    train the MLP on exact SNRs of an independent catalogue (seed 1001), then
    screen the catalogue (seed 7) at SNR 12 with the [8, 16] recompute band and
    compute the confirmed events' Fisher matrices; compare with the plain
    pipeline (exact SNR of every event, Fisher of those above 12)."""
    import torch
    from gwfast_amd import screen as SC
    t0 = time.perf_counter()
    rows_t, snr_t = SC.training_set(net, n_train, seed=1001, res=res)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    sc = SC.SNRScreen(seed=7)
    hist = sc.train(rows_t, snr_t, epochs=epochs, batch=4096, lr=2e-3, seed=3)
    del rows_t, snr_t
    from gwfast_amd import catalog, _engine as E
    rows = E.to_dev(E.param_rows(catalog.draw_bbh(n_cat, seed=7), False))
    # warm-up of both pipelines on the catalogue itself, so that neither timed
    # run pays for first-time allocations (records workspace, output tensors)
    SC.screened_fisher(net, sc, rows, threshold, band, res)
    SC.plain_fisher(net, rows, threshold, res)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    a = SC.screened_fisher(net, sc, rows, threshold, band, res)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    b = SC.plain_fisher(net, rows, threshold, res)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    aud = SC.audit(sc, rows, b["snr_all"], threshold, band)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    sc.classify(rows, band[0], band[1])
    ev1.record()
    torch.cuda.synchronize()
    t_mlp = ev0.elapsed_time(ev1) * 1e-3
    same = bool(torch.equal(a['idx'], b['idx']) and torch.equal(a['F'], b['F']))
    cov_confirmed = cov_leg(a['F'], steps=2)
    ts, tp = t2 - t1, t3 - t2
    mlp_tflops = SC.FLOP_PER_INFERENCE * n_cat / t_mlp / 1e12
    return {'n_catalogue': n_cat, 'threshold': threshold, 'band': list(band),
            'train': {'n_events': n_train, 'epochs': epochs, 'steps': hist['steps'], 'time_s': hist['time_s'],
                      'snr_generation_s': t_gen, 'final_loss': hist['loss'][-1]},
            'rmse_log_snr': aud['rmse_log_snr'], 'max_abs_err_log_snr': aud['max_abs_err_log_snr'],
            'false_negatives': aud['false_negatives'], 'false_positives_to_fisher': aud['false_positives_to_fisher'],
            'n_true_above': aud['n_true_above'], 'n_below_band': aud['n_below_band'], 'n_band': aud['n_band'],
            'n_above_band': aud['n_above_band'],
            'screened': {'time_s': ts, 'n_exact_snr': a['n_exact_snr'], 'n_fisher': a['n_fisher'],
                         'confirmed_fisher_per_s': int(a['idx'].numel()) / ts},
            'plain': {'time_s': tp, 'n_exact_snr': b['n_exact_snr'], 'n_fisher': b['n_fisher'],
                      'confirmed_fisher_per_s': int(b['idx'].numel()) / tp},
            'gain_end_to_end': tp / ts, 'fisher_identical_to_plain': same, 'cov_confirmed': cov_confirmed,
            'mlp_inference': {'kernel_s': t_mlp, 'events_per_s': n_cat / t_mlp, 'bf16_tflops': mlp_tflops,
                              'frac_of_bf16_dense_peak': mlp_tflops / 2500.0}}


def catalogue_leg(net, n=100_000, res=1000, seed=21):
    """The catalogue driver (forecast.compute_errs + to_file, the chain of
    run/calculate_forecasts_from_catalog.py:421-630 at its res=1000): network
    SNR, threshold 12, Fisher of the detected events, CheckFisher, CovMatr,
    90 % sky areas and the driver's files, on 10^5 BBH events in one chunk;
    the device chain (parameters from host arrays, results back to host
    arrays) and the file writing (text and .npy, as the driver) timed apart."""
    import tempfile
    import torch
    from gwfast_amd import catalog, forecast as FC
    ev = catalog.draw_bbh(n, seed=seed)
    FC.compute_errs({k: np.asarray(v)[:1000] for k, v in ev.items()}, net, res=res)  # warm-up
    torch.cuda.synchronize()
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        r = FC.compute_errs(ev, net, res=res)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        FC.to_file(r, d, '_0_to_%d' % n)
        t2 = time.perf_counter()
    nd = int(len(r['idxs_det']))
    return {'events_per_s': n / (t2 - t0), 'n_events': n, 'n_detected': nd, 'res': res,
            'device_chain_s': t1 - t0, 'files_s': t2 - t1, 'events_per_s_device_chain': n / (t1 - t0),
            'path': 'forecast.compute_errs (gwf_snr, gwf_fisher on the detected events, gwf_check_fisher, '
                    'gwf_cov_invert, gwf_sigma_sky_area) + forecast.to_file'}


def fisher_steps(wf, net, n, steps, warmup, seed=0):
    """The timed step on n events: gwf_prepare (per-event records) then
    gwf_fisher_prepared (the fused Fisher kernel), parameters resident in
    HBM.  Returns (F, rows, wall seconds for the steps, mean Fisher-kernel
    milliseconds from HIP events on the library's stream)."""
    import torch
    from gwfast_amd import _engine as E, _native as N, catalog
    from gwfast_amd import signal as S
    ev = catalog.draw_bbh(n, seed=seed)
    rows = E.to_dev(E.param_rows(ev, False))
    sigs = list(net.signals.values())
    F = torch.empty((11, 11, n), dtype=torch.float64, device=rows.device)
    ws, nb = E.workspace(n)
    dets = S._dets(sigs)
    wfc = wf.c_struct()
    st = N.stream_ptr()
    L = N.lib()

    def step(t_fisher=None):
        N.check(L.gwf_prepare(wfc, dets, len(sigs), N.ptr(rows), rows.stride(0), n, RES, N.ptr(ws), nb, st), L)
        if t_fisher is not None:
            t_fisher[0].record()
        N.check(L.gwf_fisher_prepared(wfc, dets, len(sigs), n, RES, -1, N.ptr(F), N.ptr(ws), nb, st), L)
        if t_fisher is not None:
            t_fisher[1].record()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    return F, rows, wall, kern_ms


def c2_leg(wf, net, n=100_000, steps=5, warmup=2):
    """Config C2: the same step on 10^5 events."""
    F, _, wall, kern_ms = fisher_steps(wf, net, n, steps, warmup)
    ok = bool(np.all(np.isfinite(F.cpu().numpy())))
    achieved = GRAM_FLOP_PER_FISHER * n / (kern_ms * 1e-3) / 1e12
    return {'value': n * steps / wall, 'unit': 'Fisher matr/s', 'n_events': n, 'steps': steps,
            'ms_per_step': 1e3 * wall / steps, 'kernel_ms': kern_ms, 'achieved_tflops': achieved,
            'frac_of_fp64_peak': achieved / FP64_PEAK_TFLOPS, 'finite': ok,
            'path': 'gwf_prepare + gwf_fisher_prepared, as the main step'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    # the GPU's clocks ramp over the first ~20 ms of sustained load (about a
    # dozen steps here): 30 warm-up steps let the 50 timed ones see the
    # steady state a catalogue run works in
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=30)
    ap.add_argument('--n-events', type=int, default=N_EV)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-per-core', type=int, default=512, help='CPU baseline events per core')
    ap.add_argument('--no-c2', action='store_true', help='skip the 10^5-event (C2) leg')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-screen', action='store_true', help='skip the learned-screen (10^6 catalogue) leg')
    ap.add_argument('--no-extra', action='store_true',
                    help='skip the PCIe-inclusive, per-arm, ragged (C3) and catalogue-driver legs')
    ap.add_argument('--screen-only', action='store_true', help='run only the learned-screen leg')
    ap.add_argument('--screen-events', type=int, default=1_000_000)
    ap.add_argument('--screen-train', type=int, default=1_000_000)
    ap.add_argument('--screen-epochs', type=int, default=8)
    ap.add_argument('--copy-calibration-mib', type=int, default=0,
                    help='first run a device copy of this many MiB (a known-size kernel for PMC byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: there is no multi-GPU path (events would shard across GPUs '
              'with no exchange during compute). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    # the CPU baseline first: its process pool forks before the GPU is touched
    cpu = None if (args.no_cpu_baseline or args.screen_only) else cpu_baseline(args.cpu_per_core)

    import torch
    from gwfast_amd import _engine as E, _native as N, catalog
    from gwfast_amd import signal as S

    torch.cuda.set_device(0)
    if args.screen_only:
        _, net = build_net()
        print(json.dumps({'screen': screen_leg(net, n_cat=args.screen_events, n_train=args.screen_train,
                                                epochs=args.screen_epochs)}), flush=True)
        return
    if args.copy_calibration_mib:
        src = torch.ones(args.copy_calibration_mib << 18, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    wf, net = build_net()
    n = args.n_events
    F, rows, wall, kern_ms = fisher_steps(wf, net, n, args.steps, args.warmup)
    ev = catalog.draw_bbh(n, seed=0)
    ms_per_step = 1e3 * wall / args.steps
    value = n * args.steps / wall
    achieved = GRAM_FLOP_PER_FISHER * n / (kern_ms * 1e-3) / 1e12
    Fh = F.cpu().numpy()
    assert np.all(np.isfinite(Fh))
    out = {
        'metric': 'Fisher matr/s, IMRPhenomD 3-det 11-par, rel err <=1e-6 vs FP64, vs plain',
        'value': value, 'unit': 'Fisher matr/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64', 'data': 'synthetic (seeded BBH catalogue)',
        'config': {'workload': 'IMRPhenomD network Fisher: ETS triangle + CE1Id + CE2NM (5 IFOs), Earth rotation, '
                               '11 params, %d events x %d freqs, FP64, fused path (gwf_prepare + gwf_fisher_prepared)' % (n, RES),
                   'n_events': n, 'n_freqs': RES, 'n_params': 11, 'n_ifo': 5, 'parallelism': 'single GPU'},
        'roofline': {'bound': 'mfma', 'achieved': achieved, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / FP64_PEAK_TFLOPS,
                     'traffic': K_FISHER_TRAFFIC_10K if n == N_EV else None,
                     'kernel': 'k_fisher_pq<IMRPhenomD,4,11>', 'kernel_ms': kern_ms,
                     'note': 'FP64 dense peak; algorithmic flops = Gram 4 P^2 N_f N_ifo per Fisher'},
    }
    if not args.no_c2 and n != 100_000:
        out['c2_1e5'] = c2_leg(wf, net)
    out['cov'] = cov_leg(F)
    out['cov']['share_of_fisher_step'] = out['cov']['cov_invert_ms'] / ms_per_step
    if not args.no_plain:
        pv, Fp = plain_path(net, rows, n)
        Fp = Fp.cpu().numpy()
        d = np.sqrt(np.abs(np.einsum('iin->in', Fp)))
        err = float(np.max(np.abs(Fh - Fp) / (d[:, None, :] * d[None, :, :])))
        out['plain_fp64'] = {'value': pv, 'unit': 'Fisher matr/s', 'speedup_fused_vs_plain': value / pv,
                             'max_norm_err_fused_vs_plain': err,
                             'path': 'gwf_grid + per-arm gwf_strain (J in HBM) + gwf_gram, FP64'}
        sv, Fs = plain_path(net, rows, n, mode='split_fp32')
        Fs = Fs.cpu().numpy()
        out['plain_split_fp32_gram'] = {
            'value': sv, 'unit': 'Fisher matr/s', 'speedup_vs_plain_fp64': sv / pv,
            'max_norm_err_vs_plain_fp64': float(np.max(np.abs(Fs - Fp) / (d[:, None, :] * d[None, :, :]))),
            'path': 'as plain_fp64 with gwf_gram_split (FP32 MFMA 16x16x4, hi/lo operands, FP64 reduction)'}
        del Fs
    if not args.no_extra:
        out['pcie_inclusive'] = pcie_leg(net, ev, n)
        out['per_arm_return_all'] = per_arm_leg(net, rows, n)
        out['ragged_c3'] = ragged_leg()
        out['catalogue_1e5'] = catalogue_leg(net)
    if not args.no_screen:
        del F
        out['screen_1e6'] = screen_leg(net, n_cat=args.screen_events, n_train=args.screen_train,
                                       epochs=args.screen_epochs)
    if cpu is not None:
        out['cpu_baseline'] = cpu
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
