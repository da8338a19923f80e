"""Benchmark of the fused Welch path on one MI355X.

Workload (headline): welch over x = 64 channels x 2^20 float32 samples, each
channel the sum of 3 unit sinusoids (freqs uniform in [0, fs/2), random phase)
plus N(0,1) noise, fs = 48000, nperseg = 1024, noverlap = 512, hann,
detrend='constant', density scaling, mean average -> 513 bins per channel.
Inputs are generated on the device from a seed (synthetic) and are resident
in HBM before the timed region; three input buffers are rotated so the 256 MB
last-level cache cannot serve the 256 MiB input.

One step = one ``mlx_signal_amd.welch`` call on a device buffer.  Every timed
region follows at least W warm-up steps and at least 80 ms of back-to-back
work, so it measures the sustained rate (an idle GPU runs its first ~30-40 ms
of work at lower clocks).  Prints ONE
JSON line (value = samples/s over the whole job).  Secondary forms (csd,
coherence at the same shape, stft / spectrogram 16 x 2^20) are reported under
"secondary".  The CPU baseline is the reference's dispatch='scipy' route
(scipy.signal.welch on the same float32 data, one process) timed on a bounded
sample on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-secondary]
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "mlx-signal_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

HBM_PEAK_GBS = 8000.0
FS = 48000.0
C, NS, NPERSEG, NOVERLAP = 64, 1 << 20, 1024, 512
METRIC = "Welch samples/s + HBM GB/s (64ch×2^20, nperseg=1024); serving p50/p99 latency"


def welch_bytes(B, n, bins):
    """Algorithmic bytes: input read once + output written once."""
    return 4 * B * n + 4 * B * bins


def make_inputs(torch, rows, n, seed, dev, with_y=True):
    """x = 3 sinusoids + N(0,1); y = FIR32(x) + 0.1 N(0,1) (causal, taps N(0,1)/sqrt(32))."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    t = torch.arange(n, device=dev, dtype=torch.float64)
    x = torch.randn(rows, n, device=dev, generator=g, dtype=torch.float32)
    f = torch.rand(rows, 3, device=dev, generator=g, dtype=torch.float64) * (FS / 2)
    ph = torch.rand(rows, 3, device=dev, generator=g, dtype=torch.float64) * (2 * np.pi)
    for k in range(3):
        x += torch.sin(2 * np.pi * f[:, k : k + 1] / FS * t[None, :] + ph[:, k : k + 1]).float()
    if not with_y:
        return x, None
    taps = torch.randn(32, device=dev, generator=g, dtype=torch.float32) / np.sqrt(32.0)
    xp = torch.nn.functional.pad(x[:, None, :], (31, 0))
    y = torch.nn.functional.conv1d(xp, taps.flip(0)[None, None, :])[:, 0, :]
    y = y + 0.1 * torch.randn(rows, n, device=dev, generator=g, dtype=torch.float32)
    return x.contiguous(), y.contiguous()


def max_err_rows(got, ref):
    """Largest per-bin error of each row relative to that row's FP64 peak
    (the parity metric), maxed over the rows checked."""
    got = np.asarray(got).reshape(len(ref), -1).astype(np.complex128)
    ref = np.asarray(ref).reshape(len(ref), -1).astype(np.complex128)
    return float(max(np.max(np.abs(g - r)) / np.max(np.abs(r)) for g, r in zip(got, ref)))


WARM_MS = 80.0  # back-to-back work before any timed region


def warm_up(torch, fn, bufs, warmup, min_ms=WARM_MS):
    """At least `warmup` steps and at least `min_ms` of synchronized back-to-back
    work.  An idle MI355X runs its first ~30-40 ms of sustained work at lower
    clocks (the fused welch kernel takes ~105 us cold, 86 us sustained), so a
    short warm-up times the ramp rather than the sustained rate.  Returns the
    number of warm-up steps run."""
    n = 0
    for i in range(warmup):
        fn(bufs[n % len(bufs)])
        n += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < min_ms:
        for _ in range(8):
            fn(bufs[n % len(bufs)])
            n += 1
        torch.cuda.synchronize()
    return n


def time_steps(torch, fn, bufs, steps, warmup, kernel_timing=False):
    """Time K back-to-back steps after warm_up(), synchronized on both sides:
    (wall ms per step, HIP-event ms per step over the whole region on the
    current stream, which is the stream the kernels launch on, None, fused
    kernel ms).  Nothing else is recorded inside that region.  With
    kernel_timing, a second region of the same K steps follows at once, in
    which the library hands a HIP event pair to each step's fused transform
    launch (hipExtLaunchKernel stamps the pair with that dispatch's start and
    end); the mean of those K durations is the fused kernel ms."""
    from mlx_signal_amd import _kernels

    warm_up(torch, fn, bufs, warmup)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for i in range(steps):
        fn(bufs[i % len(bufs)])
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = None
    if kernel_timing:
        _kernels.kernel_timing_arm(steps)
        for i in range(steps):
            fn(bufs[i % len(bufs)])
        torch.cuda.synchronize()
        tot, n = _kernels.kernel_timing_read()
        assert n == steps, f"timed {n} fused kernels over {steps} steps"
        kernel_ms = tot / n
    return wall / steps * 1e3, e0.elapsed_time(e1) / steps, None, kernel_ms


TRAFFIC_PROFILE = "profiles/welch_64x2p20_traffic.json"
VALU_PROFILE = "profiles/welch_64x2p20_valu_counters.json"


def compute_bound(kernel_ms):
    """The headline kernel against the bound that binds it, VALU issue.  The
    committed counter pass gives the VALU wave-instructions per launch (a
    property of the build, the same on every box) and the shader clock of the
    box it ran on; a CU issues at most one per cycle.  The shader clock this
    load sustains differs between boxes (1.97-2.15 GHz measured), and a run
    cannot read it without counters, so the line reports the clock at which
    this run's kernel time equals 100% issue (issue_clock_ghz) beside the
    counter pass's own clock and floor.  None if the summary is missing."""
    p = ROOT / VALU_PROFILE
    try:
        c = json.loads(p.read_text())
    except (ValueError, OSError):
        return None
    valu_per_cu = c["sq_insts_valu_per_launch"] / c["cus"]
    floor_us = valu_per_cu / (c["shader_clock_ghz"] * 1e3)
    kernel_us = kernel_ms * 1e3
    return {"bound": "valu-issue", "valu_per_pair": c["valu_per_pair"],
            "valu_per_cu": valu_per_cu, "kernel_us": kernel_us,
            "issue_clock_ghz": valu_per_cu / (kernel_us * 1e3),
            "counter_pass": {"shader_clock_ghz": c["shader_clock_ghz"],
                             "valu_issue_per_cu_cycle": c["valu_issue_per_cu_cycle"],
                             "floor_us": floor_us, "dispatch_us": c["dispatch_us"]},
            "source": f"{VALU_PROFILE}: committed rocprofv3 counter pass of this kernel and "
                      "workload (instruction count, and the clock of the box it ran on); "
                      "kernel_us is this run's"}


def traffic_from_profile():
    """HBM bytes per launch of the dominant kernel, read from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated on copies of
    known size), or None.  It is a committed measurement of the same kernel
    and workload, not a counter read in this run."""
    p = ROOT / TRAFFIC_PROFILE
    if p.exists():
        try:
            return json.loads(p.read_text()).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            return None
    return None


def cpu_baseline(min_seconds=10.0, max_reps=30):
    """scipy.signal.welch (the reference's dispatch='scipy' route, restated in
    oracle.scipy_route) on the headline workload itself, 64 x 2^20 float32,
    one process, repeated until about min_seconds of CPU work; median per call."""
    import oracle

    rng = np.random.default_rng(0)
    xs = rng.standard_normal((C, NS)).astype(np.float32)
    kw = dict(fs=FS, window="hann", nperseg=NPERSEG, noverlap=NOVERLAP)
    oracle.scipy_route("welch", xs[:1], **kw)
    ts = []
    t_start = time.perf_counter()
    while len(ts) < max_reps and (time.perf_counter() - t_start < min_seconds or len(ts) < 3):
        t0 = time.perf_counter()
        oracle.scipy_route("welch", xs, **kw)
        ts.append(time.perf_counter() - t0)
    t = statistics.median(ts)
    return {
        "value": C * NS / t,
        "unit": "samples/s",
        "cores": 1,
        "kind": "port",
        "sample": f"scipy.signal.welch float32 {C}ch x 2^20, nperseg=1024, noverlap=512 (the "
                  f"headline workload; reference dispatch='scipy' route), {len(ts)} calls in "
                  f"{sum(ts):.1f} s, median {t * 1e3:.1f} ms per call; host "
                  f"{os.cpu_count()} logical CPUs, 1 used",
    }


_CPU_X = None  # the CPU sample, inherited by forked workers (no pickling of the data)


def _cpu_welch_rows(bounds):
    import scipy.signal

    a, b = bounds
    return scipy.signal.welch(_CPU_X[a:b], fs=FS, window="hann", nperseg=NPERSEG,
                              noverlap=NOVERLAP)[1]


def cpu_all_cores():
    """The same scipy route on the headline workload sharded by channel over
    the host cores (one forked process per core, up to 16), median of 3.
    Runs before this process initialises the GPU, so the forked workers
    inherit no HIP state; the pool is shut down (workers joined) on exit."""
    import concurrent.futures
    import multiprocessing

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))  # the host share of one GPU
    global _CPU_X
    rows = C
    rng = np.random.default_rng(0)
    _CPU_X = rng.standard_normal((rows, NS)).astype(np.float32)
    edges = np.linspace(0, rows, cores + 1).astype(int)
    chunks = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    ctx = multiprocessing.get_context("fork")
    ex = concurrent.futures.ProcessPoolExecutor(len(chunks), mp_context=ctx)
    try:
        list(ex.map(_cpu_welch_rows, chunks))
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            list(ex.map(_cpu_welch_rows, chunks))
            ts.append(time.perf_counter() - t0)
    finally:
        ex.shutdown(wait=True, cancel_futures=True)
    _CPU_X = None
    t = statistics.median(ts)
    return {"value": rows * NS / t, "unit": "samples/s", "cores": cores,
            "sample": f"scipy.signal.welch float32 {rows}ch x 2^20 (the headline workload) "
                      f"sharded by channel over {cores} processes, median of 3: {t * 1e3:.1f} ms"}


def bench_precision(torch, msig, dev, steps=20):
    """Precision forms: welch over 256 channels x 2^16 (N(0,1) plus one tone
    per channel), nperseg 256, noverlap 128, hann: the FP32 kernels vs the
    matrix-core FP16 / BF16 hi+lo forms, with max and RMS per-bin error
    relative to each channel's FP64 peak (8 channels against the oracle)."""
    import oracle
    from mlx_signal_amd import _kernels

    rows, n = 256, 1 << 16
    kw = dict(fs=FS, window="hann", nperseg=256, noverlap=128)
    bufs = []
    for seed in range(3):
        g = torch.Generator(device=dev)
        g.manual_seed(100 + seed)
        t = torch.arange(n, device=dev, dtype=torch.float32)
        f0 = torch.rand(rows, 1, device=dev, generator=g) * 0.5
        bufs.append(torch.randn(rows, n, device=dev, generator=g) + torch.sin(2 * np.pi * f0 * t))
    ref = oracle.welch(bufs[0][:8].cpu().numpy(), **kw)[1]
    peak = np.abs(ref).max(axis=-1, keepdims=True)
    nbytes = welch_bytes(rows, n, 129)
    out = {"workload": "welch 256ch x 2^16 f32, nperseg=256, noverlap=128, hann, density, mean",
           "algorithmic_bytes": nbytes}
    for precision in ("fp32", "fp16_split", "bf16_split"):
        with msig.config_context(precision=precision):
            def step(x):
                return msig.welch(x, **kw)

            w, e, _, k = time_steps(torch, step, bufs, steps, 3, kernel_timing=True)
            p = step(bufs[0])[1][:8].cpu().numpy()
        rel = np.abs(p - ref) / peak
        out[precision] = {"ms_event": e, "ms_wall": w, "ms_kernel": k,
                          "samples_per_s": rows * n / (e * 1e-3),
                          "kernel_GBps": nbytes / (k * 1e-3) / 1e9,
                          "kernel_roofline_frac": nbytes / (k * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "max_rel_err_vs_fp64_peak": float(rel.max()),
                          "rms_rel_err_vs_fp64_peak": float(np.sqrt(np.mean(rel ** 2)))}
    out["mfma_launches"] = {k: v for k, v in _kernels.launches.items() if "split" in k}
    del bufs
    torch.cuda.empty_cache()
    return out


def bench_small(torch, msig, dev, steps=50):
    """Small welch: 1 x 2^16 float32 N(0,1), nperseg 256,
    noverlap 128 (launch-latency bound), beside scipy on one host core."""
    import oracle

    rng = np.random.default_rng(0)
    xh = rng.standard_normal(1 << 16).astype(np.float32)
    kw = dict(fs=FS, window="hann", nperseg=256, noverlap=128)
    x = torch.from_numpy(xh).to(dev)
    w, e, _, _ = time_steps(torch, lambda b: msig.welch(b, **kw), [x], steps, 5)
    p = msig.welch(x, **kw)[1].cpu().numpy()
    ref = oracle.welch(xh, **kw)[1]
    oracle.scipy_route("welch", xh, **kw)
    ts = []
    for _ in range(20):
        t0 = time.perf_counter()
        oracle.scipy_route("welch", xh, **kw)
        ts.append(time.perf_counter() - t0)
    return {"workload": "welch 1ch x 2^16 f32, nperseg=256, noverlap=128, hann",
            "ms_event": e, "ms_wall": w, "samples_per_s_event": (1 << 16) / (e * 1e-3),
            "max_err_vs_fp64_peak": float(np.abs(p - ref).max() / np.abs(ref).max()),
            "cpu_scipy_ms": statistics.median(ts) * 1e3}


def ragged_requests(torch, dev, n_inst=512, seed=7):
    """512 independent welch/csd instances: channels U{1..64}, lengths
    log-uniform in [2^12, 2^20], nperseg in {256, 1024, 4096}, noverlap =
    nperseg/2, hann or blackman, float32 N(0,1) + one tone, from a seed."""
    from mlx_signal_amd.batch import SpectralRequest

    rng = np.random.default_rng(seed)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    reqs = []
    for _ in range(n_inst):
        ch = int(rng.integers(1, 65))
        n = int(2 ** rng.uniform(12, 20))
        nps = int(rng.choice([256, 1024, 4096]))
        win = str(rng.choice(["hann", "blackman"]))
        kind = str(rng.choice(["welch", "csd"]))
        f0 = float(rng.uniform(0, 0.5))
        t = torch.arange(n, device=dev, dtype=torch.float32)
        x = torch.randn(ch, n, device=dev, generator=g) + torch.sin(2 * np.pi * f0 * t)
        y = (0.5 * x + torch.randn(ch, n, device=dev, generator=g)) if kind == "csd" else None
        reqs.append(SpectralRequest(kind, x, y, fs=FS, window=win, nperseg=nps,
                                    noverlap=nps // 2))
    return reqs


def bench_ragged(torch, msig, dev, reps=3):
    """Ragged batch of 512 instances: one plain call per request vs one ragged batch
    (autotuned, and replayed from a HIP graph), and the same batch with one
    launch per (nfft, kind) class instead of the single persistent launch."""
    import oracle
    from mlx_signal_amd.batch import SpectralBatch
    from mlx_signal_amd.tuning import cached_tuning

    reqs = ragged_requests(torch, dev)
    torch.cuda.synchronize()

    def plain():
        for r in reqs:
            kw = dict(fs=r.fs, window=r.window, nperseg=r.nperseg, noverlap=r.noverlap)
            if r.kind == "welch":
                msig.welch(r.x, **kw)
            else:
                msig.csd(r.x, r.y, **kw)

    def timed(fn):
        warm_up(torch, lambda _: fn(), [None], 1)
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return statistics.median(ts)

    ms_plain = timed(plain)
    tuning = cached_tuning((256, 1024, 4096))
    b = SpectralBatch(reqs, tuning=tuning)
    ms_batch = timed(b.run)
    g = b.graph()
    ms_graph = timed(g.replay)
    # the same requests with one persistent launch per (nfft, kind) class
    tuning_pc = cached_tuning((256, 1024, 4096), per_class_launches=True)
    bp = SpectralBatch(reqs, tuning=tuning_pc, per_class_launches=True)
    gp = bp.graph()
    ms_graph_per_class = timed(gp.replay)
    launches_per_class = bp.launches
    # accuracy of both plans: 6 of the smallest requests against the FP64 oracle
    small = sorted(range(len(reqs)), key=lambda i: reqs[i].x.numel())[:6]
    err = err_pc = 0.0
    for i in small:
        r = reqs[i]
        kw = dict(fs=r.fs, window=r.window, nperseg=r.nperseg, noverlap=r.noverlap)
        xh = r.x.cpu().numpy()
        ref = (oracle.welch(xh, **kw)[1] if r.kind == "welch"
               else oracle.csd(xh, r.y.cpu().numpy(), **kw)[1])
        peak = np.max(np.abs(ref))
        err = max(err, float(np.max(np.abs(b.outputs[i].cpu().numpy() - ref)) / peak))
        err_pc = max(err_pc, float(np.max(np.abs(bp.outputs[i].cpu().numpy() - ref)) / peak))
    bp.close()
    del gp, bp
    out = {
        "instances": len(reqs), "samples": b.samples, "algorithmic_bytes": b.bytes,
        "launches_per_batch": b.launches, "tuning": {str(k): v for k, v in tuning.items()},
        "ms_one_call_per_request": ms_plain, "ms_batch": ms_batch, "ms_graph": ms_graph,
        "speedup_batch": ms_plain / ms_batch, "speedup_graph": ms_plain / ms_graph,
        "ms_graph_per_class_launches": ms_graph_per_class,
        "tuning_per_class": {str(k): v for k, v in tuning_pc.items()},
        "launches_per_class_batch": launches_per_class,
        "max_err_vs_fp64_peak_sampled_per_class": err_pc,
        "samples_per_s_graph": b.samples / (ms_graph * 1e-3),
        "GBps_graph": b.bytes / (ms_graph * 1e-3) / 1e9,
        "max_err_vs_fp64_peak_sampled": err,
    }
    b.close()
    del reqs, b, g
    torch.cuda.empty_cache()
    return out


def serving_schedule(torch, dev, n_req=2000, rate=2000.0, seed=11):
    """Poisson(rate) arrivals of welch/csd requests: channels U{1..256}, lengths
    log-uniform in [2^14, 2^22], nperseg in {256, 512, 1024, 4096}; payloads are
    slices of a device-resident pool of seeded N(0,1) float32 samples."""
    from mlx_signal_amd.batch import SpectralRequest

    rng = np.random.default_rng(seed)
    pool_n = 1 << 31  # 8 GiB
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pool = torch.randn(pool_n, device=dev, generator=g)
    t = 0.0
    sched = []
    for _ in range(n_req):
        t += float(rng.exponential(1.0 / rate))
        ch = int(rng.integers(1, 257))
        n = int(2 ** rng.uniform(14, 22))
        nps = int(rng.choice([256, 512, 1024, 4096]))
        kind = str(rng.choice(["welch", "csd"]))
        size = ch * n
        ox = int(rng.integers(0, pool_n - size))
        x = pool[ox : ox + size].view(ch, n)
        y = None
        if kind == "csd":
            oy = int(rng.integers(0, pool_n - size))
            y = pool[oy : oy + size].view(ch, n)
        sched.append((t, SpectralRequest(kind, x, y, fs=FS, nperseg=nps)))
    return pool, sched


def _plain_serve(torch, msig, sched):
    """One plain call per request, in arrival order, on one stream."""
    t0 = time.perf_counter()
    pending = list(sched)
    inflight = []
    done = []
    i = 0
    while i < len(pending) or inflight:
        now = time.perf_counter() - t0
        while i < len(pending) and pending[i][0] <= now:
            ta, r = pending[i]
            kw = dict(fs=r.fs, window=r.window, nperseg=r.nperseg)
            if r.kind == "welch":
                msig.welch(r.x, **kw)
            else:
                msig.csd(r.x, r.y, **kw)
            ev = torch.cuda.Event()
            ev.record()
            inflight.append((ta, ev))
            i += 1
        still = []
        for ta, ev in inflight:
            if ev.query():
                done.append((ta, time.perf_counter() - t0))
            else:
                still.append((ta, ev))
        inflight = still
    return done


def _stats(done):
    lat = sorted(d - a for a, d in done)
    span = max(d for _, d in done) - min(a for a, _ in done)
    pick = lambda p: lat[min(len(lat) - 1, int(round(p / 100 * (len(lat) - 1))))]
    return {"requests": len(lat), "p50_ms": pick(50) * 1e3, "p99_ms": pick(99) * 1e3,
            "throughput_req_s": len(lat) / span}


def bench_serving(torch, msig, dev):
    """Serving: Poisson 2000 req/s; continuous batching vs one plain
    call per request, plus saturation throughput (all requests at t = 0)."""
    from mlx_signal_amd.serving import SpectralServer, latency_stats
    from mlx_signal_amd.tuning import cached_tuning

    pool, sched = serving_schedule(torch, dev)
    torch.cuda.synchronize()
    tuning = cached_tuning((256, 512, 1024, 4096))
    nbytes = sum(4 * r.x.numel() * (2 if r.y is not None else 1) for _, r in sched)
    nsamp = sum(r.x.numel() for _, r in sched)
    # warm both paths (code objects, cached windows / grids)
    warm = [(0.0, r) for _, r in sched[:8]]
    SpectralServer(tuning=tuning).serve(warm)
    _plain_serve(torch, msig, warm)
    torch.cuda.synchronize()
    out = {"requests": len(sched), "offered_req_s": 2000.0, "samples": nsamp,
           "input_bytes": nbytes, "offered_GBps": nbytes / sched[-1][0] / 1e9}
    srv = SpectralServer(hbm_budget=96 << 30, streams=3, max_inflight=3, tuning=tuning)
    recs = srv.serve(sched)
    out["batched"] = latency_stats(recs)
    out["batched"]["batches"] = srv.batches
    # accuracy: the 6 smallest served results against the FP64 oracle
    import oracle

    torch.cuda.synchronize()
    small = sorted(range(len(sched)), key=lambda i: sched[i][1].x.numel())[:6]
    err = 0.0
    for i in small:
        r = sched[i][1]
        kw = dict(fs=r.fs, nperseg=r.nperseg)
        xh = r.x.cpu().numpy()
        ref = (oracle.welch(xh, **kw)[1] if r.kind == "welch"
               else oracle.csd(xh, r.y.cpu().numpy(), **kw)[1])
        err = max(err, max_err_rows(recs[i].result.cpu().numpy(), ref))
    out["batched"]["max_err_vs_fp64_peak_sampled"] = err
    del recs
    torch.cuda.synchronize()
    out["plain"] = _stats(_plain_serve(torch, msig, sched))
    torch.cuda.synchronize()
    burst = [(0.0, r) for _, r in sched]
    t0 = time.perf_counter()
    SpectralServer(hbm_budget=96 << 30, streams=3, max_inflight=3, tuning=tuning).serve(burst)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    _plain_serve(torch, msig, burst)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    out["saturation"] = {"batched_s": tb, "plain_s": tp, "speedup": tp / tb,
                         "batched_GBps": nbytes / tb / 1e9, "plain_GBps": nbytes / tp / 1e9,
                         "batched_samples_per_s": nsamp / tb}
    del pool, sched, burst
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run 3 device copies of a 256 MiB buffer (a known-size kernel "
                         "that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--precision-only", action="store_true",
                    help="run only the 256ch x 2^16 precision comparison (FP32 vs matrix-core "
                         "FP16/BF16 split) and print its JSON; for profiling that kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (the path "
              "would shard by channel row with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)

    # CPU legs first, before this process touches the GPU: the all-core leg
    # forks workers, which must not inherit HIP runtime state
    cpu = cpu_all = None
    if not args.no_cpu and not args.precision_only:
        cpu = cpu_baseline()
        if not args.no_secondary:
            cpu_all = cpu_all_cores()

    import torch

    import mlx_signal_amd as msig
    import oracle
    from mlx_signal_amd import _hip

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    _hip.load()
    msig.set_config(dispatch="gpu")
    if args.precision_only:
        print(json.dumps(bench_precision(torch, msig, dev, steps=args.steps)), flush=True)
        return

    bins = NPERSEG // 2 + 1
    xs = [make_inputs(torch, C, NS, seed, dev, with_y=False)[0] for seed in range(3)]
    torch.cuda.synchronize()

    def step(x):
        return msig.welch(x, fs=FS, window="hann", nperseg=NPERSEG, noverlap=NOVERLAP)

    ms_wall, ms_ev, ev_all, ms_kernel = time_steps(torch, step, xs, args.steps, args.warmup,
                                                   kernel_timing=True)
    B = welch_bytes(C, NS, bins)
    # roofline of the dominant kernel (the fused transform + segment-sum
    # kernel; the finalize kernel reads only the block partials)
    achieved = B / (ms_kernel * 1e-3) / 1e9

    # accuracy gate on the measured data: 3 channels vs the FP64 oracle
    _, p = step(xs[0])
    hrows = [0, C // 2 - 1, C - 1]
    xr = xs[0][hrows].cpu().numpy()
    ref = oracle.welch(xr, fs=FS, window="hann", nperseg=NPERSEG, noverlap=NOVERLAP)[1]
    err = max_err_rows(p[hrows].cpu().numpy(), ref)

    if args.copy_ref:
        src = xs[0]
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del dst

    secondary = {}
    if not args.no_secondary:
        x2, y2 = make_inputs(torch, C, NS, 11, dev)
        x3, y3 = make_inputs(torch, C, NS, 12, dev)
        pairs = [(x2, y2), (x3, y3)]
        kw = dict(fs=FS, window="hann", nperseg=NPERSEG, noverlap=NOVERLAP)
        rows = [0, C // 2 - 1, C - 1]
        xh, yh = x2[rows].cpu().numpy(), y2[rows].cpu().numpy()
        for name, fn, nbytes, ref in [
            ("csd", lambda b: msig.csd(b[0], b[1], **kw), 8 * C * NS + 8 * C * bins,
             lambda: oracle.csd(xh, yh, **kw)[1]),
            ("coherence", lambda b: msig.coherence(b[0], b[1], **kw), 8 * C * NS + 4 * C * bins,
             lambda: oracle.coherence(xh, yh, **kw)[1]),
        ]:
            w, e, _, _ = time_steps(torch, fn, pairs, max(5, args.steps // 2), 2)
            got = fn(pairs[0])[1][rows].cpu().numpy()
            secondary[name] = {"ms_event": e, "ms_wall": w, "GBps": nbytes / (e * 1e-3) / 1e9,
                               "pairs_per_s": C * NS / (e * 1e-3),
                               "max_err_vs_fp64_peak": max_err_rows(got, ref()),
                               "err_rows": rows}
        del x2, y2, x3, y3, pairs
        xs16 = [x[:16] for x in xs]
        nseg_stft = (NS + 1024 - 1024) // 512 + 1
        nseg_sp = (NS - 1024) // 512 + 1
        srows = [0, 15]
        xs16h = xs16[0][srows].cpu().numpy()
        for name, fn, nbytes, ref in [
            ("stft_16x2^20", lambda b: msig.stft(b, nperseg=1024),
             4 * 16 * NS + 8 * 16 * nseg_stft * bins,
             lambda: oracle.stft(xs16h, nperseg=1024)[2]),
            ("spectrogram_16x2^20", lambda b: msig.spectrogram(b, nperseg=1024, noverlap=512),
             4 * 16 * NS + 4 * 16 * nseg_sp * bins,
             lambda: oracle.spectrogram(xs16h, nperseg=1024, noverlap=512)[2]),
        ]:
            w, e, _, _ = time_steps(torch, fn, xs16, max(5, args.steps // 2), 2)
            got = fn(xs16[0])[2][srows].cpu().numpy()
            secondary[name] = {"ms_event": e, "ms_wall": w, "GBps": nbytes / (e * 1e-3) / 1e9,
                               "samples_per_s": 16 * NS / (e * 1e-3),
                               "max_err_vs_fp64_peak": max_err_rows(got, ref()),
                               "err_rows": srows}
        # istft of the 16 x 2^20 stft (hop = N/2: inverse and overlap-add in one kernel)
        zs = [msig.stft(b, nperseg=1024)[2] for b in xs16]
        nseg_i = zs[0].shape[-1]
        w, e, _, _ = time_steps(torch, lambda z: msig.istft(z, nperseg=1024), zs,
                                max(5, args.steps // 2), 2)
        xr = msig.istft(zs[0], nperseg=1024)[1]
        rt = float((xr[:, :NS] - xs16[0]).abs().max() / xs16[0].abs().max())
        secondary["istft_16x2^20"] = {
            "ms_event": e, "ms_wall": w,
            "GBps": (8 * 16 * nseg_i * bins + 4 * 16 * NS) / (e * 1e-3) / 1e9,
            "samples_per_s": 16 * NS / (e * 1e-3), "round_trip_max_err_vs_peak": rt}
        del zs, xr
        # welch median average (fused power frames + the per-bin median kernel)
        w, e, _, _ = time_steps(
            torch, lambda b: msig.welch(b, fs=FS, nperseg=NPERSEG, noverlap=NOVERLAP,
                                        average="median"), xs, max(5, args.steps // 2), 2)
        pm = msig.welch(xs[0][:2], fs=FS, nperseg=NPERSEG, noverlap=NOVERLAP, average="median")[1]
        refm = oracle.welch(xs[0][:2].cpu().numpy(), fs=FS, nperseg=NPERSEG, noverlap=NOVERLAP,
                            average="median")[1]
        secondary["welch_median_64x2^20"] = {
            "ms_event": e, "ms_wall": w, "samples_per_s": C * NS / (e * 1e-3),
            "max_err_vs_fp64_peak": float(np.abs(pm.cpu().numpy() - refm).max() / np.abs(refm).max())}
        # end-to-end: NumPy in -> tensor -> host out (the reference's "e2e" column)
        xh = xs[0].cpu().numpy()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            _, pe = msig.welch(xh, fs=FS, nperseg=NPERSEG, noverlap=NOVERLAP)
            pe.cpu()
            ts.append(time.perf_counter() - t0)
        secondary["welch_e2e_numpy_ms"] = statistics.median(ts) * 1e3
        del xs16
        secondary["small_1x2^16"] = bench_small(torch, msig, dev)
        secondary["precision_256x2^16"] = bench_precision(torch, msig, dev)
        secondary["ragged_512"] = bench_ragged(torch, msig, dev)
        secondary["serving_2000rps"] = bench_serving(torch, msig, dev)

    out = {
        "metric": METRIC,
        "value": C * NS / (ms_ev * 1e-3),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_min_ms": WARM_MS,
        "ms_per_step": ms_wall,
        "ms_per_step_event": ms_ev,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated: 3 sinusoids + N(0,1) noise, seeded)",
        "config": {"workload": "welch 64ch x 2^20 f32, fs=48000, nperseg=1024, noverlap=512, "
                               "hann, detrend=constant, density, mean -> 513 bins",
                   "channels": C, "samples_per_channel": NS, "nperseg": NPERSEG,
                   "noverlap": NOVERLAP, "parallelism": "single GPU"},
        "ms_dominant_kernel": ms_kernel,
        "dominant_kernel": "k_mean<1024,16,PSD,real,half-hop> (fused window+FFT+|X|^2+segment sum)",
        "hbm_GBps": achieved,
        "step_GBps": B / (ms_ev * 1e-3) / 1e9,
        "max_err_vs_fp64_peak": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(),
                     "traffic_source": f"{TRAFFIC_PROFILE}: committed rocprofv3 --pmc "
                                       "FETCH_SIZE/WRITE_SIZE passes of this kernel and "
                                       "workload (not measured in this run)"},
        "compute_bound": compute_bound(ms_kernel),
        "secondary": secondary,
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if cpu_all is not None:
        secondary["cpu_all_cores"] = cpu_all
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
