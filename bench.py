#!/usr/bin/env python3
"""Benchmark of the slab C2C FFTND path on one MI355X.

Workload (the metric's configuration): FFTND C2C complex128 512x512x512 at
one rank, input std::mt19937_64 seed 43 uniform [-1,1).  One step = forward
execute + backward execute + normalize through the reference-style C API
(each call host-synchronous, as in SHAFFT).  GFLOP/s counts 5 N log2 N per
direction (2 directions per step).

Also reported, on the same JSON line:
  * roofline: the dominant kernel (the forward pass kernel), its average
    launch duration from HIP events recorded on the plan's stream around the
    forward executes of the timed region, and its algorithmic bytes per
    launch (one read + one write of the 2^27 complex128 grid = 4.295 GB);
  * serving: the served half of the metric -- p50/p99 latency of the mixed
    stream (Poisson arrivals at 2000 jobs/s, sizes 2^10..2^26, 70% complex64)
    from resident plans, its saturated throughput against the plain
    per-call path on the same jobs, and the same for the sub-stream of jobs
    up to 2^20 elements;
  * cpu_baseline: FFTW is not installed, so the labelled fallback
    scipy.fft (pocketfft) on the host threads, forward, backward and step
    timed separately plus a 1-thread forward, with the oracle's radix-2 FFT
    as a secondary figure.

Other workloads: --workload fft1d | batch | serve | serve-passes (per-pass rates
of the serving stream's largest jobs: --prec c2c|z2z|any, --count jobs).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                       [--no-serving] [--dims 512,512,512] [--prec z2z|c2c]
                       [--calibrate-copy]
"""
import argparse
import ctypes as ct
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "C2C GFLOP/s (5N log2 N) & % HBM peak, 512^3 fp64; served p50/p99 latency"
HBM_PEAK_GBS = 8000.0


class Hip:
    """Minimal HIP runtime binding for streams and events (timing only)."""

    def __init__(self):
        self.L = ct.CDLL("libamdhip64.so")
        L = self.L
        L.hipStreamCreate.argtypes = [ct.POINTER(ct.c_void_p)]
        L.hipEventCreate.argtypes = [ct.POINTER(ct.c_void_p)]
        L.hipEventRecord.argtypes = [ct.c_void_p, ct.c_void_p]
        L.hipEventSynchronize.argtypes = [ct.c_void_p]
        L.hipEventElapsedTime.argtypes = [ct.POINTER(ct.c_float), ct.c_void_p, ct.c_void_p]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamSynchronize.argtypes = [ct.c_void_p]

    def stream(self):
        s = ct.c_void_p()
        assert self.L.hipStreamCreate(ct.byref(s)) == 0
        return s.value

    def event(self):
        e = ct.c_void_p()
        assert self.L.hipEventCreate(ct.byref(e)) == 0
        return e.value

    def record(self, e, s):
        assert self.L.hipEventRecord(ct.c_void_p(e), ct.c_void_p(s)) == 0

    def elapsed_ms(self, a, b):
        self.L.hipEventSynchronize(ct.c_void_p(b))
        f = ct.c_float()
        assert self.L.hipEventElapsedTime(ct.byref(f), ct.c_void_p(a), ct.c_void_p(b)) == 0
        return f.value

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0


def _median_times(fn, reps, warm):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def _host_threads():
    """Host threads the CPU legs may use: the job's CPU share (OMP_NUM_THREADS
    when set), not every core the machine reports."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else (os.cpu_count() or 1)


def cpu_baseline(x, dims, prec, reps=10, warm=2):
    """The reference's CPU path (its FFTW backend: guru plan over every axis,
    out-of-place execute, OpenMP normalize; src/cpp/fft/fftw_method/
    fftnd_method.cpp:34-276, common/normalize.cpp:6-32) is not runnable here:
    FFTW is not installed on this host.  The labelled fallback is scipy.fft
    (pocketfft) with workers = the host threads, on the full grid: forward,
    backward and the step (forward + backward with the 1/N normalisation
    folded into ifftn), median of `reps` after `warm` warm-ups; then a
    1-thread forward on two transforms; and the oracle's own radix-2 FFT
    ("port") on a few forward transforms as a secondary figure."""
    n = int(np.prod(dims))
    flop = 5 * n * math.log2(n)
    a = x.reshape(dims)
    ncores = _host_threads()
    out = {"unit": "GFLOP/s", "cores": ncores}
    try:
        import scipy
        import scipy.fft as sfft
    except ImportError:
        sfft = None
    if sfft is not None:
        box = {}
        fwd = lambda: box.__setitem__("y", sfft.fftn(a, workers=ncores))  # noqa: E731
        t_f, _ = _median_times(fwd, reps, warm)
        y = box.pop("y")
        bwd = lambda: sfft.ifftn(y, norm="forward", workers=ncores)  # noqa: E731
        t_b, _ = _median_times(bwd, reps, warm)
        step = lambda: sfft.ifftn(sfft.fftn(a, workers=ncores), workers=ncores)  # noqa: E731
        t_s, _ = _median_times(step, max(3, reps // 2), 1)
        del y
        t_1, _ = _median_times(lambda: sfft.fftn(a, workers=1), 2, 0)
        out.update({
            "value": round(2 * flop / t_s / 1e9, 3),
            "kind": "port",
            "library": f"pocketfft (scipy {scipy.__version__}) fallback; FFTW unavailable",
            "forward_ms": round(t_f * 1e3, 1),
            "backward_ms": round(t_b * 1e3, 1),
            "step_ms": round(t_s * 1e3, 1),
            "forward_gflops": round(flop / t_f / 1e9, 3),
            "backward_gflops": round(flop / t_b / 1e9, 3),
            "forward_gflops_1thread": round(flop / t_1 / 1e9, 3),
            "sample": f"full {'x'.join(map(str, dims))} grid in FP64; step = fftn + ifftn (1/N folded in), "
                      f"median of {max(3, reps // 2)} after 1 warm-up; forward and backward: median of {reps} "
                      f"after {warm} warm-ups, workers={ncores}; 1 thread: median of 2 forward transforms",
        })
    import oracle  # the CPU-baseline leg is the only bench use of the oracle

    a128 = a.astype(np.complex128, copy=False)
    t_p, ts = _median_times(lambda: oracle.fftn_pow2(a128, None, -1), 3, 0)
    port = {"forward_gflops": round(flop / t_p / 1e9, 3), "cores": oracle.threads(),
            "sample": "median of 3 forward transforms of the full grid, the oracle's FP64 radix-2 FFT "
                      "with OpenMP over lines"}
    if sfft is None:
        out.update({"value": port["forward_gflops"], "kind": "port",
                    "sample": "scipy not importable; FFTW unavailable: " + port["sample"]})
    else:
        out["oracle_port"] = port
    return out


def profiled_traffic(dims, prec):
    """HBM bytes per launch of the forward pass kernel measured with rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE (scaled by a known-size copy), read from the
    committed profile summary when it covers this workload."""
    name = f"fft_fixed_pass_{'x'.join(map(str, dims))}_{'z2z' if prec else 'c2c'}.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def bench_batch(args, hip):
    """Batched uneven workload: `count` complex64 2-D/3-D transforms, axes
    log-uniform over the 7-smooth lengths in [16, 256].  Reports the batched
    graph launch against the plain per-call path (one FFTND plan per shape,
    host-synchronous execute per transform, as the reference API runs)."""
    from shafft_amd import api
    from shafft_amd.workloads import draw_shapes

    shapes = draw_shapes(args.count, 45)
    prec = api.C2C
    bp = api.BatchPlan(shapes, prec)
    tot = bp.total
    x = api.synthetic_input(tot, 45, prec, api.UNIFORM)  # std::mt19937_64 seed 45
    a, o, w = (api.DeviceBuffer(tot, prec) for _ in range(3))
    a.upload(x)
    stream = hip.stream()
    bp.set_stream(stream)
    bp.set_buffers(a.ptr, o.ptr, w.ptr)
    tune = bp.autotune() if not args.no_autotune else "defaults"
    flops = sum(5 * int(np.prod(sh)) * math.log2(int(np.prod(sh))) for sh in shapes)
    for _ in range(args.warmup):
        bp.execute(api.FORWARD)
    bp.synchronize()
    e0, e1 = hip.event(), hip.event()
    t0 = time.perf_counter()
    hip.record(e0, stream)
    for _ in range(args.steps):
        bp.execute(api.FORWARD)
    hip.record(e1, stream)
    bp.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    ms = hip.elapsed_ms(e0, e1) / args.steps
    info = bp.info()
    # FP32 vs split-precision matrix-core radix-8/16 stages: time and error
    # against an FP64 FFT (numpy, host) on every 16th transform
    sample = list(range(0, len(shapes), 16))
    xs_s = [x[bp.offsets[i]:bp.offsets[i] + int(np.prod(shapes[i]))].reshape(shapes[i]) for i in sample]
    refs = [np.fft.fftn(v.astype(np.complex128)) for v in xs_s]
    modes = {}
    wanted = args.batch_modes.split(",")
    for mode, name in ((api.FP32, "fp32"), (api.SPLIT_BF16, "bf16_split"), (api.SPLIT_FP16, "fp16_split")):
        if name not in wanted:
            continue
        bp.set_precision_mode(mode)
        # each mode runs the schedule the autotuner picks for it (the split
        # stages use more registers than the FP32 ones)
        mode_tune = bp.autotune() if not args.no_autotune else ""
        mode_sched = mode_tune.strip().splitlines()[-1] if mode_tune else "defaults"
        for _ in range(args.warmup):
            bp.execute(api.FORWARD)
        bp.synchronize()
        hip.record(e0, stream)
        for _ in range(args.steps):
            bp.execute(api.FORWARD)
        hip.record(e1, stream)
        bp.synchronize()
        mms = hip.elapsed_ms(e0, e1) / args.steps
        errs = []
        for i, ref in zip(sample, refs):
            y = api.download_ptr(o.ptr + bp.offsets[i] * 8, ref.size, prec).reshape(ref.shape)
            errs.append(float(np.linalg.norm(y - ref) / np.linalg.norm(ref)))
        modes[name] = {"gpu_ms": round(mms, 4), "gflops": round(flops / (mms * 1e-3) / 1e9, 1),
                       "schedule": mode_sched,
                       "relL2_vs_fp64_median": float(f"{np.median(errs):.3e}"),
                       "relL2_vs_fp64_max": float(f"{np.max(errs):.3e}")}
    bp.set_precision_mode(api.FP32)
    if not args.no_autotune and len(wanted) > 1:
        bp.autotune()  # back to the FP32 schedule for the plain-path comparison
    if args.skip_plain:
        plain = float("nan")
    else:
        plain = _batch_plain(shapes, prec, bp, a, o, stream, hip)
    out = {
        "metric": "batched uneven FFTND throughput (GFLOP/s, 5 N log2 N per transform)",
        "value": round(flops / (wall * 1e-3 * 1e3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(shapes)} complex64 2-D/3-D transforms, axes log-uniform over the 64 7-smooth "
                        f"lengths in [16,256], one persistent graph launch per step (forward)",
            "elements": tot,
            "gpu_ms_per_batch": round(ms, 4),
            "plain_per_call_ms": None if args.skip_plain else round(plain * 1e3, 2),
            "speedup_vs_plain": None if args.skip_plain else round(plain / wall, 2),
            "plain_gflops": None if args.skip_plain else round(flops / plain / 1e9, 2),
            "hbm_gbs": round(bp.info()["bytes"] / (ms * 1e-3) / 1e9, 1),
            "passes": info["passes"],
            "items": info["items"],
            "grid": info["grid"],
            "schedule": tune.strip().splitlines()[-1] if tune else "",
            "precision_modes": modes,
            "error_sample": f"{len(sample)} transforms (every 16th), forward, vs numpy FP64 FFT",
        },
    }
    print(json.dumps(out), flush=True)


def _batch_plain(shapes, prec, bp, a, o, stream, hip):
    """Plain per-call path: one FFTND plan per shape, host-synchronous
    execute per transform, as the reference API runs."""
    from shafft_amd import api

    plans = []
    for i, sh in enumerate(shapes):
        p = api.FFTND()
        p.init([1] * len(sh), sh, prec)
        p.plan()
        p.set_stream(stream)
        off = bp.offsets[i] * 8
        p.set_buffers(a.ptr + off, o.ptr + off)
        plans.append(p)
    for p in plans[:64]:
        p.execute(api.FORWARD)
    hip.sync()
    t0 = time.perf_counter()
    for p in plans:
        p.execute(api.FORWARD)   # host-synchronous, result swapped into data
    hip.sync()
    return time.perf_counter() - t0


def bench_fft1d(args, hip):
    """FFT1D of length n at one rank (four-step / Bluestein), forward
    execute through the seam with the input preserved; Gaussian inputs."""
    from shafft_amd import api

    n = args.n
    prec = api.Z2Z if args.prec == "z2z" else api.C2C
    x = api.synthetic_input(n, 44, prec, api.GAUSSIAN)  # std::mt19937_64 seed 44, Box-Muller
    p = api.OneDBackendPlan(n, prec)
    stream = hip.stream()
    p.set_stream(stream)
    npass = p.num_passes()
    pass_bytes = [p.pass_bytes(i) for i in range(npass)]
    # buffers cover every pass's extents (Bluestein intermediates span M)
    ext = max(max(p.pass_extents(i)) for i in range(npass)) if npass else n
    a, b = api.DeviceBuffer(ext, prec), api.DeviceBuffer(ext, prec)
    a.upload(x)
    del x
    if args.calibrate_copy:
        esz = 16 if prec == api.Z2Z else 8
        hip.L.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        assert hip.L.hipMemcpy(ct.c_void_p(b.ptr), ct.c_void_p(a.ptr), ct.c_size_t(n * esz), 3) == 0
        hip.sync()
        print(f"calibration copy: {n * esz} bytes read, {n * esz} bytes written", file=sys.stderr)
    for _ in range(args.warmup):
        p.execute(a.ptr, b.ptr, api.FORWARD)
    p.synchronize()
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(args.steps):
        p.execute(a.ptr, b.ptr, api.FORWARD)
    hip.record(e1, stream)
    p.synchronize()
    ms = hip.elapsed_ms(e0, e1) / args.steps
    # per pass, each launched alone between events
    per = []
    for i in range(npass):
        hip.record(e0, stream)
        for _ in range(args.steps):
            p.execute_pass(i, a.ptr, ext, b.ptr, ext, api.FORWARD)
        hip.record(e1, stream)
        per.append(hip.elapsed_ms(e0, e1) / args.steps)
    flop = 5 * n * math.log2(n)
    out = {
        "metric": "FFT1D GFLOP/s (5 N log2 N)",
        "value": round(flop / (ms * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == api.Z2Z else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"FFT1D {'complex128' if prec == api.Z2Z else 'complex64'} N={n}, one rank, forward, "
                        f"input preserved",
            "passes": npass,
            "plan": p.describe().strip().splitlines(),
            "pass_ms": [round(v, 4) for v in per],
            "pass_gbs": [round(b / (v * 1e-3) / 1e9, 1) for b, v in zip(pass_bytes, per)],
            "hbm_frac": round(sum(pass_bytes) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "compulsory_hbm_frac": round(2 * n * (16 if prec == api.Z2Z else 8) / (ms * 1e-3) / 1e9
                                         / HBM_PEAK_GBS, 4),
        },
    }
    print(json.dumps(out), flush=True)


def _serve_run(srv, reqs, inp, outs, paced: bool):
    """Submit `reqs` to the server: at their Poisson arrival times (paced)
    or all at once.  Output buffers are reused round-robin, so request i
    waits for request i - len(outs) before it is submitted; latency is
    counted from the scheduled arrival, so that wait shows up in it.
    Returns the ServedJob results in request order."""
    from shafft_amd import api

    k = len(outs)
    ids, res = [], [None] * len(reqs)
    t0 = api.Server.now_ns() + 2_000_000
    while time.monotonic_ns() < t0:
        pass
    for i, r in enumerate(reqs):
        arr = t0 + int(r.t * 1e9) if paced else t0
        if paced:
            while time.monotonic_ns() < arr:
                pass
        if i >= k:
            res[i - k] = srv.wait(ids[i - k])
        o = outs[i % k]
        if r.kind == 1:
            ids.append(srv.submit_1d(r.n, r.prec, r.direction, inp[r.prec].ptr, o.ptr, arrival_ns=arr))
        else:
            ids.append(srv.submit_nd(r.dims, r.prec, r.direction, inp[r.prec].ptr, o.ptr, arrival_ns=arr))
    for i in range(max(0, len(reqs) - k), len(reqs)):
        res[i] = srv.wait(ids[i])
    bad = [(reqs[i].describe(), r.message) for i, r in enumerate(res) if r.status != 0]
    if bad:
        raise RuntimeError(f"served jobs failed: {bad[:3]}")
    return res


def _lat_stats(res):
    lat = np.sort(np.array([r.latency_ms for r in res]))
    rank = lambda q: float(lat[max(0, int(math.ceil(q * len(lat))) - 1)])  # noqa: E731
    span = (max(r.done for r in res) - min(r.arrival for r in res)) * 1e-9
    return {"p50_ms": round(rank(0.5), 3), "p99_ms": round(rank(0.99), 3), "mean_ms": round(float(lat.mean()), 3),
            "max_ms": round(float(lat[-1]), 3), "span_s": span}


def _serve_inputs(reqs):
    from shafft_amd import api

    max_e = max(r.elems for r in reqs)
    inp = {}
    for prec in (api.C2C, api.Z2Z):
        inp[prec] = api.DeviceBuffer(max_e, prec)
        inp[prec].upload(api.synthetic_input(max_e, 46, prec, api.UNIFORM))  # std::mt19937_64 seed 46
    return max_e, inp


def _plain_run(reqs, hip, src, dst, with_planning=False):
    """The plain per-call path on the same jobs: the reference call sequence
    (src/cpp/shafft_internal.cpp:62-127), one plan per shape and one
    host-synchronous execute per job, one job at a time.  Plans are made
    beforehand (not timed) unless with_planning, where every job runs init,
    plan, execute and destroy as for a shape not seen before."""
    from shafft_amd import api

    stream = hip.stream()

    def make(r):
        if r.kind == 1:
            p = api.FFT1D()
            p.init(r.n, r.n, 0, r.prec)
        else:
            p = api.FFTND()
            p.init([1] * len(r.dims), r.dims, r.prec)
        p.plan()
        p.set_stream(stream)
        return p

    key = lambda r: (r.kind, r.n, tuple(r.dims or ()), r.prec)  # noqa: E731
    plans = {}
    if not with_planning:
        for r in reqs:
            if key(r) not in plans:
                plans[key(r)] = make(r)
    hip.sync()
    t0 = time.perf_counter()
    for r in reqs:
        p = make(r) if with_planning else plans[key(r)]
        p.set_buffers(src.ptr, dst.ptr)
        p.execute(api.BACKWARD if r.direction else api.FORWARD)  # host-synchronous
        if with_planning:
            p.release()
    hip.sync()
    return time.perf_counter() - t0


def serve_measure(reqs, hip, slots=4, arena_gb=24, outstanding=48, full=True, planners=8):
    """Serves `reqs` (each job reads a shared input and writes one of
    `outstanding` output buffers): a cold paced run that builds the plans,
    a paced run from resident plans (latency), every job queued at once
    (throughput; median of three runs), and the plain per-call path on the
    same jobs.  full adds
    the cold saturated run and the plain path with planning per call."""
    from shafft_amd import api

    max_e, inp = _serve_inputs(reqs)
    outs = [api.DeviceBuffer(max_e, api.Z2Z) for _ in range(outstanding)]
    gflop = sum(r.gflop for r in reqs)
    srv = api.Server(slots=slots, arena_bytes=arena_gb << 30, planners=planners)
    cold = _lat_stats(_serve_run(srv, reqs, inp, outs, paced=True))
    st_cold = srv.stats()
    srv.reset_stats()
    warm = _lat_stats(_serve_run(srv, reqs, inp, outs, paced=True))
    # saturated: every job queued at once, three runs, the median by span
    sats = []
    for _ in range(3):
        srv.reset_stats()
        r = _serve_run(srv, reqs, inp, outs, paced=False)
        sats.append((_lat_stats(r)["span_s"], r, srv.stats(), srv.batch_stats()))
    sats.sort(key=lambda t: t[0])
    _, sat_res, st, (nb, nbj) = sats[1]
    sat = _lat_stats(sat_res)
    srv.close()
    gpu_busy_s = sum(r.gpu_ns for r in sat_res) * 1e-9
    # the plain path may overwrite its source (execute leaves the result in
    # whichever buffer it ends in); the served runs above are done with it
    plain = _plain_run(reqs, hip, inp[api.Z2Z], outs[0])
    out = {
        "jobs": len(reqs),
        "total_gflop": round(gflop, 2),
        "p50_ms": warm["p50_ms"],
        "p99_ms": warm["p99_ms"],
        "mean_ms": warm["mean_ms"],
        "max_ms": warm["max_ms"],
        "paced_gflops": round(gflop / warm["span_s"], 2),
        "cold_p50_ms": cold["p50_ms"],
        "cold_p99_ms": cold["p99_ms"],
        "cold_plan_builds": st_cold["plan_builds"],
        "saturated_gflops": round(gflop / sat["span_s"], 2),
        "saturated_req_per_s": round(len(reqs) / sat["span_s"], 1),
        "saturated_p50_ms": sat["p50_ms"],
        "saturated_p99_ms": sat["p99_ms"],
        "gpu_busy_frac": round(gpu_busy_s / sat["span_s"], 3),
        "plain_per_call_s": round(plain, 4),
        "plain_per_call_gflops": round(gflop / plain, 2),
        "plain_per_call_req_per_s": round(len(reqs) / plain, 1),
        "speedup_vs_plain": round(plain / sat["span_s"], 3),
        "micro_batched_jobs": nbj,
        "micro_batches": nb,
        "plans_resident": st["plans"],
        "table_mb": round(st["table_bytes"] / 2**20, 2),
        "arena_peak_gb": round(st["arena_peak"] / 2**30, 3),
    }
    if full:
        srv = api.Server(slots=slots, arena_bytes=arena_gb << 30, planners=planners)
        cold_sat = _lat_stats(_serve_run(srv, reqs, inp, outs, paced=False))
        srv.close()
        plain_planned = _plain_run(reqs, hip, inp[api.Z2Z], outs[0], with_planning=True)
        out.update({"cold_saturated_s": round(cold_sat["span_s"], 4),
                    "plain_with_planning_s": round(plain_planned, 4),
                    "cold_speedup_vs_plain_with_planning": round(plain_planned / cold_sat["span_s"], 3)})
    return out


STREAM_DESC = ("Poisson arrivals at {rate:g}/s, sizes log-uniform 2^{lo}..2^{hi} (7-smooth), 70% complex64 / "
               "30% complex128, fwd/bwd 50/50, half FFTND 2-D/3-D, half FFT1D")


def serving_summary(hip, requests=2000, rate=2000.0):
    """The serving half of the metric: the mixed stream at its configured
    range (2^10..2^26) and the sub-stream of jobs up to 2^20 elements, where
    per-call overheads rather than HBM bandwidth bound the plain path."""
    from shafft_amd.workloads import draw_stream

    full = serve_measure(draw_stream(requests, seed=46, rate=rate), hip, full=True)
    # up to 256 small jobs outstanding (16 MiB output buffers each): the
    # saturated client keeps enough requests queued for groups to form
    small = serve_measure(draw_stream(requests, seed=46, rate=rate, hi=20), hip, outstanding=256, full=False)
    return {
        "workload": f"{requests} jobs, " + STREAM_DESC.format(rate=rate, lo=10, hi=26),
        **{k: full[k] for k in ("p50_ms", "p99_ms", "cold_p50_ms", "cold_p99_ms", "saturated_gflops",
                                "saturated_req_per_s", "plain_per_call_gflops", "speedup_vs_plain",
                                "gpu_busy_frac", "micro_batched_jobs", "micro_batches", "cold_plan_builds",
                                "cold_saturated_s", "plain_with_planning_s",
                                "cold_speedup_vs_plain_with_planning")},
        "target_speedup_vs_plain": 3.0,
        "substream_le_2p20": {
            "workload": f"{requests} jobs, " + STREAM_DESC.format(rate=rate, lo=10, hi=20),
            **{k: small[k] for k in ("p50_ms", "p99_ms", "saturated_req_per_s", "saturated_gflops",
                                     "plain_per_call_req_per_s", "plain_per_call_gflops", "speedup_vs_plain",
                                     "micro_batched_jobs", "micro_batches")},
        },
    }


def bench_serve(args, hip):
    """Serving stream: Poisson arrivals of FFTND/FFT1D jobs, sizes
    log-uniform in 2^lo..2^hi elements (7-smooth), 70% complex64 / 30%
    complex128, forward or backward.  Reports the p50/p99 latency at the
    stated arrival rate (plans built on demand, then resident), the
    throughput with every job queued at once, and the plain per-call path
    (the reference call sequence: one plan per shape, host-synchronous
    execute, one job at a time) on the same jobs."""
    from shafft_amd.workloads import draw_stream

    reqs = draw_stream(args.requests, seed=46, rate=args.rate, hi=args.serve_hi)
    outstanding = args.outstanding or (256 if args.serve_hi <= 20 else 48)
    m = serve_measure(reqs, hip, slots=args.slots, arena_gb=args.arena_gb, outstanding=outstanding,
                      planners=args.planners)
    out = {
        "metric": "served mixed-size stream throughput (GFLOP/s, 5 N log2 N per job); p50/p99 latency",
        "value": m["saturated_gflops"],
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(m["total_gflop"] / m["saturated_gflops"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(reqs)} jobs, " + STREAM_DESC.format(rate=args.rate, lo=10, hi=args.serve_hi),
            "slots": args.slots,
            "planner_threads": args.planners,
            "outstanding_buffers": outstanding,
            **m,
        },
    }
    print(json.dumps(out), flush=True)


def bench_serve_passes(args, hip):
    """Per-pass HBM rates of the serving stream's largest distinct jobs: each
    pass of each job launched `steps` times back to back between HIP events
    on its own stream (algorithmic bytes 2 * lines * n * sizeof(T) per pass).
    Reports every pass and the totals per precision (all passes, and the
    generic-kernel passes alone)."""
    from shafft_amd import api
    from shafft_amd.workloads import draw_stream

    reqs = [r for r in draw_stream(args.requests, seed=46, rate=args.rate) if args.prec == "any" or
            r.prec == (api.Z2Z if args.prec == "z2z" else api.C2C)]
    reqs.sort(key=lambda r: -r.elems * (16 if r.prec else 8))
    jobs, seen = [], set()
    for r in reqs:
        key = (r.kind, r.n, tuple(r.dims or ()), r.prec)
        if key not in seen:
            seen.add(key)
            jobs.append(r)
        if len(jobs) >= args.count:
            break
    stream = hip.stream()
    e0, e1 = hip.event(), hip.event()
    cap = max(r.elems for r in jobs) * 2  # complex128 elements (Bluestein-free: 7-smooth sizes)
    a, b = api.DeviceBuffer(cap, api.Z2Z), api.DeviceBuffer(cap, api.Z2Z)
    rows, tot = [], {}
    for r in jobs:
        p = (api.OneDBackendPlan(r.n, r.prec) if r.kind == 1 else
             api.NDBackendPlan(list(r.dims), list(range(len(r.dims))), r.prec))
        p.set_stream(stream)
        desc = p.describe().strip().splitlines()
        elems = cap * 16 // (16 if r.prec else 8)
        for i in range(p.num_passes()):
            p.execute_pass(i, a.ptr, elems, b.ptr, elems, r.direction)
            hip.record(e0, stream)
            for _ in range(args.steps):
                p.execute_pass(i, a.ptr, elems, b.ptr, elems, r.direction)
            hip.record(e1, stream)
            ms = hip.elapsed_ms(e0, e1) / args.steps
            by = p.pass_bytes(i)
            generic = "fixed" not in desc[i]
            t = tot.setdefault(r.prec, [0, 0.0, 0, 0.0])
            t[0] += by
            t[1] += ms
            if generic:
                t[2] += by
                t[3] += ms
            rows.append({"job": r.describe(), "pass": desc[i], "us": round(ms * 1e3, 1),
                         "gbs": round(by / (ms * 1e-3) / 1e9, 1)})
    for row in rows:
        print(f"{row['job']:38s} {row['pass'][:100]:100s} {row['us']:8.1f} us {row['gbs']:8.1f} GB/s", flush=True)
    out = {"metric": "per-pass HBM rate of the serving stream's largest jobs (GB/s, algorithmic bytes)",
           "unit": "GB/s", "jobs": len(jobs), "steps": args.steps}
    for pr, t in sorted(tot.items()):
        name = "complex128" if pr else "complex64"
        out[name] = {"all_passes_gbs": round(t[0] / (t[1] * 1e-3) / 1e9, 1),
                     "generic_passes_gbs": round(t[2] / max(t[3] * 1e-3, 1e-12) / 1e9, 1)}
    print(json.dumps(out), flush=True)


def bench_job_pass(args, hip):
    """One pass of one job (FFT1D of --n points, or FFTND of --dims over every
    axis), launched `steps` times back to back between HIP events: the unit a
    counter run (rocprofv3 --pmc) is taken on."""
    from shafft_amd import api

    prec = api.Z2Z if args.prec == "z2z" else api.C2C
    if args.dims_nd:
        dims = [int(v) for v in args.dims_nd.split(",")]
        p = api.NDBackendPlan(dims, list(range(len(dims))), prec)
        elems = int(np.prod(dims))
        name = "FFTND " + "x".join(map(str, dims))
    else:
        p = api.OneDBackendPlan(args.n, prec)
        elems = args.n
        name = f"FFT1D N={args.n}"
    stream = hip.stream()
    p.set_stream(stream)
    ext = max(max(p.pass_extents(i)) for i in range(p.num_passes()))
    a, b = api.DeviceBuffer(ext, prec), api.DeviceBuffer(ext, prec)
    i = args.pass_index
    desc = p.describe().strip().splitlines()[i]
    for _ in range(args.warmup):
        p.execute_pass(i, a.ptr, ext, b.ptr, ext, api.FORWARD)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(args.steps):
        p.execute_pass(i, a.ptr, ext, b.ptr, ext, api.FORWARD)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1) / args.steps
    by = p.pass_bytes(i)
    print(json.dumps({"metric": "one pass of one transform (GB/s, algorithmic bytes)", "job": name,
                      "precision": "complex128" if prec == api.Z2Z else "complex64", "elements": elems,
                      "pass": desc, "us": round(ms * 1e3, 1), "gbs": round(by / (ms * 1e-3) / 1e9, 1),
                      "steps": args.steps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dims", type=str, default="512,512,512")
    ap.add_argument("--prec", type=str, default="z2z", choices=["c2c", "z2z", "any"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="fftnd",
                    choices=["fftnd", "batch", "serve", "fft1d", "serve-passes", "job-pass"])
    ap.add_argument("--dims-nd", type=str, default="", help="job-pass: FFTND dims (else FFT1D of --n)")
    ap.add_argument("--pass-index", type=int, default=0, help="job-pass: which pass of the plan")
    ap.add_argument("--n", type=int, default=1 << 27, help="fft1d length")
    ap.add_argument("--requests", type=int, default=2000, help="jobs in the serve workload")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve arrival rate (jobs/s)")
    ap.add_argument("--slots", type=int, default=4,
                    help="server stream slots (4: one per HIP hardware queue of the process)")
    ap.add_argument("--planners", type=int, default=8, help="server plan-building threads")
    ap.add_argument("--outstanding", type=int, default=0,
                    help="serve output buffers (max job size each; 0: 256 up to 2^20-element jobs, else 48)")
    ap.add_argument("--arena-gb", type=int, default=24, help="server scratch arena (GiB)")
    ap.add_argument("--serve-hi", type=int, default=26, help="serve: largest job is 2^serve_hi elements")
    ap.add_argument("--no-serving", action="store_true",
                    help="fftnd: skip the serving-stream summary on the JSON line")
    ap.add_argument("--count", type=int, default=4096, help="transforms in the batch workload")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-modes", default="fp32,bf16_split,fp16_split",
                    help="radix-8/16 stage precisions timed in the batch workload")
    ap.add_argument("--skip-plain", action="store_true", help="batch: skip the plain per-call comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one device-to-device copy of the data buffer first (known-size "
                         "reference kernel for HBM counter scaling under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "shafft_amd runs one process on one GPU; multi-GPU "
                                   "(slab-sharded) runs are not implemented"}))
        sys.exit(2)

    if args.workload == "batch":
        bench_batch(args, Hip())
        return
    if args.workload == "serve":
        bench_serve(args, Hip())
        return
    if args.workload == "fft1d":
        bench_fft1d(args, Hip())
        return
    if args.workload == "serve-passes":
        bench_serve_passes(args, Hip())
        return
    if args.workload == "job-pass":
        bench_job_pass(args, Hip())
        return

    from shafft_amd import api

    dims = tuple(int(v) for v in args.dims.split(","))
    prec = api.Z2Z if args.prec == "z2z" else api.C2C
    esz = 16 if prec == api.Z2Z else 8
    n = int(np.prod(dims))
    hip = Hip()
    stream = hip.stream()

    x = api.synthetic_input(n, 43, prec, api.UNIFORM)  # std::mt19937_64 seed 43, uniform [-1,1)

    plan = api.FFTND()
    plan.init([1] * len(dims), dims, prec)
    plan.plan()
    plan.set_stream(stream)
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    plan.set_buffers(a.ptr, b.ptr)

    # seam handle with the same pass program, for per-pass byte counts
    seam = api.NDBackendPlan(dims, list(range(len(dims))), prec)
    npass = seam.num_passes()
    pass_bytes = [seam.pass_bytes(i) for i in range(npass)]
    del seam

    if args.calibrate_copy:
        hip.L.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        assert hip.L.hipMemcpy(ct.c_void_p(b.ptr), ct.c_void_p(a.ptr), ct.c_size_t(n * esz), 3) == 0
        hip.sync()
        print(f"calibration copy: {n * esz} bytes read, {n * esz} bytes written", file=sys.stderr)

    for _ in range(args.warmup):
        plan.execute(api.FORWARD)
        plan.execute(api.BACKWARD)
        plan.normalize()
    hip.sync()

    ev = [(hip.event(), hip.event()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        hip.record(ev[i][0], stream)
        plan.execute(api.FORWARD)
        hip.record(ev[i][1], stream)
        plan.execute(api.BACKWARD)
        plan.normalize()
    hip.sync()
    dt = time.perf_counter() - t0
    fwd_ms = [hip.elapsed_ms(e0, e1) for e0, e1 in ev]

    flops_step = 2 * 5 * n * math.log2(n)
    ms_step = dt * 1e3 / args.steps
    gflops = flops_step * args.steps / dt / 1e9
    # dominant kernel: the forward pass kernel, npass launches per forward execute
    avg_launch_ms = float(np.mean(fwd_ms)) / npass
    bytes_launch = float(np.mean(pass_bytes))
    achieved = bytes_launch / (avg_launch_ms * 1e-3) / 1e9
    step_bytes = 2 * sum(pass_bytes) + 2 * n * esz  # fwd + bwd + normalize
    out = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == api.Z2Z else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"FFTND C2C {'complex128' if prec == api.Z2Z else 'complex64'} "
                        f"{'x'.join(map(str, dims))}, one rank; step = forward + backward + normalize",
            "dims": list(dims),
            "passes_per_direction": npass,
            "forward_ms": round(float(np.mean(fwd_ms)), 4),
            "forward_gflops": round(5 * n * math.log2(n) / (np.mean(fwd_ms) * 1e-3) / 1e9, 2),
            "step_hbm_frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "forward_hbm_frac": round(sum(pass_bytes) / (np.mean(fwd_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "pass_bytes": pass_bytes,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(dims, prec),
            "kernel": f"fft_fixed_pass<{'double' if prec == api.Z2Z else 'float'},{dims[-1]},-1> (forward)",
            "bytes_per_launch": int(bytes_launch),
            "avg_launch_ms": round(avg_launch_ms, 4),
        },
    }
    del a, b, plan
    if not args.no_serving:
        out["serving"] = serving_summary(hip)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(x, dims, prec)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
