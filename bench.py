#!/usr/bin/env python3
"""Benchmark of the MoFFT 1-D C2C path on one MI355X.

Default workload (one "step"): 256 FP64 transforms of N = 2^16, forward then
backward, in place, on a 256 MiB device buffer with iid N(0,1) components -
the FP64 forward+backward in-place configuration the metric is quoted on.
Prints ONE JSON line:

  value      GFLOP/s under the 5*N*log2(N) convention (tools/benchmark.cpp:709),
             both directions counted, whole job, inputs resident in HBM: K
             steps back to back between two HIP events
  p50/p99    per-step latency from a run of K steps before that one, each
             step between its own events
  roofline   dominant kernel vs the 8 TB/s HBM3E peak.  Algorithmic bytes are
             one read + one write of every point per transform direction
             (2*N*16 B per FP64 transform); a plan with P HBM passes is
             charged 1/P of that per pass launch.  The dominant kernel is
             the pass with the longest launches, timed with HIP events
             between consecutive pass launches on the launch stream over
             the same number of executions as the timed region
             (`event_launch_us`); its launch duration in the timed region
             (`avg_launch_us`) is its share of the timed step in proportion
             to those bracketed durations (an event pair adds a gap to the
             launch it brackets); `achieved` = algorithmic bytes per launch
             / avg_launch_us.  `traffic` is the
             measured HBM bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE,
             profiles/), or null when no counter summary is present.
  protocol   the reference benchmark's sampling (tools/benchmark.cpp:65-130)
             on the same step: warm-up of at least 1000 steps and 0.25 s,
             repetitions sized so one sample takes about 2 ms, 31 samples;
             median ms per step, 95 % bootstrap CI of the median (2000
             resamples, seed 0x4d6f464654 + samples, :119-130) and the
             coefficient of variation
  cpu_baseline  the reference's executor (radix stages as DFT-matrix
             contractions, transposes between stages; tables made at plan
             time) restated in C and compiled for the host
             (oracle/mofft_cpu_port.c), one core, on a bounded sample of the
             same workload.  FFTW 3.3.11, the comparison the reference uses,
             has no source offline ("fftw": "FFTW baseline unavailable");
             beside it, clearly labelled, numpy/scipy's pocketfft on the
             same sample, one thread and all host threads
  configs    the other configurations of the workload set, measured in the
             same run: c1 (FP32 N=1024 x 64, latency), c3 (FP32 N=2^4..2^20,
             2^24 points each, plain and matrix-core form, fraction of the
             HBM peak), c4 (4096 uneven FP32 transforms in one persistent
             launch), c5 (the serving loop on a short Poisson window,
             p50/p99 latency); --no-configs leaves them out

Other workloads (--workload c1|c3|c4|c5) print their own lines for study;
the benchmark's result is the default workload's line.  c4: 4096 FP32 transforms of N = 2^U{6..18}
(seeded), half forward, half backward, mixed placement, in one persistent
launch replayed from a captured HIP graph, beside the same transforms
launched one plan execution each.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "copy"])
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="c5: requests/s")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="c5: arrival window")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-log2", default="6,18",
                    help="c4: lengths 2^U{a..b}")
    ap.add_argument("--owned-tiles", type=int, default=0)
    ap.add_argument("--batch-bits", type=int, default=32, choices=[32, 64])
    ap.add_argument("--item-tiles", type=int, default=0)
    ap.add_argument("--share-all", action="store_true",
                    help="c4: every multi-pass transform shared between workgroups")
    ap.add_argument("--batch-grid", type=int, default=0,
                    help="c4: persistent grid size (0: resident workgroups per CU x CUs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--sweep-max", type=int, default=20)
    ap.add_argument("--no-configs", action="store_true",
                    help="default workload only: leave out the c1/c3/c4/c5 summaries")
    ap.add_argument("--configs-serve-seconds", type=float, default=3.0,
                    help="arrival window of the c5 summary in the default run")
    ap.add_argument("--samples", type=int, default=31,
                    help="protocol samples (tools/benchmark.cpp takes >= 7, 30 formal)")
    return ap.parse_args()


def percentile(values, q):
    v = sorted(values)
    if not v:
        return float("nan")
    k = (len(v) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def cpu_baseline_port(n, bits, seconds):
    """The reference executor on one host core (oracle/mofft_cpu_port.c:
    DFT-matrix stages and transposes, coefficients and twiddles made at plan
    time), forward then backward of single transforms of length n with the
    GPU plan's radix chain, until `seconds` of CPU work are spent; the plans
    are made before the timed loop, as the reference makes them once."""
    import numpy as np
    import oracle
    from mofft_amd import describe
    chain = describe(n, 1, bits).chain
    dt = np.complex64 if bits == 32 else np.complex128
    rng = np.random.default_rng(7)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dt)
    y = np.empty_like(x)
    z = np.empty_like(x)
    fwd = oracle.PortPlan(n, chain, -1, bits)
    bwd = oracle.PortPlan(n, chain, 1, bits)
    fwd.execute(x, y)
    bwd.execute(y, z)
    count = 0
    t0 = time.perf_counter()
    while True:
        fwd.execute(x, y)
        bwd.execute(y, z)
        count += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    fwd.destroy()
    bwd.destroy()
    flops = 2 * 5.0 * n * math.log2(n) * count
    return {"value": round(flops / el / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"{count} FP{bits} transforms of N={n}, forward+backward, "
                      f"radix chain {chain}, DFT-matrix stages with tables made "
                      f"at plan time (oracle/mofft_cpu_port.c), {el:.1f} s on one "
                      f"host core"}


def cpu_pocketfft(n, batch, bits, seconds):
    """numpy/scipy's pocketfft (the FFTW stand-in, labelled as such) on the
    same transforms: batches of `batch` forward+backward on one thread and
    on every host thread (at most 16, the box's CPU share)."""
    import numpy as np
    import scipy.fft as sfft
    dt = np.complex64 if bits == 32 else np.complex128
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(dt)
    threads = max(1, min(16, os.cpu_count() or 1))
    out = {"library": "scipy.fft (pocketfft) " + sfft.__name__, "fftw": "FFTW baseline unavailable"}
    for label, w in (("threads_1", 1), ("threads_all", threads)):
        sfft.ifft(sfft.fft(x, workers=w), workers=w)
        count = 0
        t0 = time.perf_counter()
        while True:
            sfft.ifft(sfft.fft(x, workers=w), workers=w)
            count += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
        out[label] = {"value": round(2 * 5.0 * n * math.log2(n) * batch * count / el / 1e9, 3),
                      "unit": "GFLOP/s", "threads": w,
                      "sample": f"{count} x (batch {batch} FP{bits} N={n} forward+backward), {el:.1f} s"}
    return out


def load_traffic(kernel_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        entry = json.load(open(path)).get(kernel_key)
    except Exception:
        return None
    if isinstance(entry, dict):
        return entry.get("bytes_per_launch")
    return entry


def measure_copy(torch, dev, stream, nbytes):
    """Device-to-device copy bandwidth (read + write bytes / s) of a plain
    copy kernel, the practical ceiling the roofline fraction is read against."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        b.copy_(a)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return round(2.0 * nbytes / (ms * 1e-3) / 1e9, 1)


def protocol_samples(torch, stream, step, samples):
    """tools/benchmark.cpp:65-130 on one step function: warm-up until at
    least 1000 steps and 0.25 s (at most 10^6), repetitions per sample sized
    to ~2 ms from the fastest single step, `samples` samples timed with HIP
    events on the launch stream; median, bootstrap 95 % CI of the median
    (2000 resamples, seed 0x4d6f464654 + samples) and the coefficient of
    variation, in ms per step."""
    import numpy as np
    t0 = time.perf_counter()
    warm = 0
    while warm < 1000000 and (warm < 1000 or time.perf_counter() - t0 < 0.25):
        step()
        warm += 1
        if warm % 64 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    warm_s = time.perf_counter() - t0
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fastest = float("inf")
    for _ in range(5):
        e0.record(stream)
        step()
        e1.record(stream)
        torch.cuda.synchronize()
        fastest = min(fastest, e0.elapsed_time(e1) * 1e-3)
    reps = int(min(1000000, max(1, math.ceil(0.002 / fastest))))
    per = []
    for _ in range(samples):
        e0.record(stream)
        for _ in range(reps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        per.append(e0.elapsed_time(e1) / reps)
    x = np.array(per)
    rng = np.random.Generator(np.random.MT19937(0x4D6F464654 + len(per)))
    boots = np.median(x[rng.integers(0, len(x), (2000, len(x)))], axis=1)
    return {"median_ms_per_step": round(float(np.median(x)), 5),
            "ci95_ms": [round(float(np.quantile(boots, 0.025)), 5),
                        round(float(np.quantile(boots, 0.975)), 5)],
            "cv": round(float(np.std(x) / np.mean(x)), 4),
            "samples": len(per), "repetitions": reps,
            "warmup_steps": warm, "warmup_s": round(warm_s, 3)}


def run_gpu(args):
    import torch
    import mofft_amd as mf

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    g = torch.Generator(device=dev)
    g.manual_seed(20261203)

    if args.workload == "c2":
        n, batch, bits = 1 << 16, 256, 64
        fwd = mf.Plan(n, batch, 64, mf.FORWARD, in_place=True)
        bwd = mf.Plan(n, batch, 64, mf.BACKWARD, in_place=True)
        buf = torch.view_as_complex(torch.randn(batch, n, 2, dtype=torch.float64,
                                                device=dev, generator=g))
        plans = [(fwd, buf, buf), (bwd, buf, buf)]
        directions = 2
        workload = "fp64 c2c forward+backward in-place N=65536 batch=256"
        dtype = "f64"
    elif args.workload == "c1":
        n, batch, bits = 1024, 64, 32
        fwd = mf.Plan(n, batch, 32, mf.FORWARD, in_place=False)
        x = torch.rand(batch, n, 2, dtype=torch.float32, device=dev, generator=g) * 2 - 1
        x = torch.view_as_complex(x)
        y = torch.empty_like(x)
        plans = [(fwd, x, y)]
        directions = 1
        workload = "fp32 c2c forward out-of-place N=1024 batch=64"
        dtype = "f32"
    elif args.workload == "copy":
        # calibration for the HBM byte counters: one 256 MiB device copy
        nbytes = 256 << 20
        gbs = measure_copy(torch, dev, stream, nbytes)
        print(json.dumps({"workload": "copy calibration", "bytes_read": nbytes,
                          "bytes_written": nbytes, "copy_gbs": gbs}), flush=True)
        return None, None
    elif args.workload == "c4":
        print(json.dumps(run_batch(args)), flush=True)
        return None, None
    elif args.workload == "c5":
        print(json.dumps(run_serve(args)), flush=True)
        return None, None
    else:
        run_sweep(args)
        return None, None

    desc = plans[0][0].description()
    passes = desc.pass_count
    chunks = math.ceil(batch / desc.chunk_transforms)
    launches_per_step = sum(p.description().pass_count * chunks for p, _, _ in plans)

    def step():
        for p, a, b in plans:
            p.execute_async(a, b, sh)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # latency percentiles: K steps, each between its own events
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    step_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    # throughput: another K steps back to back between two events (an event
    # between steps stops a step's first kernel from overlapping the
    # previous step's last one)
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for i in range(args.steps):
        step()
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    ms = total_ms / args.steps
    flops = directions * 5.0 * n * math.log2(n) * batch
    gflops = flops / (ms * 1e-3) / 1e9
    cb = 8 if bits == 32 else 16
    alg_bytes_step = directions * 2.0 * n * batch * cb
    alg_per_launch = alg_bytes_step / launches_per_step
    # per-kernel launch durations: the same executions with a device
    # timestamp between consecutive pass launches (HIP events on the launch
    # stream); the dominant kernel is the pass with the longest launches
    per_pass = [0.0] * passes
    for _ in range(args.steps):
        for p, a, b in plans:
            for i, sec in enumerate(p.execute_timed(a, b, sh)):
                per_pass[i] += sec
    launches_per_pass = args.steps * len(plans) * chunks
    pass_us = [1e6 * v / launches_per_pass for v in per_pass]
    dom = max(range(passes), key=lambda i: pass_us[i])
    kinfo = mf.kernels()[desc.pass_kernel[dom]]
    kernel_key = (f"f{bits}_L{kinfo.length}_c{kinfo.columns_per_block}_"
                  f"{'rows' if kinfo.rows else 'cols'}")
    # The events between launches add their own gap to every launch they
    # bracket (the same kernel runs 4-8 % shorter back to back, as in the
    # timed region and in a kernel trace of it).  The dominant launch's
    # duration in the timed region is its share of the timed step, in
    # proportion to the bracketed durations of the step's launches.
    step_launch_us = sum(pass_us) * chunks * len(plans)
    launch_us = pass_us[dom] * (ms * 1e3) / step_launch_us
    achieved = alg_per_launch / (launch_us * 1e-6) / 1e9
    traffic = load_traffic(kernel_key)
    copy_gbs = measure_copy(torch, dev, stream, 256 << 20)
    proto = protocol_samples(torch, stream, step, args.samples)
    proto["gflops_at_median"] = round(flops / (proto["median_ms_per_step"] * 1e-3) / 1e9, 2)
    line = {
        "metric": "batched C2C FFT GFLOP/s (5N log2N) + % HBM roofline; p50/p99 latency",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "p50_ms": round(percentile(step_ms, 0.5), 4),
        "p99_ms": round(percentile(step_ms, 0.99), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded on device)",
        "config": {"workload": workload, "length": n, "batch": batch,
                   "precision_bits": bits, "passes": passes,
                   "pass_lengths": desc.passes, "chunk_transforms": desc.chunk_transforms,
                   "radices": desc.chain, "parallelism": "single GPU"},
        "hbm_gbs": round(alg_bytes_step / (ms * 1e-3) / 1e9, 1),
        "copy_gbs": copy_gbs,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": kernel_key,
                     "kernel_template": f"fft_pass_kernel<{'float' if bits == 32 else 'double'}, "
                                        f"{kinfo.length}, {kinfo.threads_per_column}, "
                                        f"{kinfo.columns_per_block}, {'true' if kinfo.rows else 'false'}, "
                                        + ", ".join(str(r) for r in kinfo.stages[:kinfo.stage_count]) + ">",
                     "launches_per_step": launches_per_step,
                     "avg_launch_us": round(launch_us, 3),
                     "event_launch_us": [round(v, 3) for v in pass_us]},
        "wall_s": round(wall, 3),
        "protocol": proto,
    }
    if args.workload == "c1":
        # the same execution replayed from a captured HIP graph
        p0, a0, b0 = plans[0]
        with p0.capture(a0, b0) as gr:
            for _ in range(args.warmup):
                gr.launch(sh)
            torch.cuda.synchronize()
            gs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(args.steps)]
            for e0, e1 in gs:
                e0.record(stream)
                gr.launch(sh)
                e1.record(stream)
            torch.cuda.synchronize()
            g_ms = [e0.elapsed_time(e1) for e0, e1 in gs]
        line["graph_p50_ms"] = round(percentile(g_ms, 0.5), 4)
        line["graph_p99_ms"] = round(percentile(g_ms, 0.99), 4)
    for p, _, _ in plans:
        p.destroy()
    return line, (n, bits)


def run_batch(args):
    """Uneven batch in one persistent launch (mofft_batch_*)."""
    import numpy as np
    import torch
    import mofft_amd as mf
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    g = torch.Generator(device=dev)
    g.manual_seed(4096)
    rng = np.random.default_rng(2026)
    count = args.batch_count
    ka, kb = (int(v) for v in args.batch_log2.split(","))
    lengths = [1 << int(k) for k in rng.integers(ka, kb + 1, count)]
    dirs = [mf.FORWARD if d else mf.BACKWARD for d in rng.integers(0, 2, count)]
    inplace = [bool(v) for v in rng.integers(0, 2, count)]
    bits = args.batch_bits
    fdt = torch.float32 if bits == 32 else torch.float64
    ins = [torch.view_as_complex(torch.randn(n, 2, device=dev, generator=g, dtype=fdt))
           for n in lengths]
    outs = [a if ip else torch.empty_like(a) for a, ip in zip(ins, inplace)]
    bp = mf.BatchPlan(lengths, dirs, ins, outs, bits, graph=True, grid=args.batch_grid,
                      owned_tiles=args.owned_tiles, item_tiles=args.item_tiles,
                      share_all=args.share_all)
    d = bp.description()

    def timed(fn, reps):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        es = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for a, b in es:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b in es]

    for _ in range(args.warmup):
        bp.execute_async(sh)
    t = timed(lambda: bp.execute_async(sh), args.steps)
    bp.check()
    ms = sum(t) / len(t)
    # the same transforms, one single-transform plan execution each
    plans = {}
    for n, dr, ip in set(zip(lengths, dirs, inplace)):
        plans[(n, dr, ip)] = mf.Plan(n, 1, bits, dr, in_place=ip)

    def each():
        for i, n in enumerate(lengths):
            plans[(n, dirs[i], inplace[i])].execute_async(ins[i], outs[i], sh)
    te = timed(each, 3)
    ms_each = sum(te) / len(te)
    gbs = d.algorithmic_bytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "uneven batch GFLOP/s (5N log2N) in one persistent launch",
        "value": round(d.flops / (ms * 1e-3) / 1e9, 2), "unit": "GFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "p50_ms": round(percentile(t, 0.5), 4),
        "p99_ms": round(percentile(t, 0.99), 4), "higher_is_better": True,
        "dtype": f"f{bits}", "data": "synthetic (seeded on device)",
        "config": {"workload": f"{count} fp{bits} transforms N=2^U{{{ka}..{kb}}} 50/50 fwd/bwd "
                               "mixed placement, one persistent launch (hipGraph)",
                   "jobs": count, "distinct_lengths": d.distinct_lengths,
                   "multi_pass_jobs": d.multi_pass_jobs, "shared_jobs": d.shared_jobs,
                   "owned_tiles": d.owned_tiles, "item_tiles": d.item_tiles,
                   "items": d.item_count,
                   "segments": d.segment_count, "grid": d.grid,
                   "blocks_per_cu": d.blocks_per_cu, "graph": bool(d.graph)},
        "hbm_gbs": round(gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)},
        "per_transform_launches_ms": round(ms_each, 4),
        "speedup_vs_per_transform": round(ms_each / ms, 2),
    }
    bp.destroy()
    for p in plans.values():
        p.destroy()
    return line


def run_serve(args):
    """Serving loop: Poisson arrivals, N = 2^U{8..24}, batch U{1..64}, 80%
    FP32 / 20% FP64, inputs generated on the device from per-request seeds,
    admission against a device-memory budget (mofft_server_run_trace)."""
    import numpy as np
    import mofft_amd as mf
    rng = np.random.default_rng(2000)
    count = int(args.serve_rate * args.serve_seconds)
    ks = rng.integers(8, 25, count)
    bs = rng.integers(1, 65, count)
    prec = np.where(rng.random(count) < 0.8, 32, 64)
    dirs = np.where(rng.random(count) < 0.5, mf.FORWARD, mf.BACKWARD)
    seeds = rng.integers(1, 2**62, count)
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, count))
    reqs = [mf.Request(1 << int(k), int(b), int(p), int(d), None, int(sd), i)
            for i, (k, b, p, d, sd) in enumerate(zip(ks, bs, prec, dirs, seeds))]
    budget = int(args.serve_budget_gb * (1 << 30))
    with mf.Server(budget, streams=args.serve_streams) as srv:
        # resident plans for every (length, precision, direction) first
        warm = [mf.Request(1 << k, 1, p, d, None, 1, 0)
                for k in range(8, 25) for p in (32, 64) for d in (mf.FORWARD, mf.BACKWARD)]
        srv.run_trace(warm, [0.0] * len(warm))
        t0 = time.perf_counter()
        cs = srv.run_trace(reqs, arrivals.tolist())
        wall = time.perf_counter() - t0
        st = srv.stats()
    ok = [c for c in cs if c.status == mf.SUCCESS]
    lat = [1e3 * (c.done_seconds - c.submit_seconds) for c in ok]
    first = min(c.submit_seconds for c in ok)
    last = max(c.done_seconds for c in ok)
    flops = sum(5.0 * r.length * math.log2(r.length) * r.batch
                for r, c in zip(reqs, cs) if c.status == mf.SUCCESS) - 0.0
    span = last - first
    alg = sum(2.0 * r.length * r.batch * (8 if r.precision_bits == 32 else 16)
              for r, c in zip(reqs, cs) if c.status == mf.SUCCESS)
    line = {
        "metric": "served C2C FFT GFLOP/s (5N log2N); p50/p99 request latency",
        "value": round(flops / span / 1e9, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32/f64 (80/20)",
        "data": "synthetic (generated on device from per-request seeds)",
        "config": {"workload": f"Poisson {args.serve_rate:g} req/s for {args.serve_seconds:g} s, "
                               "N=2^U{8..24}, batch U{1..64}, 80% fp32 / 20% fp64",
                   "requests": count, "budget_gb": args.serve_budget_gb,
                   "streams": args.serve_streams},
        "p50_ms": round(percentile(lat, 0.5), 3),
        "p99_ms": round(percentile(lat, 0.99), 3),
        "completed": len(ok), "rejected": count - len(ok),
        "completed_per_s": round(len(ok) / span, 1),
        "offered_per_s": args.serve_rate,
        "hbm_gbs": round(alg / span / 1e9, 1),
        "peak_bytes_in_use_gb": round(st.peak_bytes_in_use / 2**30, 2),
        "peak_inflight": st.peak_inflight,
        "resident_plans": st.resident_plans,
        "wall_s": round(wall, 2),
    }
    return line


def mfma32_chains(mf):
    """Plan chains (outermost radix first) of the whole-transform
    matrix-core kernels with radix-32 sub-stages, by length."""
    out = {}
    for k in mf.kernels():
        st = list(k.stages[:k.stage_count])
        if k.variant == mf.MFMA_BANK and k.rows and k.precision_bits == 32 and 32 in st:
            out[k.length] = st[::-1]
    return out


def _sweep_plan(mf, form, n, batch, direction, chains):
    if form != "mfma32":
        return mf.Plan(n, batch, 32, direction, in_place=False, mfma=form == "mfma")
    p = mf.Plan(n, batch, 32, direction, in_place=False, radices=chains[n])
    if mf.lib().mofft_plan_set_kernel_bank_f32(p.handle, mf.MFMA_BANK) != 0:
        raise RuntimeError(f"kernel bank {mf.MFMA_BANK} refused for N={n}")
    p.set_blocked_transpose_mask(0)
    return p


def run_sweep(args, emit=True):
    """FP32 forward out-of-place, N = 2^k, batch = 2^24 / N: the plain FP32
    form, the matrix-core split-precision form (radix-16 sub-stages) and,
    for the lengths with such a kernel, the radix-32 matrix-core chain (one
    line per N and form).  `rt_err` is the forward->backward round-trip
    relative L2 error measured on the device (the oracle errors per N come
    from the tests)."""
    import torch
    import mofft_amd as mf
    chains = mfma32_chains(mf)
    dev = torch.device("cuda:0")
    sh = torch.cuda.current_stream(dev).cuda_stream
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    total = 1 << 24
    x = torch.view_as_complex(torch.rand(total, 2, device=dev, generator=g) * 2 - 1)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    rows = []
    for k in range(4, args.sweep_max + 1):
        n = 1 << k
        batch = total // n
        for form in ("plain", "mfma") + (("mfma32",) if n in chains else ()):
            p = _sweep_plan(mf, form, n, batch, mf.FORWARD, chains)
            q = _sweep_plan(mf, form, n, batch, mf.BACKWARD, chains)
            for _ in range(10):
                p.execute_async(x, y, sh)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 30
            e0.record()
            for _ in range(reps):
                p.execute_async(x, y, sh)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            q.execute_async(y, z, sh)
            torch.cuda.synchronize()
            rt = (torch.linalg.vector_norm(z / n - x) / torch.linalg.vector_norm(x)).item()
            gf = 5.0 * n * k * batch / (ms * 1e-3) / 1e9
            gbs = 2.0 * total * 8 / (ms * 1e-3) / 1e9
            row = {"workload": "fp32 forward out-of-place sweep", "form": form,
                   "length": n, "batch": batch, "passes": p.description().passes,
                   "chain": p.description().chain,
                   "ms": round(ms, 4), "gflops": round(gf, 1),
                   "hbm_gbs": round(gbs, 1),
                   "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                   "rt_err": float(f"{rt:.3e}")}
            rows.append(row)
            if emit:
                print(json.dumps(row), flush=True)
            p.destroy()
            q.destroy()
    return rows


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (--gpus {args.gpus} requested); "
              "batched transforms would shard across GPUs with no exchange",
              file=sys.stderr)
        sys.exit(2)
    line, shape = run_gpu(args)
    if line is None:
        return
    if args.workload == "c2" and not args.no_configs:
        line["configs"] = run_configs(args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_port(shape[0], shape[1], args.cpu_seconds)
        line["cpu_pocketfft"] = cpu_pocketfft(shape[0], 16, shape[1], min(5.0, args.cpu_seconds))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def run_configs(args):
    """The other configurations, measured in the same run and summarised
    for the headline line (each progress line goes to stderr)."""
    import argparse as ap
    out = {}
    c1 = ap.Namespace(**vars(args))
    c1.workload = "c1"
    line, _ = run_gpu(c1)
    out["c1"] = {k: line[k] for k in ("value", "unit", "ms_per_step", "p50_ms", "p99_ms",
                                      "graph_p50_ms", "graph_p99_ms", "hbm_gbs")}
    out["c1"]["protocol"] = line["protocol"]
    print("configs: c1 done", file=sys.stderr, flush=True)
    rows = run_sweep(args, emit=False)
    sweep = {}
    for r in rows:
        e = sweep.setdefault(r["length"], {"passes": r["passes"]})
        if r["form"] == "mfma32":
            e["mfma32_chain"] = r["chain"]
        e[f"{r['form']}_ms"] = r["ms"]
        e[f"{r['form']}_hbm_frac"] = r["hbm_frac"]
        e[f"{r['form']}_rt_err"] = r["rt_err"]
    out["c3"] = {"workload": "fp32 forward out-of-place, N=2^k, batch 2^24/N",
                 "per_length": {str(k): v for k, v in sorted(sweep.items())},
                 "hbm_frac_ge_0.70": [n for n, v in sorted(sweep.items())
                                      if v.get("plain_hbm_frac", 0) >= 0.70]}
    print("configs: c3 done", file=sys.stderr, flush=True)
    c4 = run_batch(args)
    out["c4"] = {k: c4[k] for k in ("value", "unit", "ms_per_step", "p50_ms", "p99_ms",
                                    "hbm_gbs", "speedup_vs_per_transform")}
    out["c4"]["roofline_frac"] = c4["roofline"]["frac"]
    print("configs: c4 done", file=sys.stderr, flush=True)
    c5a = ap.Namespace(**vars(args))
    c5a.serve_seconds = args.configs_serve_seconds
    c5 = run_serve(c5a)
    out["c5"] = {k: c5[k] for k in ("value", "unit", "p50_ms", "p99_ms", "completed",
                                    "rejected", "completed_per_s", "offered_per_s",
                                    "hbm_gbs", "peak_bytes_in_use_gb", "wall_s")}
    out["c5"]["window_s"] = args.configs_serve_seconds
    print("configs: c5 done", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
