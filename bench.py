#!/usr/bin/env python3
"""Benchmarks of the HighMap grid-erosion hot path on one MI355X.

Default (the headline): vpipes + thermal on a 4096 x 4096 grid.  One step =
10 hydraulic_vpipes iterations followed by one thermal step (talus 1/4096,
then extrapolate_borders) on one continuous device-resident state, with the
1024^2-workload parameters (water_height 1e-3 with volume top-up, dt 0.5,
k_capacity 0.1, k_erode 0.05, k_depose 0.05, evaporation 0.01).  The default
50 timed steps are one full 500-iteration job.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--storage f32|bf16|f16]
    python bench.py --workload batch|serve|precision|thermal256|vpipes1024|flow|snow|tiles|copy|ceiling|noise

Each run prints one JSON line.  For the headline, `value` is hydraulic
cell-iterations per second (nx*ny*iterations / wall time of the timed steps,
thermal included in the time) with the state already resident in HBM;
`roofline` is measured on the fused vpipes kernel alone with HIP events on
the library's stream; `access_pattern_ceiling` times the same kernel's
memory operations alone (k_vp_pattern); `cpu_baseline` times the FP32 CPU
restatement (oracle/cpu_bench) on a bounded sample of the same workload.
Unless --no-extras: the kernel's rate with BF16 and FP16 state, noise_fbm's
rate (PERLIN, the input generator), the
thermal 256^2 x 50 and vpipes 1024^2 x 200 workloads (both variants, error
against the FP64 oracle), the 96-map batch and the served Poisson stream.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(water_height=1e-3, evap_rate=0.01, dt=0.5, k_capacity=0.1, k_erode=0.05,
                k_depose=0.05)
STORAGE = {"f32": 0, "bf16": 1, "f16": 2}
# Algorithmic bytes per hydraulic cell-iteration: every persistent plane read
# once and written once (z FP32; water, sediment and 4 fluxes in storage).
ALG_BYTES = {"f32": 8 + 6 * 4 * 2, "bf16": 8 + 6 * 2 * 2, "f16": 8 + 6 * 2 * 2}


def cpu_threads():
    """Every host CPU this process may run on."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        return max(1, os.cpu_count() or 1)


def cpu_bench(workload, nx, ny, iters, reps=3, threads=None):
    exe = os.path.join(ROOT, "oracle", "cpu_bench")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    threads = threads or cpu_threads()
    out = subprocess.run([exe, workload, str(nx), str(ny), str(iters), str(threads), str(reps)],
                         capture_output=True, text=True, check=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def cpu_model():
    """Model name, vendor family/model/stepping, clock and CPU counts from
    /proc/cpuinfo."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    info.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    parts = [info.get("model name", "unknown")]
    fam = ", ".join("%s %s" % (k, info[k]) for k in ("vendor_id", "cpu family", "model",
                                                      "stepping") if k in info)
    if fam:
        parts.append("(%s)" % fam)
    if "cpu MHz" in info:
        parts.append("%s MHz" % info["cpu MHz"])
    parts.append("%d logical CPUs, %d in this process's affinity set"
                 % (os.cpu_count() or 0, cpu_threads()))
    return " ".join(parts)


def cpu_baseline(n, iters=20, reps=5):
    r = cpu_bench("vpipes_thermal", n, n, iters, reps)
    # one thread on a shorter sample of the same grid (no thermal step falls in it)
    r1 = cpu_bench("vpipes_thermal", n, n, 4, 1, 1)
    return {
        "value": round(r["gcell_iter_s"], 5),
        "unit": "Gcell-iter/s",
        "cores": r["threads"],
        "kind": "port",
        "sample": "%dx%d fBm seed 2, %d vpipes iterations with a thermal step every 10, "
                  "FP32 CPU restatement of the reference passes built -O3 -Ofast -ffast-math "
                  "(serial FP32 volume sum as the reference's), OpenMP %d threads, median "
                  "of %d runs after 1 warm-up (%.2f s each)"
                  % (n, n, iters, r["threads"], reps, r["median_s"]),
        "value_1_thread": round(r1["gcell_iter_s"], 5),
        "sample_1_thread": "same grid, 4 vpipes iterations, 1 thread, 1 run after 1 warm-up "
                           "(%.2f s)" % r1["median_s"],
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
    }


def load_traffic(key, chunk_rows, prefetch):
    """HBM bytes per launch of the dominant kernel, from the rocprof
    FETCH_SIZE/WRITE_SIZE runs in profiles/traffic.json for this workload and
    launch shape (traffic depends on the shape: halo and pre-roll rows), or
    None when that shape was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        e = t[key]
        sh = e["shapes"]["%dx%d" % (chunk_rows, prefetch)]
        return int(sh["fetch_kib"] * e["fetch_scale"] * 1024 + sh["write_kib"] * e["write_scale"] * 1024)
    except (OSError, KeyError, ValueError):
        return None


def _gpu():
    import torch

    import highmap_amd as hm

    hm.init_gpu(0)
    return torch, hm, hm._default_ctx


def device_terrain(torch, hm, ctx, nx, ny, seed):
    """remap(noise_fbm(PERLIN, kw 4, seed)) to [0,1], generated on the GPU."""
    z = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    hm._check(hm.lib().hmg_noise_fbm_dev(ctx.handle, z.data_ptr(), nx, ny, hm.NoiseType.PERLIN,
                                         4.0, 4.0, seed, 8, 0.7, 0.5, 2.0, None, None, None,
                                         None, None))
    hm.remap(z.data_ptr(), nx * ny, 0.0, 1.0)
    torch.cuda.synchronize()
    return z


# ------------------------------------------------------------- headline


def run_headline(args):
    torch, hm, ctx = _gpu()
    n = args.n
    stream = torch.cuda.ExternalStream(ctx.stream)
    z = device_terrain(torch, hm, ctx, n, n, 2)
    if not args.chunk_rows and args.mode == "fused":
        # the online launch-shape probes (iterations 20-55 of a session) run
        # on a throw-away session of the same input; the measured session
        # then starts from the process-wide cache and probes nothing
        pre = hm.Session(z, storage=STORAGE[args.storage], mode=args.mode, **WORKLOAD)
        pre.vpipes(60)
        pre.close()
    sess = hm.Session(z, storage=STORAGE[args.storage], mode=args.mode, **WORKLOAD)
    if args.chunk_rows:
        sess.set_tuning(args.chunk_rows, args.prefetch)
    talus = 1.0 / n

    def step():
        sess.vpipes(10)
        sess.thermal(talus, 1)

    for _ in range(args.warmup):
        step()
    sess.sync()
    torch.cuda.synchronize()
    probing_in_timed_steps = sess.tuning()[2]

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    e1.record(stream)
    e1.synchronize()
    sess.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    elapsed = max(wall, e0.elapsed_time(e1) / 1e3)
    iters = 10 * args.steps
    value = n * n * iters / elapsed / 1e9

    # dominant kernel alone: consecutive vpipes launches between events
    m = 50
    e0.record(stream)
    sess.vpipes(m)
    e1.record(stream)
    e1.synchronize()
    launch_s = e0.elapsed_time(e1) / 1e3 / m
    alg = ALG_BYTES[args.storage] * n * n
    achieved = alg / launch_s / 1e9
    # the thermal step alone (one streaming launch with the borders folded
    # in), on the same state
    m_th = 20
    e0.record(stream)
    for _ in range(m_th):
        sess.thermal(talus, 1)
    e1.record(stream)
    e1.synchronize()
    thermal_us = e0.elapsed_time(e1) * 1e3 / m_th
    chunk_rows, prefetch, _ = sess.tuning()
    max_backtrace = sess.max_backtrace()
    sess.close()
    copy_gbs, copy_variant = measured_copy_gbs(torch)
    pattern_us = (vp_pattern_us(torch, n, n, args.storage, chunk_rows)
                  if args.mode == "fused" else None)

    kernel = (("k_vp_stream_ga" if prefetch == 2 else "k_vp_stream") if args.mode == "fused"
              else "k_vp_* (plain passes)")
    out = {
        "metric": "Gcell-iter/s vpipes+thermal 4096^2",
        "value": round(value, 3),
        "unit": "Gcell-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else "f32 math, %s storage" % args.storage,
        "data": "synthetic: remap(noise_fbm PERLIN, kw 4, seed 2) to [0,1], generated on the GPU",
        "config": {
            "workload": "vpipes+thermal %dx%d: per step 10 hydraulic_vpipes iterations then 1 "
                        "thermal step (talus 1/%d) on one device-resident state" % (n, n, n),
            "nx": n,
            "ny": n,
            "hydraulic_iterations_per_step": 10,
            "params": WORKLOAD,
            "storage": args.storage,
            "mode": args.mode,
            "launch_shape": {"chunk_rows": chunk_rows, "prefetch": prefetch,
                             "source": "set" if args.chunk_rows else
                             "online tuning, cached from a 60-iteration run of the same input",
                             "probes_in_timed_steps": probing_in_timed_steps},
        },
        # how far (cells) the last state's sediment backtraces reach: the
        # gathers of this workload land hundreds of cells away
        "max_backtrace_cells": round(max_backtrace, 2),
        "thermal_step_us": round(thermal_us, 2),
        "pct_hbm_peak_step": round(100.0 * (ALG_BYTES[args.storage] * n * n * iters +
                                            8 * n * n * args.steps) / elapsed / 1e9
                                   / HBM_PEAK_GBS, 2),
        "roofline": {
            "kernel": kernel,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic("vpipes_%s_%s_%d" % (args.mode, args.storage, n),
                                    chunk_rows, prefetch),
            "alg_bytes_per_launch": alg,
            "avg_launch_us": round(launch_s * 1e6, 2),
        },
    }
    # the same traffic against what a plain device copy (read + write) reaches
    # on this GPU in this process: the practical ceiling for this mix
    tr = out["roofline"]["traffic"]
    out["measured_copy"] = {"copy_hbm_gbs": round(copy_gbs, 1),
                            "copy_kernel": dict(copy_variant, source="hbm_probe.hip copy, "
                                                "fastest variant, 2 GiB buffers"),
                            "traffic_gbs": round(tr / launch_s / 1e9, 1) if tr else None,
                            "traffic_frac_of_copy": round(tr / launch_s / 1e9 / copy_gbs, 4)
                            if tr else None,
                            "alg_frac_of_copy": round(achieved / copy_gbs, 4)}
    if pattern_us:
        # the same memory operations and decomposition without the erosion
        # arithmetic (k_vp_pattern): the ceiling of this access pattern
        out["access_pattern_ceiling"] = {
            "kernel": "k_vp_pattern (hbm_probe.hip)", "chunk_rows": chunk_rows,
            "avg_launch_us": round(pattern_us, 2),
            "alg_gbs": round(alg / (pattern_us * 1e-6) / 1e9, 1),
            "alg_frac_of_peak": round(alg / (pattern_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_frac_of_pattern": round(pattern_us / (launch_s * 1e6), 4)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n)
    if not args.no_extras:
        out["storage_variants"] = storage_variants(args)
        out["noise_fbm_perlin"] = noise_rates(args)["1"]
        # rates only (device-generated inputs); their bit-exactness against
        # the oracle at these sizes is the GPU tests'
        out["config1_thermal256"] = run_thermal256(args, cpu=False, big_grid=False, check=False)
        out["config2_vpipes1024"] = run_vpipes1024(args, cpu=False, check=False)["results"]
        out["batch"] = batch_summary(args)
        out["served"] = served_summary(args)
    return out


def noise_rates(args, types=(1,)):
    """hmap::gpu::noise_fbm on the device at n x n (8 octaves, wavenumber
    4, seed 2, no control or warping planes): HIP events around the call on
    the library's stream, best of 3; Gcell/s per noise type.  VALU-bound
    (hash sines and gradient sincos), reported as a rate, not against HBM."""
    torch, hm, ctx = _gpu()
    n = args.n
    stream = torch.cuda.ExternalStream(ctx.stream)
    z = torch.empty((n, n), dtype=torch.float32, device="cuda")
    out = {}
    for t in types:
        best = None
        for _ in range(3):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            hm._check(hm.lib().hmg_noise_fbm_dev(ctx.handle, z.data_ptr(), n, n, t, 4.0, 4.0, 2,
                                                 8, 0.7, 0.5, 2.0, None, None, None, None, None))
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        out[str(t)] = {"ms": round(best, 4), "gcell_s": round(n * n / (best / 1e3) / 1e9, 2)}
    del z
    return out


def run_noise(args):
    """noise_fbm at n x n for every implemented hmap::NoiseType."""
    r = noise_rates(args, types=(1, 2, 3, 4, 6, 7, 9, 10))
    return {"metric": "Gcell/s noise_fbm %dx%d, 8 octaves" % (args.n, args.n), "unit": "Gcell/s",
            "n_gpus": 1, "value": r["1"]["gcell_s"], "by_type": r}


def storage_variants(args):
    """The headline's fused kernel with BF16 and FP16 state (FP32
    arithmetic): per-launch time at the tuned shape (a 60-iteration session
    of the same input probes the shapes first), HIP events around 50
    consecutive launches on the library's stream after 300 iterations
    (with the headline's thermal steps)."""
    torch, hm, ctx = _gpu()
    n = args.n
    stream = torch.cuda.ExternalStream(ctx.stream)
    z = device_terrain(torch, hm, ctx, n, n, 2)
    out = {}
    for storage in ("bf16", "f16"):
        pre = hm.Session(z, storage=STORAGE[storage], **WORKLOAD)
        pre.vpipes(60)
        pre.close()
        sess = hm.Session(z, storage=STORAGE[storage], **WORKLOAD)
        # past the water's spreading from rest (wide sediment backtraces make
        # the first ~150 iterations slower), on the headline's state: 10
        # hydraulic iterations then a thermal step, 30 times
        for _ in range(30):
            sess.vpipes(10)
            sess.thermal(1.0 / n, 1)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        m = 50
        e0.record(stream)
        sess.vpipes(m)
        e1.record(stream)
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / m
        chunk_rows, prefetch, _ = sess.tuning()
        sess.close()
        alg = ALG_BYTES[storage] * n * n
        out[storage] = {"avg_launch_us": round(us, 2),
                        "kernel_gcell_iter_s": round(n * n / (us * 1e-6) / 1e9, 2),
                        "alg_bytes_per_launch": alg,
                        "achieved_gbs": round(alg / (us * 1e-6) / 1e9, 1),
                        "frac": round(alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                        "launch_shape": {"chunk_rows": chunk_rows, "prefetch": prefetch}}
    del z
    torch.cuda.empty_cache()
    return out


def batch_summary(args):
    """The 96-map batch (config of run_batch) in a few numbers."""
    r = run_batch(args)
    return {"workload": r["config"]["workload"],
            "gcell_iter_s": r["value"],
            "per_map_plain_gcell_iter_s": r["per_map_plain_gcell_iter_s"],
            "per_map_fused_gcell_iter_s": r["per_map_fused_gcell_iter_s"],
            "speedup_vs_per_map_plain": r["speedup_vs_per_map_plain"],
            "graph_gcell_iter_s": r["graph_gcell_iter_s"],
            "persistent_gcell_iter_s": r["persistent_gcell_iter_s"],
            "launch_shape": r["config"]["tuning"], "storage": args.storage}


def served_summary(args, requests=2000, sub_requests=500, sub_rate=40.0):
    """The served Poisson stream (2000 requests at 200/s, BF16 state, 3 x
    4096^2 active cells at most) against sequential plain calls, and a
    sub-capacity load point (the first 500 requests of the same generator
    at 40/s), all replayed in real time."""
    from highmap_amd import Storage, VpipesParams, serve, workloads

    import highmap_amd as hm

    hm.init_gpu(0)
    params = VpipesParams.make(**WORKLOAD)
    cap = 3 * 4096 * 4096
    serve.warm_up(Storage.BF16, params)
    m = serve.measure(hm, Storage.BF16, params, workloads.serve_requests(n=requests), cap)
    m.pop("_main")
    low = serve.measure(hm, Storage.BF16, params,
                        workloads.serve_requests(n=sub_requests, rate=sub_rate), cap,
                        sequential=False)
    low.pop("_main")
    return {"workload": "Poisson stream, rate 200/s (seed 12345), %d requests: side "
                        "log-uniform 256-4096, 50-500 iterations, 70%% vpipes+thermal / 30%% "
                        "thermal-only, BF16 state" % requests,
            "p50_s": m["p50_latency_s"], "p99_s": m["p99_latency_s"],
            "gcell_iter_s": m["achieved_gcell_iter_s"],
            "offered_gcell_iter_s": m["offered_gcell_iter_s"],
            "max_active_cells": cap,
            "sequential_plain": m["sequential_plain"],
            "alone_equals_mix": m["alone_equals_mix"],
            "sub_capacity": {"requests": sub_requests, "rate_per_s": sub_rate,
                             "offered_gcell_iter_s": low["offered_gcell_iter_s"],
                             "gcell_iter_s": low["achieved_gcell_iter_s"],
                             "p50_s": low["p50_latency_s"], "p99_s": low["p99_latency_s"]}}


# ---------------------------------------------------------------- batch


FORMS = {0: "stream", 1: "graph", 2: "persistent"}


def run_batch(args):
    """96 maps (seeds 100..195, shapes and iteration counts from the batch
    manifest) with the autotuned launch shape and form: one batched launch
    per iteration captured in a hipGraph, or one persistent cooperative
    launch with a grid barrier between iterations (both are timed; `value`
    is the form the autotuner picked), against the same maps run one after
    another with per-map plain launches (and per-map fused launches)."""
    from highmap_amd import workloads

    torch, hm, ctx = _gpu()
    maps = workloads.batch_manifest()
    if args.maps:
        maps = maps[: args.maps]
    zs = [device_terrain(torch, hm, ctx, m["nx"], m["ny"], m["seed"]).cpu().numpy()
          for m in maps]
    cell_iters = sum(m["nx"] * m["ny"] * m["iterations"] for m in maps)
    st = STORAGE[args.storage]

    # warm-up and autotune (shape, then graph against persistent launch) on
    # a throw-away copy: 20 ordinary iterations, then the fixed and the
    # occupancy-fitted shapes (up to 24) and the 2 forms, 2 probe
    # iterations each
    b = hm.Batch(storage=st, **WORKLOAD)
    for z, m in zip(zs, maps):
        b.add(z, 100)
    tuned = b.autotune(probe=2)
    form = b.tuned_form()
    b.close()

    def batched(f):
        b = hm.Batch(storage=st, **WORKLOAD)
        for z, m in zip(zs, maps):
            b.add(z, m["iterations"])
        b.set_tuning(*tuned)
        t0 = time.perf_counter()
        dev_ms = b.run(form=f)
        dt = time.perf_counter() - t0
        b.close()
        return dt, dev_ms

    graph_s, graph_ms = batched(hm.Batch.GRAPH)
    persist_s, persist_ms = batched(hm.Batch.PERSISTENT)
    batched_s, dev_ms = (persist_s, persist_ms) if form == hm.Batch.PERSISTENT else (graph_s,
                                                                                    graph_ms)

    def per_map(mode):
        sess = []
        for z in zs:
            sess.append(hm.Session(z, storage=st, mode=mode, **WORKLOAD))
        ctx.sync()
        t0 = time.perf_counter()
        for s, m in zip(sess, maps):
            s.vpipes(m["iterations"])
        ctx.sync()
        dt = time.perf_counter() - t0
        for s in sess:
            s.close()
        return dt

    plain_s = per_map("plain")
    fused_s = per_map("fused")
    return {
        "metric": "Gcell-iter/s batched vpipes, 96 uneven maps",
        "value": round(cell_iters / batched_s / 1e9, 3),
        "unit": "Gcell-iter/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32" if args.storage == "f32" else "f32 math, %s storage" % args.storage,
        "data": "synthetic: remap(noise_fbm PERLIN, kw 4, seed s) per map, s = 100..195",
        "config": {"workload": "batch manifest: %d maps, %d cells, %d cell-iterations"
                               % (len(maps), sum(m["nx"] * m["ny"] for m in maps), cell_iters),
                   "params": WORKLOAD, "tuning": {"chunk_rows": tuned[0], "prefetch": tuned[1],
                                                  "form": FORMS[form]}},
        "batched_s": round(batched_s, 4),
        "batched_device_ms": round(dev_ms, 2),
        "graph_gcell_iter_s": round(cell_iters / graph_s / 1e9, 3),
        "persistent_gcell_iter_s": round(cell_iters / persist_s / 1e9, 3),
        "per_map_plain_s": round(plain_s, 4),
        "per_map_plain_gcell_iter_s": round(cell_iters / plain_s / 1e9, 3),
        "per_map_fused_s": round(fused_s, 4),
        "per_map_fused_gcell_iter_s": round(cell_iters / fused_s / 1e9, 3),
        "speedup_vs_per_map_plain": round(plain_s / batched_s, 2),
    }


def run_batch_forms(args):
    """Small maps in the three ways to run them: one batch as a hipGraph of
    per-iteration launches, the same batch as one persistent cooperative
    launch (grid barrier between iterations), and the maps as sessions one
    after another.  Batches use the shape the autotuner picks on a
    throw-away copy; rates from the device time of each batch."""
    torch, hm, ctx = _gpu()
    rows = []
    for n, iters, nmaps in ((1024, 200, 1), (512, 200, 4), (256, 300, 16), (2048, 100, 1)):
        zs = [device_terrain(torch, hm, ctx, n, n, 2 + k).cpu().numpy() for k in range(nmaps)]
        b = hm.Batch(**WORKLOAD)
        for z in zs:
            b.add(z, 40)
        tuned = b.autotune(probe=2)
        form = b.tuned_form()
        b.close()
        cells = n * n * nmaps * iters
        r = {"side": n, "maps": nmaps, "iterations": iters,
             "tuning": {"chunk_rows": tuned[0], "prefetch": tuned[1], "form": FORMS[form]}}
        for f in (hm.Batch.GRAPH, hm.Batch.PERSISTENT):
            b = hm.Batch(**WORKLOAD)
            b.set_tuning(*tuned)
            for z in zs:
                b.add(z, iters)
            ms = b.run(form=f)
            b.close()
            r[FORMS[f] + "_gcell_iter_s"] = round(cells / (ms * 1e-3) / 1e9, 2)
            r[FORMS[f] + "_us_per_iteration"] = round(ms * 1e3 / iters, 1)
        sess = [hm.Session(z, **WORKLOAD) for z in zs]
        ctx.sync()
        t0 = time.perf_counter()
        for x in sess:
            x.vpipes(iters)
        ctx.sync()
        r["sessions_gcell_iter_s"] = round(cells / (time.perf_counter() - t0) / 1e9, 2)
        for x in sess:
            x.close()
        print("batchforms %d^2 x %d done" % (n, nmaps), file=sys.stderr, flush=True)
        rows.append(r)
    return {"metric": "Gcell-iter/s of small maps: batched graph, persistent, sessions",
            "unit": "Gcell-iter/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
            "data": "synthetic: remap(noise_fbm PERLIN, kw 4, seed 2 + k)",
            "config": {"workload": "small maps in batches", "params": WORKLOAD},
            "results": rows}


# ------------------------------------------------------------ precision


def run_precision(args):
    """FP32 / BF16 / FP16 storage of water, sediment and fluxes against the
    FP64 CPU oracle on one continuous vpipes+thermal state (thermal every 10
    iterations): max |z - z64| / range(z64) and the drift of sum(z+s) at
    checkpoints, for the workload parameters and the reference defaults."""
    import numpy as np

    from oracle import pyoracle as O

    torch, hm, ctx = _gpu()
    n, total = args.n, args.iters
    checkpoints = [c for c in (10, 20, 50, 100, 200, 500, 1000) if c < total] + [total]
    z0 = device_terrain(torch, hm, ctx, n, n, 2).cpu().numpy()
    mass0 = float(z0.astype(np.float64).sum())
    out = {"metric": "max |z - z_fp64| / range and sum(z+s) drift vs the FP64 oracle",
           "unit": "relative", "n_gpus": 1, "config": {"nx": n, "ny": n, "iterations": total,
                                                        "thermal_every": 10},
           "results": {}}
    for pname, params in (("workload", WORKLOAD), ("reference_defaults", {})):
        o = O.VpState(z0.astype(np.float64), O.vp_params(**params), dtype=np.float64)
        sess = {k: hm.Session(z0, storage=v, **params) for k, v in STORAGE.items()}
        done = 0
        rows = []
        t0 = time.perf_counter()
        for cp in checkpoints:
            while done < cp:
                step = min(10 - done % 10, cp - done)
                o.run(step)
                for s in sess.values():
                    s.vpipes(step)
                done += step
                if done % 10 == 0:
                    o.thermal_step(1.0 / n, 1)
                    for s in sess.values():
                        s.thermal(1.0 / n, 1)
            rng = float(o.z.max() - o.z.min())
            m64 = float((o.z + o.s).sum())
            row = {"iteration": cp, "fp64_mass_drift": (m64 - mass0) / mass0}
            for k, s in sess.items():
                z, d, sed, u, v = s.export()
                row[k] = {"max_abs_err_rel_range": float(np.abs(z - o.z).max() / rng),
                          "mass_vs_fp64_rel": float(((z.astype(np.float64) + sed).sum() - m64)
                                                    / m64)}
            rows.append(row)
            print("precision %s: iteration %d done" % (pname, cp), file=sys.stderr, flush=True)
        for s in sess.values():
            s.close()
        out["results"][pname] = {"params": params or "hmap::gpu::hydraulic_vpipes defaults",
                                 "checkpoints": rows,
                                 "oracle_s": round(time.perf_counter() - t0, 1)}

    # The chunked form the reference API can express, once for the record:
    # hydraulic_vpipes(iterations=10) then thermal(1), repeated, with water,
    # sediment and fluxes reset by every call.  FP32 drop-in calls (host
    # arrays) against the FP64 oracle run the same way.
    t0 = time.perf_counter()
    z32 = z0.copy()
    z64 = z0.astype(np.float64)
    p64 = O.vp_params(**WORKLOAD)
    rows = []
    for c in range(1, total // 10 + 1):
        hm.hydraulic_vpipes(z32, iterations=10, **WORKLOAD)
        hm.thermal(z32, 1.0 / n, 1)
        z64 = O.hydraulic_vpipes(z64, p64, None, 10, np.float64)[0]
        z64 = O.thermal(z64, 1.0 / n, 1, np.float64)
        if 10 * c in checkpoints:
            print("precision chunked: iteration %d done" % (10 * c), file=sys.stderr, flush=True)
            rng = float(z64.max() - z64.min())
            rows.append({"iteration": 10 * c,
                         "f32": {"max_abs_err_rel_range": float(np.abs(z32 - z64).max() / rng)}})
    out["results"]["chunked_dropin"] = {
        "params": WORKLOAD, "form": "%d x (hydraulic_vpipes(iterations=10), thermal(1))"
        % (total // 10), "checkpoints": rows, "oracle_s": round(time.perf_counter() - t0, 1)}
    return out


# ------------------------------------------------ configs 1 and 2 (drop-in)


def thermal256_checks(z0, z):
    """The 256^2 thermal result against the FP32 oracle (bit for bit) and
    the FP64 oracle (error, interior sum drift)."""
    import numpy as np

    from oracle import pyoracle as O

    o64 = O.thermal(z0.astype(np.float64), 2.0 / 256, 50, np.float64)

    def interior_sum(a):
        return float(a[1:-1, 1:-1].astype(np.float64).sum())

    s0 = interior_sum(z0)
    return {"bitexact_vs_fp32_oracle": bool(np.array_equal(z, O.thermal(z0, 2.0 / 256, 50))),
            "max_abs_err_rel_range_vs_fp64": float(np.abs(z - o64).max() / (o64.max() - o64.min())),
            # relative change of the interior sum over the 50 steps (the
            # exchange with border cells moves height across the interior's
            # edge), for the GPU result and the FP64 oracle
            "interior_sum_drift": (interior_sum(z) - s0) / s0,
            "interior_sum_drift_fp64": (interior_sum(o64) - s0) / s0}


def run_thermal(args):
    """hmap::gpu::thermal at n x n, device-resident (fBm seed 1 generated on
    the device): a uniform talus 1/n and a talus plane (uniform random in
    [0, 2/n), torch's generator, seed 5), 1 and 50 steps per call; HIP
    events around the call on the library's stream, best of 3.  GB/s counts
    the algorithmic bytes per cell-step: 8 (z read and written) with a
    scalar, 12 (and the talus read) with a plane."""
    torch, hm, ctx = _gpu()
    n = args.n
    z0 = device_terrain(torch, hm, ctx, n, n, 1)
    g = torch.Generator(device="cuda").manual_seed(5)
    plane = torch.rand((n, n), device="cuda", generator=g) * (2.0 / n)
    z = torch.empty_like(z0)
    st = torch.cuda.ExternalStream(ctx.stream)
    out = {"metric": "Gcell-iter/s thermal %dx%d (device-resident)" % (n, n), "unit": "Gcell-iter/s",
           "n_gpus": 1}
    for name, tp, ts, bpc in (("uniform", None, 1.0 / n, 8), ("plane", plane.data_ptr(), 0.0, 12)):
        for steps in (1, 50):
            best = None
            for _ in range(3):
                z.copy_(z0)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                hm._check(hm.lib().hmg_thermal_dev(ctx.handle, z.data_ptr(), n, n, tp, ts, steps,
                                                   None))
                e1.record(st)
                e1.synchronize()
                ms = e0.elapsed_time(e1)
                best = ms if best is None else min(best, ms)
            rate = n * n * steps / (best / 1e3) / 1e9
            out["%s_x%d" % (name, steps)] = {"us_per_step": round(best * 1e3 / steps, 2),
                                             "gcell_iter_s": round(rate, 2),
                                             "gbs": round(rate * bpc, 1)}
    out["value"] = out["uniform_x50"]["gcell_iter_s"]
    return out


def run_thermal256(args, cpu=True, big_grid=True, check=True):
    """hmap::gpu::thermal drop-in on 256^2 fBm (seed 1), talus 2/256, 50
    iterations: host-pointer call (transfers included) and device-resident,
    against the CPU restatement (1 thread and all threads; cpu) and the FP64
    oracle (check), and the same call at 4096^2 device-resident (big_grid).
    Without check the input is generated on the device and only the rates
    are reported (the oracle is not used)."""
    import numpy as np

    torch, hm, ctx = _gpu()
    if check:
        from oracle import pyoracle as O

        z0 = O.fbm_terrain(256, 256, 1)
    else:
        z0 = device_terrain(torch, hm, ctx, 256, 256, 1).cpu().numpy()
    z = z0.copy()
    hm.thermal(z, 2.0 / 256, 50)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        z = z0.copy()
        hm.thermal(z, 2.0 / 256, 50)
    call_s = (time.perf_counter() - t0) / reps
    zd = torch.from_numpy(z0).cuda()
    hm._check(hm.lib().hmg_thermal_dev(ctx.handle, zd.data_ptr(), 256, 256, None, 2.0 / 256, 50,
                                       None))
    t0 = time.perf_counter()
    for _ in range(reps):
        hm._check(hm.lib().hmg_thermal_dev(ctx.handle, zd.data_ptr(), 256, 256, None, 2.0 / 256,
                                           50, None))
    dev_s = (time.perf_counter() - t0) / reps
    cells = 256 * 256 * 50
    out = {"metric": "Gcell-iter/s thermal 256^2 x 50 (drop-in)",
           "value": round(cells / call_s / 1e9, 4), "unit": "Gcell-iter/s", "n_gpus": 1,
           "device_resident_gcell_iter_s": round(cells / dev_s / 1e9, 4)}
    if check:
        out.update(thermal256_checks(z0, z))
    if big_grid:

        # the same call at 4096^2, device-resident (two Jacobi steps per
        # streaming launch): events around the call, best of 3
        big = 4096
        zb = device_terrain(torch, hm, ctx, big, big, 1)
        st = torch.cuda.ExternalStream(ctx.stream)
        big_ms = None
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            hm._check(hm.lib().hmg_thermal_dev(ctx.handle, zb.data_ptr(), big, big, None,
                                               2.0 / big, 50, None))
            e1.record(st)
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            big_ms = ms if big_ms is None else min(big_ms, ms)
        big_rate = big * big * 50 / (big_ms / 1e3) / 1e9
        # 8 B per cell-step is one read and one write of z per step; the
        # two-step launches move half of it
        out.update({"resident_4096_gcell_iter_s": round(big_rate, 2),
                    "resident_4096_ms": round(big_ms, 3),
                    "resident_4096_gbs_at_8B_per_cell_step": round(big_rate * 8, 1)})
    if cpu:
        c1 = cpu_bench("thermal", 256, 256, 50, 5, 1)
        cn = cpu_bench("thermal", 256, 256, 50, 5)
        out.update({"cpu_1_thread_gcell_iter_s": c1["gcell_iter_s"],
                    "cpu_all_threads_gcell_iter_s": cn["gcell_iter_s"],
                    "cpu_threads": cn["threads"]})
    return out


def run_flow(args):
    """hmap::gpu::flow_simulation on 4096^2 fBm (seed 2) with the parameters of
    the reference's example (water_height 1e-2, uniform depth map, dt 0.5,
    flux diffusion 0.01), 200 iterations: device-resident and drop-in (host
    arrays), checked bit for bit against the FP32 oracle; roofline of the
    fused flow kernel at 44 B per cell-iteration (z read; d and 4 fluxes read
    and written)."""
    import numpy as np

    from oracle import pyoracle as O

    torch, hm, ctx = _gpu()
    n, iters = args.n, 200
    stream = torch.cuda.ExternalStream(ctx.stream)
    z = device_terrain(torch, hm, ctx, n, n, 2)
    dm = torch.ones_like(z)
    d = torch.empty_like(z)

    def resident(it):
        hm._check(hm.lib().hmg_flow_simulation_dev(ctx.handle, z.data_ptr(), n, n, 1e-2,
                                                   dm.data_ptr(), it, 0.5, 1, 0.01, 0.0,
                                                   d.data_ptr()))

    resident(5)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    best = 1e9
    for _ in range(3):
        e0.record(stream)
        resident(iters)
        e1.record(stream)
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / 1e3)
    zh = z.cpu().numpy()
    t0 = time.perf_counter()
    got = hm.flow_simulation(zh, 1e-2, np.ones_like(zh), iters)
    dropin = time.perf_counter() - t0
    print("flow: GPU done, running the oracle", file=sys.stderr, flush=True)
    want = O.flow_simulation(zh, 1e-2, np.ones_like(zh), iters)
    c = cpu_bench("flow", n, n, 10, 3)
    alg = 44 * n * n * iters
    return {"metric": "Gcell-iter/s flow_simulation 4096^2 x 200", "unit": "Gcell-iter/s",
            "n_gpus": 1, "value": round(n * n * iters / best / 1e9, 3),
            "dropin_gcell_iter_s": round(n * n * iters / dropin / 1e9, 3),
            "bitexact_vs_fp32_oracle": bool(np.array_equal(got, want)),
            "roofline": {"kernel": "k_flow_stream", "bound": "hbm",
                         "achieved": round(alg / best / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / best / 1e9 / HBM_PEAK_GBS, 4),
                         "avg_launch_us": round(best / iters * 1e6, 2)},
            "cpu_baseline": {"value": round(c["gcell_iter_s"], 5), "unit": "Gcell-iter/s",
                             "cores": c["threads"], "kind": "port",
                             "sample": "4096^2, 10 iterations, median of 3"}}


def run_snow(args):
    """hmap::gpu::snow_simulation.  (i) The example configuration
    (examples/ex_snow_simulation: 256^2 fBm seed 2, snow 0.1, uniform fall,
    no melting, talus 2.5/256, 500 iterations, reference defaults with the
    thermal post-filter) as drop-in calls on host arrays, checked bit for
    bit against the FP32 oracle.  (ii) 4096^2, 100 iterations, device
    resident without the post-filter, for the step kernel's roofline at
    24 B per cell-iteration (z, s, talus, melting map and fall map read, s
    written)."""
    import ctypes as C

    import numpy as np

    from oracle import pyoracle as O

    torch, hm, ctx = _gpu()
    n0, it0 = 256, 500
    z0 = O.fbm_terrain(n0, n0, 2)
    talus0 = np.float32(2.5) / np.float32(n0)
    hm.snow_simulation(z0, 0.1, 1.0, 0.0, talus0, 5)  # warm-up
    best_dropin = 1e9
    for _ in range(3):
        t0 = time.perf_counter()
        got = hm.snow_simulation(z0, 0.1, 1.0, 0.0, talus0, it0)
        best_dropin = min(best_dropin, time.perf_counter() - t0)
    print("snow: GPU example done, running the oracle", file=sys.stderr, flush=True)
    want = O.snow_simulation(z0, 0.1, 1.0, 0.0, talus0, it0)

    n, iters = args.n, 100
    stream = torch.cuda.ExternalStream(ctx.stream)
    z = device_terrain(torch, hm, ctx, n, n, 2)
    fall = torch.ones_like(z)
    melt = torch.zeros_like(z)
    tal = torch.full_like(z, 2.5 / n)
    s = torch.empty_like(z)

    def resident(it):
        hm._check(hm.lib().hmg_snow_simulation_dev(
            ctx.handle, z.data_ptr(), n, n, 0.1, fall.data_ptr(), melt.data_ptr(),
            tal.data_ptr(), it, 0.5, 1.0, 0.5, 0.1, 0.8, 1.0, 1.0, 0, 0.2, s.data_ptr()))

    resident(3)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    best = 1e9
    for _ in range(3):
        e0.record(stream)
        resident(iters)
        e1.record(stream)
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / 1e3)
    c = cpu_bench("snow", n0, n0, it0, 3)
    alg = 24 * n * n * iters
    return {"metric": "Gcell-iter/s snow_simulation 256^2 x 500 (example)",
            "unit": "Gcell-iter/s", "n_gpus": 1,
            "value": round(n0 * n0 * it0 / best_dropin / 1e9, 3),
            "bitexact_vs_fp32_oracle": bool(np.array_equal(got, want)),
            "resident_4096_gcell_iter_s": round(n * n * iters / best / 1e9, 3),
            "roofline": {"kernel": "k_snow_step", "bound": "hbm",
                         "achieved": round(alg / best / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / best / 1e9 / HBM_PEAK_GBS, 4),
                         "avg_launch_us": round(best / iters * 1e6, 2),
                         "note": "4096^2 resident run, post-filter off; time per iteration "
                                 "includes the call's fixed work spread over 100 iterations"},
            "cpu_baseline": {"value": round(c["gcell_iter_s"], 5), "unit": "Gcell-iter/s",
                             "cores": c["threads"], "kind": "port",
                             "sample": "the example configuration, 256^2 x 500, median of 3"}}


def run_tiles(args):
    """hydraulic_vpipes run tile by tile over a hmap::VirtualArray (C++
    drop-in, highmap_amd/va_bench): a 16384^2 grid (1 GiB of float32) held
    on the host as 4096^2 tiles with 64-cell halos, 50 iterations per tile,
    each tile uploaded, eroded and downloaded by the drop-in call; the
    distributed mode (one host thread per core, one device context per
    thread up to 4) against the sequential one.  PCIe-inclusive rates."""
    exe = os.path.join(ROOT, "highmap_amd", "va_bench")
    if not os.path.exists(exe):
        raise RuntimeError("highmap_amd/va_bench is not built (python -m highmap_amd.build)")
    r = subprocess.run([exe, "16384", "4096", "64", "50"], capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0:
        raise RuntimeError("va_bench failed: " + r.stdout + r.stderr)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"metric": "Gcell-iter/s vpipes over a 16384^2 VirtualArray (4096^2 tiles)",
            "unit": "Gcell-iter/s", "n_gpus": 1, "value": res["dist_gcell_iter_s"],
            "data": "synthetic: noise_fbm PERLIN per tile over its bbox, remapped",
            "detail": res}


def run_vpipes1024(args, cpu=True, check=True):
    """hydraulic_vpipes at 1024^2 (seed 2), 200 iterations: (i) reference
    semantics (gravity 1, volume top-up) with plain and fused launches,
    (ii) gravity 9.81 with 1e-3 rain per iteration; GPU vs CPU restatement
    (cpu), error vs the FP64 oracle (check).  Without check the input is
    generated on the device and only the rates are reported."""
    import numpy as np

    torch, hm, ctx = _gpu()
    n, iters = 1024, 200
    if check:
        from oracle import pyoracle as O

        z0 = O.fbm_terrain(n, n, 2)
    else:
        z0 = device_terrain(torch, hm, ctx, n, n, 2).cpu().numpy()
    res = {}
    for name, extra in (("i_reference", {}),
                        ("ii_gravity_rain", dict(gravity=9.81, rain_per_iter=1e-3,
                                                 maintain_water_volume=0))):
        kw = dict(WORKLOAD)
        kw.update(extra)
        row = {}
        for mode in ("plain", "fused"):
            z = z0.copy()
            hm.hydraulic_vpipes(z, iterations=5, mode=mode, **kw)  # warm-up
            t0 = time.perf_counter()
            z = z0.copy()
            hm.hydraulic_vpipes(z, iterations=iters, mode=mode, **kw)
            row[mode + "_dropin_gcell_iter_s"] = round(n * n * iters / (time.perf_counter() - t0)
                                                       / 1e9, 3)
            s = hm.Session(z0, mode=mode, **kw)
            s.vpipes(2)
            s.sync()
            t0 = time.perf_counter()
            s.vpipes(iters)
            s.sync()
            row[mode + "_resident_gcell_iter_s"] = round(n * n * iters / (time.perf_counter() - t0)
                                                         / 1e9, 3)
            s.close()
        if check:
            # these parameters are chaotic: FP32 arithmetic itself (the GPU
            # equals the FP32 oracle bit for bit) separates from FP64 by
            # O(range) within tens of iterations
            o64 = O.VpState(z0.astype(np.float64), O.vp_params(**kw), dtype=np.float64).run(iters)
            row["max_abs_err_rel_range_vs_fp64"] = float(np.abs(z - o64.z).max()
                                                         / (o64.z.max() - o64.z.min()))
        res[name] = row
    out = {"metric": "Gcell-iter/s vpipes 1024^2 x 200", "unit": "Gcell-iter/s", "n_gpus": 1,
           "value": res["i_reference"]["fused_resident_gcell_iter_s"], "results": res}
    if cpu:
        c = cpu_bench("vpipes", n, n, 20, 3)
        out.update({"cpu_all_threads_gcell_iter_s": c["gcell_iter_s"], "cpu_threads": c["threads"]})
    return out


_PROBE = None


def probe_lib():
    """libhmg_probe.so: the HBM stream kernels (read, write, copy)."""
    global _PROBE
    if _PROBE is None:
        import ctypes as C

        path = os.path.join(ROOT, "highmap_amd", "libhmg_probe.so")
        if not os.path.exists(path):
            raise RuntimeError("libhmg_probe.so not built (python -m highmap_amd.build)")
        L = C.CDLL(path)
        L.hmg_probe_stream.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_void_p,
                                       C.POINTER(C.c_float)]
        L.hmg_probe_stream.restype = C.c_int
        _PROBE = L
    return _PROBE


STREAM_MODES = {"read": 0, "write": 1, "copy": 2}


def vp_pattern_us(torch, nx, ny, storage, chunk_rows, reps=30):
    """Average launch time (us) of k_vp_pattern: the memory operations of one
    fused vpipes iteration with the kernel's decomposition (strips of 64
    lanes, 60 written; chunks of `chunk_rows` rows; four wavefronts per
    workgroup in row lockstep) and no erosion arithmetic, on planes of the
    storage's element size; FP32 state with the heights and water depths in
    one float2 plane, as the FP32 sessions keep them."""
    import ctypes as C

    L = probe_lib()
    L.hmg_probe_vp_pattern.argtypes = [C.c_int] + [C.c_void_p] * 8 + [C.c_int] * 4 + [
        C.POINTER(C.c_float)]
    L.hmg_probe_vp_pattern.restype = C.c_int
    elem = 8 if storage == "f32" else 2
    dt = torch.float32 if elem == 8 else torch.float16
    n = nx * ny
    z, zo = (torch.zeros(2 * n if elem == 8 else n, dtype=torch.float32, device="cuda")
             for _ in range(2))
    D, Do = (torch.zeros(1 if elem == 8 else n, dtype=dt, device="cuda") for _ in range(2))
    S, So = (torch.zeros(n, dtype=dt, device="cuda") for _ in range(2))
    F, Fo = (torch.zeros(4 * n, dtype=dt, device="cuda") for _ in range(2))
    torch.cuda.synchronize()
    ms = C.c_float()
    rc = L.hmg_probe_vp_pattern(elem, z.data_ptr(), D.data_ptr(), F.data_ptr(), S.data_ptr(),
                                zo.data_ptr(), Do.data_ptr(), Fo.data_ptr(), So.data_ptr(), nx, ny,
                                chunk_rows, reps, C.byref(ms))
    if rc != 0:
        raise RuntimeError("hmg_probe_vp_pattern failed (%d)" % rc)
    del z, zo, D, Do, S, So, F, Fo
    torch.cuda.empty_cache()
    return ms.value * 1e3


def stream_sweep(torch, gib=4, modes=("read", "write", "copy"), unrolls=(1, 2, 4, 8),
                 nts=(0, 1), blocks_per_cu=(0, 4, 8, 16), reps=3):
    """Times the HBM stream kernels over `gib` GiB buffers (far beyond the
    256 MB of cache) for every variant; returns {mode: [(GB/s, variant)...]}
    sorted fastest first.  Bytes per launch: read gib, write gib, copy 2 gib."""
    import ctypes as C

    L = probe_lib()
    nbytes = gib << 30
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    sink = torch.zeros(1, dtype=torch.float32, device="cuda")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    torch.cuda.synchronize()
    out = {}
    for mode in modes:
        rows = []
        moved = nbytes * (2 if mode == "copy" else 1)
        for u in unrolls:
            for nt in nts:
                for bpc in blocks_per_cu:
                    ms = C.c_float()
                    rc = L.hmg_probe_stream(STREAM_MODES[mode], a.data_ptr(), b.data_ptr(),
                                            nbytes, u, nt, bpc * cus, reps, sink.data_ptr(),
                                            C.byref(ms))
                    if rc != 0:
                        raise RuntimeError("hmg_probe_stream failed (%d)" % rc)
                    rows.append((round(moved / (ms.value / 1e3) / 1e9, 1),
                                 {"unroll": u, "nontemporal": nt,
                                  "grid": "one pass" if bpc == 0 else "%d blocks/CU" % bpc}))
        rows.sort(key=lambda r: -r[0])
        out[mode] = rows
    del a, b
    torch.cuda.empty_cache()
    return out


def measured_copy_gbs(torch):
    """Read + write bytes per second of the fastest HBM copy kernel variant
    (2 GiB buffers), and that variant."""
    rows = stream_sweep(torch, gib=2, modes=("copy",), unrolls=(1, 2, 4),
                        blocks_per_cu=(0,), reps=3)["copy"]
    return rows[0]


def run_copy(args):
    """A plain device copy of known size (1 GiB, three times, caches flushed
    by a 2 GiB fill in between), used to calibrate the HBM byte counters."""
    import torch

    n = 256 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    flush = torch.empty(2 * n, dtype=torch.float32, device="cuda")
    t0 = time.perf_counter()

    def best_ms(op):
        best = None
        for _ in range(3):
            flush.fill_(1.0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            op()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        return best

    copy_ms = best_ms(lambda: b.copy_(a))  # the three profiled copies
    wall = time.perf_counter() - t0
    read_ms = best_ms(lambda: a.sum())
    fill_ms = best_ms(lambda: b.fill_(2.0))
    torch.cuda.synchronize()
    # bytes per second of the best of three: what plain streams that read and
    # write HBM in equal parts, only read, or only write reach on this GPU
    return {"metric": "device copy for counter calibration", "bytes_read_per_copy": 4 * n,
            "bytes_written_per_copy": 4 * n, "copies": 3,
            "best_copy_ms": round(copy_ms, 4),
            "copy_hbm_gbs": round(8 * n / (copy_ms / 1e3) / 1e9, 1),
            "read_hbm_gbs": round(4 * n / (read_ms / 1e3) / 1e9, 1),
            "fill_hbm_gbs": round(4 * n / (fill_ms / 1e3) / 1e9, 1),
            "wall_s": round(wall, 4)}


def run_ceiling(args):
    """What this GPU's HBM delivers to streaming kernels: read-only,
    write-only and copy over 4 GiB buffers, every variant of
    highmap_amd/csrc/hbm_probe.hip (1-8 float4 loads in flight per lane,
    nontemporal or not, one-pass or striding grids), best of 3 each."""
    import torch

    t0 = time.perf_counter()
    res = stream_sweep(torch, gib=4)
    out = {"metric": "HBM stream ceiling", "unit": "GB/s", "n_gpus": 1,
           "peak_gbs": HBM_PEAK_GBS, "buffer_gib": 4}
    for mode, rows in res.items():
        out[mode] = {"best_gbs": rows[0][0], "best_variant": rows[0][1],
                     "frac_of_peak": round(rows[0][0] / HBM_PEAK_GBS, 4),
                     "all": [[g, v["unroll"], v["nontemporal"], v["grid"]] for g, v in rows]}
    # the fused vpipes kernel's access pattern at 4096^2, FP32 and 16-bit
    # state, at the chunk lengths the online tuner picks there
    n = 4096
    out["vp_pattern"] = {}
    for storage in ("f32", "bf16"):
        for cr in (70, 72):
            us = vp_pattern_us(torch, n, n, storage, cr)
            out["vp_pattern"]["%s_%d" % (storage, cr)] = {
                "avg_launch_us": round(us, 2),
                "alg_gbs": round(ALG_BYTES[storage] * n * n / (us * 1e-6) / 1e9, 1)}
    out["wall_s"] = round(time.perf_counter() - t0, 2)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--storage", default=None, choices=sorted(STORAGE),
                    help="water/sediment/flux storage (default f32; bf16 for --workload serve)")
    ap.add_argument("--mode", default="fused", choices=["fused", "plain"])
    ap.add_argument("--chunk-rows", type=int, default=0,
                    help="headline: fixed rows per wavefront (0: online tuning)")
    ap.add_argument("--prefetch", type=int, default=1, choices=[0, 1, 2],
                    help="with --chunk-rows: 0 one row ahead, 1 two rows ahead, "
                         "2 two rows ahead with the sediment gathers one row ahead")
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--iters", type=int, default=500)
    ap.add_argument("--maps", type=int, default=0)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--max-active-cells", type=int, default=3 * 4096 * 4096,
                    help="serve: cells of the jobs that run at once (0: no cap)")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "batch", "batchforms", "serve", "precision",
                             "thermal", "thermal256", "vpipes1024", "flow", "snow", "tiles",
                             "copy", "ceiling", "noise"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline: skip the BF16/FP16 kernel rates, the thermal 256^2 and "
                         "vpipes 1024^2 workloads, the batch and the served mix")
    args = ap.parse_args()
    if args.storage is None:
        args.storage = "bf16" if args.workload == "serve" else "f32"

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X and has no "
              "multi-GPU path (requested --gpus %d, WORLD_SIZE %d)" % (args.gpus, world),
              file=sys.stderr)
        sys.exit(2)
    fn = {"headline": run_headline, "batch": run_batch, "batchforms": run_batch_forms,
          "precision": run_precision,
          "thermal": run_thermal, "thermal256": run_thermal256, "vpipes1024": run_vpipes1024,
          "flow": run_flow,
          "snow": run_snow, "tiles": run_tiles, "copy": run_copy, "ceiling": run_ceiling,
          "noise": run_noise}
    if args.workload == "serve":
        from highmap_amd import serve

        out = serve.bench(args)
    else:
        out = fn[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
