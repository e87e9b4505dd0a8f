#!/usr/bin/env python3
"""Headline benchmark: vpipes + thermal on a 4096 x 4096 grid, one MI355X.

One step = 10 hydraulic_vpipes iterations followed by one thermal step
(talus 1/4096, then extrapolate_borders) on one continuous device-resident
state, with the 1024^2-workload parameters (water_height 1e-3 with volume
top-up, dt 0.5, k_capacity 0.1, k_erode 0.05, k_depose 0.05, evaporation
0.01).  The default 50 timed steps are one full 500-iteration job.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--storage f32|bf16|f16]

Prints one JSON line.  `value` is hydraulic cell-iterations per second
(nx*ny*iterations / wall time of the timed steps, thermal included in the
time), with the state already resident in HBM.  `roofline` is measured on the
fused vpipes kernel alone with HIP events on the library's stream.
`cpu_baseline` times the FP32 CPU restatement (oracle/cpu_bench) on a
bounded sample of the same workload.
"""
import argparse
import json
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


def cpu_baseline(n, iters=20, reps=3):
    exe = os.path.join(ROOT, "oracle", "cpu_bench")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    threads = max(1, min(16, os.cpu_count() or 1))
    out = subprocess.run([exe, "vpipes_thermal", str(n), str(n), str(iters), str(threads),
                          str(reps)], capture_output=True, text=True, check=True).stdout
    r = json.loads(out.strip().splitlines()[-1])
    return {
        "value": round(r["gcell_iter_s"], 5),
        "unit": "Gcell-iter/s",
        "cores": threads,
        "kind": "port",
        "sample": "%dx%d fBm seed 2, %d vpipes iterations with a thermal step every 10, "
                  "FP32 CPU restatement of the reference passes built -O3 -Ofast -ffast-math, "
                  "OpenMP %d threads, median of %d runs after 1 warm-up (%.1f s each)"
                  % (n, n, iters, threads, reps, r["median_s"]),
    }


def load_traffic(key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        e = t[key]
        return int(e["fetch_kib"] * e["fetch_scale"] * 1024 + e["write_kib"] * e["write_scale"] * 1024)
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--storage", default="f32", choices=sorted(STORAGE))
    ap.add_argument("--mode", default="fused", choices=["fused", "plain"])
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X and has no "
              "multi-GPU path (requested --gpus %d, WORLD_SIZE %d)" % (args.gpus, world),
              file=sys.stderr)
        sys.exit(2)

    import torch

    import highmap_amd as hm

    n = args.n
    hm.init_gpu(0)
    ctx = hm._default_ctx
    stream = torch.cuda.ExternalStream(ctx.stream)
    dev = torch.device("cuda", 0)

    # input: remap(noise_fbm(PERLIN, kw 4, seed 2)) generated on the device
    z = torch.empty((n, n), dtype=torch.float32, device=dev)
    L = hm.lib()
    hm._check(L.hmg_noise_fbm_dev(ctx.handle, z.data_ptr(), n, n, hm.NoiseType.PERLIN, 4.0, 4.0,
                                  2, 8, 0.7, 0.5, 2.0, None, None, None, None, None))
    hm.remap(z.data_ptr(), n * n, 0.0, 1.0)
    torch.cuda.synchronize()

    sess = hm.Session(z, storage=STORAGE[args.storage], mode=args.mode, **WORKLOAD)
    talus = 1.0 / n

    def step():
        sess.vpipes(10)
        sess.thermal(talus, 1)

    for _ in range(args.warmup):
        step()
    sess.sync()
    torch.cuda.synchronize()

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
    gpu_ms = e0.elapsed_time(e1)
    elapsed = max(wall, gpu_ms / 1e3)
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
    sess.close()

    kernel = "k_vp_fused" if args.mode == "fused" else "k_vp_* (plain passes)"
    traffic = load_traffic("vpipes_%s_%s_%d" % (args.mode, args.storage, n))
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
        },
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
            "traffic": traffic,
            "alg_bytes_per_launch": alg,
            "avg_launch_us": round(launch_s * 1e6, 2),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
