"""Continuous-batching erosion server (hmg_server_*) and its benchmark.

Server(...) wraps the C ABI: submit() queues a job (the input terrain is
either passed or generated on the device from a seed), poll()/wait_all()
follow it, and the finished z lands in the output array given at submit.

bench(args) replays the Poisson request stream of highmap_amd.workloads
against the server in real time (BF16 state, the 1024^2-workload
parameters), then runs the same requests one after another with plain
launches, each alone, and reports throughput and p50/p99 latency
(arrival to result on the host) for both, plus whether each checked job's
result is the same alone and in the mix.
"""
import ctypes as C
import threading
import time

import numpy as np

from . import HmgError, Session, Storage, VpipesParams, _check, _ctx, _ptr, lib
from . import workloads

JOB_VPIPES_THERMAL, JOB_THERMAL = 0, 1
QUEUED, RUNNING, DONE = 1, 2, 3


class JobDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("kind", C.c_int), ("iterations", C.c_int),
                ("thermal_every", C.c_int), ("talus", C.c_float), ("has_input", C.c_int),
                ("seed", C.c_uint32)]


def _bind():
    L = lib()
    if getattr(L, "_serve_bound", False):
        return L
    vp, i32, i64, dbl, sz = C.c_void_p, C.c_int, C.c_longlong, C.c_double, C.c_size_t
    L.hmg_server_create.argtypes = [vp, sz, i32, C.POINTER(VpipesParams), C.POINTER(vp)]
    L.hmg_server_submit.argtypes = [vp, C.POINTER(JobDesc), vp, vp, C.POINTER(i64)]
    L.hmg_server_poll.argtypes = [vp, i64, C.POINTER(i32)]
    L.hmg_server_job_times.argtypes = [vp, i64, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]
    L.hmg_server_wait_all.argtypes = [vp]
    L.hmg_server_set_max_active_cells.argtypes = [vp, sz]
    L.hmg_server_stats.argtypes = [vp, C.POINTER(dbl), C.POINTER(i64), C.POINTER(dbl),
                                   C.POINTER(sz), C.POINTER(sz)]
    L.hmg_server_destroy.argtypes = [vp]
    for n in ("hmg_server_create", "hmg_server_submit", "hmg_server_poll",
              "hmg_server_job_times", "hmg_server_wait_all", "hmg_server_stats",
              "hmg_server_destroy", "hmg_server_set_max_active_cells"):
        getattr(L, n).restype = i32
    L._serve_bound = True
    return L


class Server:
    def __init__(self, storage=Storage.BF16, hbm_budget=0, params=None, ctx=None,
                 max_active_cells=0, **kw):
        L = _bind()
        self.ctx = _ctx(ctx)
        self.params = params if params is not None else VpipesParams.make(**kw)
        h = C.c_void_p()
        _check(L.hmg_server_create(self.ctx.handle, int(hbm_budget), int(storage),
                                   C.byref(self.params), C.byref(h)))
        self.handle = h
        self._out = {}
        if max_active_cells:
            _check(L.hmg_server_set_max_active_cells(h, int(max_active_cells)))

    def submit(self, nx, ny, kind, iterations, talus, seed=0, z=None, thermal_every=10):
        d = JobDesc(nx, ny, kind, iterations, thermal_every, talus, 0 if z is None else 1, seed)
        out = np.empty((ny, nx), np.float32)
        zin = None if z is None else np.ascontiguousarray(z, np.float32)
        jid = C.c_longlong()
        _check(lib().hmg_server_submit(self.handle, C.byref(d), _ptr(zin), _ptr(out),
                                       C.byref(jid)))
        self._out[jid.value] = out
        return jid.value

    def poll(self, jid):
        st = C.c_int()
        _check(lib().hmg_server_poll(self.handle, jid, C.byref(st)))
        return st.value

    def result(self, jid):
        st = self.poll(jid)
        if st < 0:
            raise HmgError(st, "job %d failed" % jid)
        if st != DONE:
            return None
        return self._out.pop(jid)

    def times(self, jid):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().hmg_server_job_times(self.handle, jid, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def wait_all(self):
        _check(lib().hmg_server_wait_all(self.handle))

    def stats(self):
        ci, tk, bs = C.c_double(), C.c_longlong(), C.c_double()
        bud, used = C.c_size_t(), C.c_size_t()
        _check(lib().hmg_server_stats(self.handle, C.byref(ci), C.byref(tk), C.byref(bs),
                                      C.byref(bud), C.byref(used)))
        return {"cell_iters": ci.value, "ticks": tk.value, "busy_s": bs.value,
                "budget": bud.value, "used": used.value}

    def close(self):
        if getattr(self, "handle", None):
            lib().hmg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_talus(kind, side):
    """talus 1/n for the mixed jobs (one thermal step per 10 hydraulic) and
    2/n for thermal-only jobs."""
    return 1.0 / side if kind == JOB_VPIPES_THERMAL else 2.0 / side


def run_alone(hm, req, storage, params, mode="plain"):
    """One request processed alone: generated terrain, then the job's
    schedule on a fresh session (plain launches by default).  Returns z."""
    import torch

    ctx = hm._default_ctx
    n = req["side"]
    z = torch.empty((n, n), dtype=torch.float32, device="cuda")
    _check(lib().hmg_noise_fbm_dev(ctx.handle, z.data_ptr(), n, n, 1, 4.0, 4.0, req["seed"], 8,
                                   0.7, 0.5, 2.0, None, None, None, None, None))
    hm.remap(z.data_ptr(), n * n, 0.0, 1.0)
    kind = JOB_VPIPES_THERMAL if req["kind"] == "vpipes+thermal" else JOB_THERMAL
    talus = job_talus(kind, n)
    if kind == JOB_THERMAL:
        _check(lib().hmg_thermal_dev(ctx.handle, z.data_ptr(), n, n, None, talus,
                                     req["iterations"], None))
        torch.cuda.synchronize()
        return z.cpu().numpy()
    s = Session(z, params=params, storage=storage, mode=mode)
    left = req["iterations"]
    while left > 0:
        k = min(10, left)
        s.vpipes(k)
        left -= k
        if k == 10:
            s.thermal(talus, 1)
    zz = s.export()[0]
    s.close()
    return zz


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def measure(hm, storage, params, reqs, cap=0, sequential=True, check_ids=None):
    """Replays `reqs` (arrival times in seconds from the start) against a
    server in real time and returns throughput and latency (arrival to the
    result on the host), and, with `sequential`, the same requests run one
    after another with plain launches, each alone, with FIFO queueing
    computed from their measured service times."""
    cell_iters = sum(r["side"] ** 2 * r["iterations"] for r in reqs)
    span = reqs[-1]["arrival_s"]
    if check_ids is None:
        check_ids = sorted({0, 1, 2, len(reqs) // 2, len(reqs) - 1})

    def served(cap):
        sv = Server(storage=storage, params=params, max_active_cells=cap)
        jobs = {}
        lock = threading.Lock()

        def submitter():
            t0 = time.monotonic()
            for r in reqs:
                dt = t0 + r["arrival_s"] - time.monotonic()
                if dt > 0:
                    time.sleep(dt)
                kind = JOB_VPIPES_THERMAL if r["kind"] == "vpipes+thermal" else JOB_THERMAL
                jid = sv.submit(r["side"], r["side"], kind, r["iterations"],
                                job_talus(kind, r["side"]), seed=r["seed"])
                with lock:
                    jobs[r["id"]] = jid

        th = threading.Thread(target=submitter)
        wall0 = time.monotonic()
        th.start()
        done = {}
        kept = {}
        while len(done) < len(reqs):
            with lock:
                pending = [(i, j) for i, j in jobs.items() if i not in done]
            for i, j in pending:
                st = sv.poll(j)
                if st == DONE or st < 0:
                    z = sv.result(j) if st == DONE else None
                    if i in check_ids:
                        kept[i] = z
                    done[i] = sv.times(j)
            time.sleep(0.0005)
        th.join()
        wall = time.monotonic() - wall0
        stats = sv.stats()
        sv.close()
        lat = [t[2] - t[0] for t in done.values()]
        first_submit = min(t[0] for t in done.values())
        last_done = max(t[2] for t in done.values())
        return {"rate": cell_iters / (last_done - first_submit) / 1e9, "lat": lat, "wall": wall,
                "stats": stats, "kept": kept}

    main = served(cap)
    out = {"requests": len(reqs), "arrival_span_s": round(span, 3),
           "offered_gcell_iter_s": round(cell_iters / span / 1e9, 2),
           "achieved_gcell_iter_s": round(main["rate"], 3),
           "p50_latency_s": round(_pct(main["lat"], 50), 4),
           "p99_latency_s": round(_pct(main["lat"], 99), 4),
           "wall_s": round(main["wall"], 3),
           "server_busy_s": round(main["stats"]["busy_s"], 3),
           "server_ticks": main["stats"]["ticks"],
           "max_active_cells": cap}
    if sequential:
        service = []
        same = {}
        for r in reqs:
            t0 = time.monotonic()
            z = run_alone(hm, r, storage, params, "plain")
            service.append(time.monotonic() - t0)
            if r["id"] in main["kept"]:
                k = main["kept"][r["id"]]
                same[str(r["id"])] = bool(k is not None and np.array_equal(k, z))
        finish = 0.0
        seq_lat = []
        for r, t in zip(reqs, service):
            finish = max(finish, r["arrival_s"]) + t
            seq_lat.append(finish - r["arrival_s"])
        out["sequential_plain"] = {"gcell_iter_s": round(cell_iters / finish / 1e9, 3),
                                   "p50_latency_s": round(_pct(seq_lat, 50), 4),
                                   "p99_latency_s": round(_pct(seq_lat, 99), 4),
                                   "total_service_s": round(sum(service), 3)}
        out["alone_equals_mix"] = same
    out["_main"] = main
    return out


def warm_up(storage, params):
    """Kernels loaded and the buffer pool primed before a timed replay."""
    sv = Server(storage=storage, params=params)
    for seed in (1, 2):
        sv.submit(256, 256, JOB_VPIPES_THERMAL, 20, 1.0 / 256, seed=seed)
    sv.wait_all()
    sv.close()


def bench(args):
    import highmap_amd as hm

    hm.init_gpu(0)
    storage = {"f32": Storage.F32, "bf16": Storage.BF16, "f16": Storage.F16}[args.storage]
    params = VpipesParams.make(water_height=1e-3, evap_rate=0.01, dt=0.5, k_capacity=0.1,
                               k_erode=0.05, k_depose=0.05)
    reqs = workloads.serve_requests(n=args.requests)
    warm_up(storage, params)
    cap = int(getattr(args, "max_active_cells", 0) or 0)
    m = measure(hm, storage, params, reqs, cap)
    main = m.pop("_main")
    other = None
    if cap:
        o = measure(hm, storage, params, reqs, 0, sequential=False)
        o.pop("_main")
        other = {"gcell_iter_s": o["achieved_gcell_iter_s"], "p50_latency_s": o["p50_latency_s"],
                 "p99_latency_s": o["p99_latency_s"]}
    return {
        "metric": "served mix: Gcell-iter/s and p50/p99 latency (continuous batching)",
        "value": round(main["rate"], 3),
        "unit": "Gcell-iter/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 math, %s storage" % {1: "bf16", 2: "f16"}.get(storage, "f32"),
        "data": "synthetic: Poisson stream (rate 200/s, seed 12345), fBm terrain per request",
        "config": {"requests": len(reqs), "arrival_span_s": m["arrival_span_s"],
                   "offered_gcell_iter_s": m["offered_gcell_iter_s"],
                   "mix": "70% vpipes+thermal (1:10) / 30% thermal-only",
                   "side": "log-uniform 256-4096 (multiples of 64)", "iterations": "50-500"},
        "p50_latency_s": m["p50_latency_s"],
        "p99_latency_s": m["p99_latency_s"],
        "wall_s": m["wall_s"],
        "server_busy_s": m["server_busy_s"],
        "server_ticks": m["server_ticks"],
        "sequential_plain": m["sequential_plain"],
        "alone_equals_mix": m["alone_equals_mix"],
        "max_active_cells": cap,
        "uncapped": other,
    }
