"""GPU parity at the benchmark workloads' real sizes, against the FP32 oracle,
bit for bit (every plane the drop-in returns: z, d, s, u, v).

  - vpipes 1024^2 x 200 (fBm seed 2), both the reference's semantics
    (gravity 1, water volume maintained) and the extension (gravity 9.81,
    rain 1e-3 per iteration, no top-up), fused (online shape tuning active)
    and plain;
  - vpipes + thermal 4096^2 (fBm seed 2, talus 1/4096 after every 10th
    hydraulic iteration, one continuous state): FP32 for 500 iterations,
    BF16 and FP16 storage for 80 (the online tuner probes iterations 20-55,
    so the tuned shape runs after that);
  - the 96-map batch manifest: every map batched equals the map alone (the
    persistent launch gives the same maps), and 12 maps chosen across sizes
    and shapes equal the oracle;
  - the first 40 requests of the served Poisson stream at their real sides
    (256-3456) with BF16 state: served equals alone for all, and five jobs
    of both kinds equal the oracle.
The device-generated inputs (noise_fbm_dev then remap_dev, which the
benchmark and the server use) are checked against the oracle's terrain
first.  The oracle runs on the host (OpenMP); the largest case, 8.4e9
cell-iterations at 4096^2, takes one to three minutes there.
"""
import json
import os

import numpy as np
import pytest

from conftest import VP_WORKLOAD

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("z", "d", "s", "u", "v")


def _assert_same(g, o, what=""):
    for name, a, b in zip(NAMES, g, o):
        assert np.array_equal(a, b), "%s %s differs: max |diff| %g, %d cells" % (
            what, name, np.abs(a - b).max(), int((a != b).sum()))


def _dev_terrain(hm, n_x, n_y, seed):
    """remap(noise_fbm(PERLIN, kw 4, seed)) generated on the device, as a
    torch tensor."""
    import torch

    from highmap_amd import _check, lib

    ctx = hm._default_ctx
    z = torch.empty((n_y, n_x), dtype=torch.float32, device="cuda")
    _check(lib().hmg_noise_fbm_dev(ctx.handle, z.data_ptr(), n_x, n_y, 1, 4.0, 4.0, seed, 8,
                                   0.7, 0.5, 2.0, None, None, None, None, None))
    hm.remap(z.data_ptr(), n_x * n_y, 0.0, 1.0)
    torch.cuda.synchronize()
    return z


@pytest.mark.parametrize("shape,seed", [((4096, 4096), 2), ((1024, 1024), 2), ((256, 256), 1),
                                        ((448, 448), 1000), ((96, 192), 150), ((333, 77), 9)])
def test_device_generated_terrain_equals_oracle(hm, orc, shape, seed):
    """hmg_noise_fbm_dev + hmg_remap_dev (the inputs of the benchmark and of
    every served job) against the oracle's noise_fbm + remap."""
    nx, ny = shape
    g = _dev_terrain(hm, nx, ny, seed).cpu().numpy()
    o = orc.fbm_terrain(nx, ny, seed)
    assert np.array_equal(g, o), np.abs(g - o).max()
    assert g.min() == 0.0 and g.max() == 1.0


# ------------------------------------------------------- vpipes 1024^2 x 200

VARIANTS = {
    "reference": dict(VP_WORKLOAD),
    "gravity_rain": dict(VP_WORKLOAD, gravity=9.81, rain_per_iter=1e-3,
                         maintain_water_volume=0),
}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_vpipes_1024_200_iterations(hm, orc, variant):
    kw = VARIANTS[variant]
    z0 = orc.fbm_terrain(1024, 1024, 2)
    st = orc.VpState(z0, orc.vp_params(**kw)).run(200)
    want = (st.z, st.d, st.s, st.u, st.v)
    for mode in ("fused", "plain"):
        s = hm.Session(z0, mode=mode, **kw)
        s.vpipes(200)
        _assert_same(s.export(), want, "%s %s" % (variant, mode))
        if mode == "fused":
            assert not s.tuning()[2], "online tuning still probing after 200 iterations"
        s.close()


# ------------------------------------------- vpipes + thermal 4096^2, 3 storages


def _interleaved(sess, o, iters, talus, checkpoints):
    """10 hydraulic iterations then one thermal step, on both states; yields
    at each checkpoint (multiples of 10)."""
    done = 0
    while done < iters:
        sess.vpipes(10).thermal(talus, 1)
        o.run(10).thermal_step(talus, 1)
        done += 10
        if done % 50 == 0:
            print("%d x %d: %d iterations done" % (o.nx, o.ny, done), flush=True)
        if done in checkpoints:
            yield done


def test_vpipes_thermal_4096_f32_500_iterations(hm, orc):
    """The headline state at its full length: 500 hydraulic iterations with
    50 thermal steps, fused with online tuning, from the device-generated
    input; the plain form gives the same final state."""
    n = 4096
    z0 = orc.fbm_terrain(n, n, 2)
    zd = _dev_terrain(hm, n, n, 2)
    s = hm.Session(zd, **VP_WORKLOAD)
    o = orc.VpState(z0, orc.vp_params(**VP_WORKLOAD))
    for it in _interleaved(s, o, 500, 1.0 / n, (10, 50, 500)):
        print("4096^2 f32: checking iteration %d" % it, flush=True)
        _assert_same(s.export(), (o.z, o.d, o.s, o.u, o.v), "iteration %d" % it)
    assert not s.tuning()[2]
    fused = s.export()
    s.close()
    p = hm.Session(z0, mode="plain", **VP_WORKLOAD)
    for _ in range(50):
        p.vpipes(10).thermal(1.0 / n, 1)
    _assert_same(p.export(), fused, "plain vs fused")
    p.close()


@pytest.mark.parametrize("storage", [1, 2])
def test_vpipes_thermal_4096_16bit_80_iterations(hm, orc, storage):
    """BF16 (1) / FP16 (2) storage at 4096^2: the online tuner's probes
    (iterations 20-55) and the tuned shape after them, against the oracle
    with the same storage rounding."""
    n = 4096
    z0 = orc.fbm_terrain(n, n, 2)
    s = hm.Session(z0, storage=storage, **VP_WORKLOAD)
    o = orc.VpState(z0, orc.vp_params(storage=storage, **VP_WORKLOAD))
    for it in _interleaved(s, o, 80, 1.0 / n, (30, 80)):
        _assert_same(s.export(), (o.z, o.d, o.s, o.u, o.v), "storage %d it %d" % (storage, it))
    assert not s.tuning()[2]
    s.close()


# ------------------------------------------------------------ 96-map batch


def _manifest():
    with open(os.path.join(HERE, "golden", "batch_manifest.json")) as f:
        return json.load(f)


def _oracle_subset(maps):
    """12 maps across the manifest: the shortest run of each square side,
    plus the shortest of each rectangular shape, until 12."""
    by_shape = {}
    for k, m in enumerate(maps):
        key = (m["nx"], m["ny"])
        if key not in by_shape or m["iterations"] < maps[by_shape[key]]["iterations"]:
            by_shape[key] = k
    squares = sorted(k for key, k in by_shape.items() if key[0] == key[1])
    rects = sorted(k for key, k in by_shape.items() if key[0] != key[1])
    pick = (squares + rects)[:12]
    assert len(pick) == 12
    return pick


def test_batch_manifest_96_maps(hm, orc):
    maps = _manifest()
    assert len(maps) == 96
    zs = [orc.fbm_terrain(m["nx"], m["ny"], m["seed"]) for m in maps]
    b = hm.Batch(**VP_WORKLOAD)
    ids = [b.add(z, m["iterations"]) for z, m in zip(zs, maps)]
    rows, pf = b.autotune(probe=2)
    b.run(use_graph=True)
    outs = [b.export(i) for i in ids]
    b.close()
    # the same batch as one persistent launch
    b = hm.Batch(**VP_WORKLOAD)
    b.set_tuning(rows, pf)
    ids = [b.add(z, m["iterations"]) for z, m in zip(zs, maps)]
    b.run(form=hm.Batch.PERSISTENT)
    for k, i in enumerate(ids):
        _assert_same(b.export(i), outs[k], "persistent map %d" % k)
    b.close()
    for k, (z, m) in enumerate(zip(zs, maps)):
        s = hm.Session(z, **VP_WORKLOAD)
        s.vpipes(m["iterations"])
        _assert_same(outs[k], s.export(), "map %d (%dx%d, %d it)" % (
            k, m["nx"], m["ny"], m["iterations"]))
        s.close()
    for k in _oracle_subset(maps):
        m = maps[k]
        st = orc.VpState(zs[k], orc.vp_params(**VP_WORKLOAD)).run(m["iterations"])
        _assert_same(outs[k], (st.z, st.d, st.s, st.u, st.v), "map %d vs oracle" % k)


# ------------------------------------------------------ served request stream


def _requests(n):
    with open(os.path.join(HERE, "golden", "serve_requests_head.json")) as f:
        return json.load(f)[:n]


# two thermal-only jobs and three vpipes+thermal jobs of the first 40
ORACLE_REQS = (6, 13, 35, 28, 32)


def test_served_stream_head_at_real_sizes(hm, orc):
    from highmap_amd import serve

    reqs = _requests(40)
    p = hm.VpipesParams.make(**VP_WORKLOAD)
    sv = serve.Server(storage=hm.Storage.BF16, params=p)
    ids = []
    for r in reqs:
        kind = serve.JOB_VPIPES_THERMAL if r["kind"] == "vpipes+thermal" else serve.JOB_THERMAL
        ids.append(sv.submit(r["side"], r["side"], kind, r["iterations"],
                             serve.job_talus(kind, r["side"]), seed=r["seed"]))
    sv.wait_all()
    got = {r["id"]: sv.result(j) for r, j in zip(reqs, ids)}
    sv.close()
    for r in reqs:
        alone = serve.run_alone(hm, r, hm.Storage.BF16, p, mode="fused")
        assert np.array_equal(got[r["id"]], alone), r
    kinds = set()
    for rid in ORACLE_REQS:
        r = reqs[rid]
        n = r["side"]
        z0 = orc.fbm_terrain(n, n, r["seed"])
        if r["kind"] == "thermal":
            want = orc.thermal(z0, 2.0 / n, r["iterations"])
        else:
            o = orc.VpState(z0, orc.vp_params(storage=1, **VP_WORKLOAD))
            left = r["iterations"]
            while left > 0:
                k = min(10, left)
                o.run(k)
                left -= k
                if k == 10:
                    o.thermal_step(1.0 / n, 1)
            want = o.z
        assert np.array_equal(got[rid], want), r
        kinds.add(r["kind"])
    assert kinds == {"thermal", "vpipes+thermal"}
