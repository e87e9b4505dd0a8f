"""Batched launches: every map of a batch ends bit-identical to the same map
run alone (and to the FP32 oracle), launched per iteration, captured in a
hipGraph or as one persistent cooperative launch, and for any launch shape
and form the autotuner picks.  A persistent run whose barrier wait gives up
(forced from the host with a wait limit of one poll) reports an error and
ends instead of hanging."""
import numpy as np
import pytest

from conftest import VP_WORKLOAD

pytestmark = pytest.mark.gpu

# (nx, ny, iterations, seed): mixed shapes (one rectangular 1:2, one tiny)
MAPS = [(128, 128, 23, 100), (96, 192, 17, 101), (192, 192, 11, 102), (64, 64, 30, 103),
        (257, 130, 9, 104), (5, 3, 14, 105), (300, 300, 6, 106)]


def _single(hm, z0, iters, storage=0):
    s = hm.Session(z0, storage=storage, **VP_WORKLOAD)
    s.vpipes(iters)
    out = s.export()
    s.close()
    return out


def _assert_same(a, b):
    for name, x, y in zip("zdsuv", a, b):
        assert np.array_equal(x, y), "%s differs (max %g)" % (name, np.abs(x - y).max())


def _inputs(orc):
    return [orc.fbm_terrain(nx, ny, seed) if nx > 8 else
            np.random.default_rng(seed).random((ny, nx), dtype=np.float32)
            for nx, ny, _, seed in MAPS]


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("storage", [0, 1])
def test_batch_equals_single_sessions(hm, orc, use_graph, storage):
    zs = _inputs(orc)
    b = hm.Batch(storage=storage, **VP_WORKLOAD)
    ids = [b.add(z, it) for z, (_, _, it, _) in zip(zs, MAPS)]
    b.run(use_graph=use_graph)
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        _assert_same(b.export(mid), _single(hm, z, it, storage))
    b.close()


def test_batch_against_oracle(hm, orc):
    zs = _inputs(orc)
    b = hm.Batch(**VP_WORKLOAD)
    ids = [b.add(z, it) for z, (_, _, it, _) in zip(zs, MAPS)]
    b.run()
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        st = orc.VpState(z, orc.vp_params(**VP_WORKLOAD)).run(it)
        _assert_same(b.export(mid), (st.z, st.d, st.s, st.u, st.v))
    b.close()


def test_autotune_keeps_results(hm, orc):
    zs = _inputs(orc)
    b = hm.Batch(**VP_WORKLOAD)
    ids = [b.add(z, it + 30) for z, (_, _, it, _) in zip(zs, MAPS)]
    rows, pf = b.autotune(probe=2)
    assert rows >= 1 and pf in (0, 1, 2)
    b.run()
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        _assert_same(b.export(mid), _single(hm, z, it + 30))
    b.close()


@pytest.mark.parametrize("rows", [8, 33])
@pytest.mark.parametrize("storage", [0, 1, 2])
def test_batch_gathers_ahead_equals_single_sessions(hm, orc, rows, storage):
    """The gathers-ahead variant of the batched kernel (rows loaded two
    ahead, sediment gathers one row early) gives the same maps."""
    zs = _inputs(orc)
    b = hm.Batch(storage=storage, **VP_WORKLOAD)
    b.set_tuning(rows, 2)
    ids = [b.add(z, it) for z, (_, _, it, _) in zip(zs, MAPS)]
    b.run()
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        _assert_same(b.export(mid), _single(hm, z, it, storage))
    b.close()


def test_batch_with_rain_map_and_zero_iterations(hm, orc):
    z1 = orc.fbm_terrain(80, 60, 7)
    z2 = orc.fbm_terrain(64, 64, 8)
    rain = np.random.default_rng(2).random((60, 80), dtype=np.float32) + 0.5
    b = hm.Batch(**VP_WORKLOAD)
    a = b.add(z1, 12, rain_map=rain)
    c = b.add(z2, 0)
    b.run()
    st = orc.VpState(z1, orc.vp_params(**VP_WORKLOAD), rain).run(12)
    _assert_same(b.export(a), (st.z, st.d, st.s, st.u, st.v))
    assert np.array_equal(b.export(c)[0], z2)
    b.close()


@pytest.mark.parametrize("storage,rows,pf", [(0, 32, 2), (0, 17, 0), (1, 48, 1), (2, 8, 2),
                                             (2, 64, 0)])
def test_persistent_batch_equals_single_sessions(hm, orc, storage, rows, pf):
    """One cooperative launch for every iteration (grid barrier between
    iterations, maps leaving as they finish) gives each map's own result."""
    zs = _inputs(orc)
    b = hm.Batch(storage=storage, **VP_WORKLOAD)
    b.set_tuning(rows, pf)
    ids = [b.add(z, it) for z, (_, _, it, _) in zip(zs, MAPS)]
    b.run(form=hm.Batch.PERSISTENT)
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        _assert_same(b.export(mid), _single(hm, z, it, storage))
    b.close()


def test_persistent_batch_with_rain_maps(hm, orc):
    """A batch holding a map with a rain map runs the kernel form that reads
    rain per cell; the maps still equal their single runs."""
    zs = _inputs(orc)
    rain = np.random.default_rng(7).random((192, 96), dtype=np.float32)
    b = hm.Batch(**dict(VP_WORKLOAD, maintain_water_volume=0))
    ids = [b.add(z, it, rain_map=rain if (nx, ny) == (96, 192) else None)
           for z, (nx, ny, it, _) in zip(zs, MAPS)]
    b.run(form=hm.Batch.PERSISTENT)
    for mid, z, (nx, ny, it, _) in zip(ids, zs, MAPS):
        s = hm.Session(z, rain_map=rain if (nx, ny) == (96, 192) else None,
                       **dict(VP_WORKLOAD, maintain_water_volume=0))
        s.vpipes(it)
        _assert_same(b.export(mid), s.export())
        s.close()
    b.close()


def test_autotuned_form_keeps_results(hm, orc):
    zs = _inputs(orc)
    b = hm.Batch(storage=1, **VP_WORKLOAD)
    ids = [b.add(z, it + 40) for z, (_, _, it, _) in zip(zs, MAPS)]
    b.autotune(probe=2)
    assert b.tuned_form() in (hm.Batch.GRAPH, hm.Batch.PERSISTENT)
    b.run(form=hm.Batch.AUTO)
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        _assert_same(b.export(mid), _single(hm, z, it + 40, 1))
    b.close()


def test_persistent_barrier_wait_gives_up_without_hanging(hm, orc, monkeypatch):
    """With HMG_GRID_SPIN_LIMIT=1 every barrier wait polls once: a workgroup
    that arrives before the last one gives up, the others see the flag and
    stop, the kernel ends, and the run reports HMG_ERR_DEVICE.  (In this
    process: a second process on the device is refused cooperative launches
    while this one holds them.)"""
    b = hm.Batch(water_height=1e-3)
    b.set_tuning(8, 1)
    for k in range(6):
        b.add(orc.fbm_terrain(256, 256, k), 40 + 10 * k)
    monkeypatch.setenv("HMG_GRID_SPIN_LIMIT", "1")
    with pytest.raises(hm.HmgError) as e:
        b.run(form=hm.Batch.PERSISTENT)
    monkeypatch.delenv("HMG_GRID_SPIN_LIMIT")
    assert e.value.code == -3 and "grid barrier" in str(e.value)
    b.close()
    # the device is fine afterwards: a persistent batch runs normally
    zs = _inputs(orc)
    b = hm.Batch(**VP_WORKLOAD)
    ids = [b.add(z, it) for z, (_, _, it, _) in zip(zs, MAPS)]
    b.run(form=hm.Batch.PERSISTENT)
    for mid, z, (_, _, it, _) in zip(ids, zs, MAPS):
        _assert_same(b.export(mid), _single(hm, z, it))
    b.close()
