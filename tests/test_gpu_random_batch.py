"""Seeded random sweep of the batched form against the FP32 oracle, bit for
bit: 2 to 8 maps of random shapes (3 x 3 up, some with a rain plane) and
iteration counts (0 to 30, so maps leave the batch at different launches),
shared random parameters and storage, a random launch shape, and each form
of hmg_batch_run (launches on the stream, one hipGraph, one persistent
cooperative launch)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("z", "d", "s", "u", "v")


def _case(k):
    rng = np.random.default_rng(9000 + k)
    p = dict(water_height=float(10 ** rng.uniform(-4, -1.5)),
             evap_rate=float(rng.uniform(0.0, 0.05)),
             dt=float(rng.uniform(0.1, 0.8)),
             k_capacity=float(10 ** rng.uniform(-2, 0)),
             k_erode=float(rng.uniform(0.0, 0.08)),
             k_depose=float(rng.uniform(0.0, 0.08)),
             k_discharge_exp=float(rng.choice([1.0, 1.0, 1.3])),
             flux_diffusion=int(rng.integers(0, 2)),
             maintain_water_volume=int(rng.integers(0, 2)),
             rain_per_iter=float(rng.choice([0.0, 1e-3])))
    maps = []
    for _ in range(int(rng.integers(2, 9))):
        maps.append((int(rng.choice([3, 17, 60, 61, 128, 190])), int(rng.integers(3, 120)),
                     int(rng.integers(0, 31)), bool(rng.integers(0, 4) == 0),
                     int(rng.integers(0, 1000))))
    storage = int(rng.integers(0, 3))
    shape = (int(rng.integers(1, 80)), int(rng.integers(0, 3)))
    form = int(rng.integers(0, 3))  # HMG_BATCH_STREAM, _GRAPH, _PERSISTENT
    return p, maps, storage, shape, form


@pytest.mark.parametrize("k", range(24))
def test_random_batch_equals_oracle(hm, orc, k):
    p, maps, storage, shape, form = _case(k)
    b = hm.Batch(storage=storage, **p)
    want = []
    ids = []
    for nx, ny, iters, rain, seed in maps:
        z = orc.fbm_terrain(nx, ny, seed)
        r = np.random.default_rng(seed).random((ny, nx), dtype=np.float32) if rain else None
        ids.append(b.add(z, iters, rain_map=r))
        st = orc.VpState(z, orc.vp_params(storage=storage, **p), r).run(iters)
        want.append((st.z, st.d, st.s, st.u, st.v))
    b.set_tuning(*shape)
    try:
        b.run(form=form)
    except hm.HmgError as e:
        # a persistent launch the runtime refuses leaves the maps unchanged
        # (HMG_ERR_UNSUPPORTED); the graph form then runs them
        assert form == hm.Batch.PERSISTENT, e
        b.run(form=hm.Batch.GRAPH)
    for mid, w, m in zip(ids, want, maps):
        for name, a, o in zip(NAMES, b.export(mid), w):
            assert np.array_equal(a, o), "case %d map %s %s: %d cells differ (storage %d form %d %s)" % (
                k, m, name, int((a != o).sum()), storage, form, p)
    b.close()
