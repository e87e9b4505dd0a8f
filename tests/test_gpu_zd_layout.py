"""FP32 sessions in the fused form keep the heights and the water depths in
one float2 plane.  The paths around it against the FP32 oracle, bit for
bit: thermal steps on such a session by the float2 streaming kernel
(uniform talus >= 0, from 4 x 4) and by the split / ordinary path / merge
fallback (a negative talus, grids below 4 x 4, zero steps);
hmg_session_z_dev's copy of the heights; and the server with FP32 state,
whose thermal-only jobs keep separate planes, each job equal to the same
job alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("z", "d", "s", "u", "v")
KW = dict(water_height=1e-3, evap_rate=0.01, dt=0.5, k_capacity=0.1, k_erode=0.05,
          k_depose=0.05)


def _input(orc, nx, ny, seed):
    if nx >= 8 and ny >= 8:
        return orc.fbm_terrain(nx, ny, seed)
    return np.random.default_rng(seed).random((ny, nx), dtype=np.float32)


@pytest.mark.parametrize("nx,ny", [(3, 3), (3, 9), (9, 3), (4, 4), (61, 5), (130, 67)])
@pytest.mark.parametrize("talus,steps", [(0.01, 1), (0.01, 3), (0.0, 2), (-1e-3, 1),
                                         (-1e-3, 2), (0.02, 0)])
def test_float2_session_thermal_matches_oracle(hm, orc, nx, ny, talus, steps):
    z0 = _input(orc, nx, ny, nx * 31 + ny)
    s = hm.Session(z0, **KW)
    o = orc.VpState(z0, orc.vp_params(**KW))
    s.vpipes(3).thermal(talus, steps).vpipes(2)
    o.run(3).thermal_step(talus, steps).run(2)
    got = s.export()
    s.close()
    for name, a, b in zip(NAMES, got, (o.z, o.d, o.s, o.u, o.v)):
        assert np.array_equal(a, b), (name, int((a != b).sum()))


def _dev_copy(ptr, n):
    import torch

    class _View:
        __cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (ptr, False),
                                    "version": 2}

    return torch.as_tensor(_View(), device="cuda").cpu().numpy().copy()


@pytest.mark.parametrize("storage", [0, 1])
def test_session_z_dev_is_the_heights(hm, orc, storage):
    """FP32 (float2 layout): a copy of the heights; BF16: the live plane."""
    import torch

    nx, ny = 257, 131
    z0 = orc.fbm_terrain(nx, ny, 4)
    s = hm.Session(z0, storage=storage, **KW)
    s.vpipes(7).thermal(0.01, 1).vpipes(3)
    p = s.z_dev_ptr()
    assert p
    torch.cuda.synchronize()
    got = _dev_copy(p, nx * ny).reshape(ny, nx)
    assert np.array_equal(got, s.export()[0])
    s.close()


def test_server_f32_jobs_equal_alone(hm, orc):
    """FP32 server state: a vpipes + thermal job (float2 sessions, batched
    launches) and thermal-only jobs (separate planes) beside each other."""
    from highmap_amd import serve

    p = hm.VpipesParams.make(**KW)
    reqs = [dict(id=0, side=192, kind="vpipes+thermal", iterations=23, seed=5),
            dict(id=1, side=128, kind="thermal", iterations=7, seed=6),
            dict(id=2, side=256, kind="vpipes+thermal", iterations=31, seed=7),
            dict(id=3, side=64, kind="thermal", iterations=0, seed=8)]
    sv = serve.Server(storage=hm.Storage.F32, params=p)
    ids = []
    for r in reqs:
        kind = serve.JOB_VPIPES_THERMAL if r["kind"] == "vpipes+thermal" else serve.JOB_THERMAL
        ids.append(sv.submit(r["side"], r["side"], kind, r["iterations"],
                             serve.job_talus(kind, r["side"]), seed=r["seed"]))
    sv.wait_all()
    got = [sv.result(j) for j in ids]
    sv.close()
    for r, g in zip(reqs, got):
        assert g is not None, r
        assert np.array_equal(g, serve.run_alone(hm, r, hm.Storage.F32, p, mode="fused")), r
