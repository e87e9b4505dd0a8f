"""snow_simulation on the GPU against the FP32 oracle, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = np.float32


def _terrain(orc, nx, ny, seed):
    return orc.fbm_terrain(nx, ny, seed)


def test_snow_example_configuration(hm, orc):
    """ex_snow_simulation: 256^2, snow 0.1, uniform fall, no melting,
    talus 2.5 / 256, 500 iterations, reference defaults (post-filter on)."""
    z = _terrain(orc, 256, 256, 2)
    talus = F(2.5) / F(256)
    z0 = z.copy()
    got = hm.snow_simulation(z, 0.1, 1.0, 0.0, talus, 500)
    want = orc.snow_simulation(z, 0.1, 1.0, 0.0, talus, 500)
    assert np.array_equal(got, want)
    assert np.array_equal(z, z0)
    assert got.max() > 0


@pytest.mark.parametrize("shape", [(3, 3), (5, 4), (63, 5), (125, 9), (96, 64), (130, 257)])
@pytest.mark.parametrize("iters,kw", [
    (0, {}), (1, {}), (2, dict(post_filter=False)),
    (37, dict(fall_iterations_ratio=0.4, post_filter=False)),
    (25, dict(dt=0.3, k_snow=0.9, k_visc=0.3, k_melt_factor=2.0, k_depth_ratio=0.05,
              k_depth_slope_ratio=3.0, thermal_talus_ratio=0.5)),
    (12, dict(fall_iterations_ratio=3.0))])
def test_snow_matches_oracle(hm, orc, shape, iters, kw):
    nx, ny = shape
    r = np.random.default_rng(nx * 31 + ny)
    z = r.random((ny, nx), dtype=np.float32) if min(nx, ny) < 8 else _terrain(orc, nx, ny, 7)
    fall = r.random((ny, nx), dtype=np.float32)
    melt = r.random((ny, nx), dtype=np.float32)
    talus = (r.random((ny, nx), dtype=np.float32) * F(0.02)).astype(np.float32)
    got = hm.snow_simulation(z, 0.08, fall, melt, talus, iters, **kw)
    want = orc.snow_simulation(z, 0.08, fall, melt, talus, iters, **kw)
    assert np.array_equal(got, want), (shape, iters, kw)


def test_snow_device_entry_matches_host_entry(hm, orc):
    import ctypes as C

    import torch

    ctx = hm.init_gpu()
    ny, nx = 200, 300
    z = _terrain(orc, nx, ny, 11)
    melt = np.random.default_rng(5).random((ny, nx), dtype=np.float32)
    want = hm.snow_simulation(z, 0.1, 1.0, melt, 0.01, 60)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda()
           for a in (z, np.ones_like(z), melt, np.full_like(z, F(0.01)))]
    out = torch.empty_like(dev[0])
    rc = hm.lib().hmg_snow_simulation_dev(
        hm._default_ctx.handle, C.c_void_p(dev[0].data_ptr()), nx, ny, 0.1,
        C.c_void_p(dev[1].data_ptr()), C.c_void_p(dev[2].data_ptr()),
        C.c_void_p(dev[3].data_ptr()), 60, 0.5, 1.0, 0.5, 0.1, 0.8, 1.0, 1.0, 1, 0.2,
        C.c_void_p(out.data_ptr()))
    assert rc == 0, hm.last_error() if hasattr(hm, "last_error") else rc
    assert np.array_equal(out.cpu().numpy(), want)
    assert ctx is not None


def test_snow_large_grid(hm, orc):
    z = _terrain(orc, 1024, 768, 3)
    got = hm.snow_simulation(z, 0.1, 1.0, 0.0, 2.5 / 1024, 20)
    want = orc.snow_simulation(z, 0.1, 1.0, 0.0, 2.5 / 1024, 20)
    assert np.array_equal(got, want)


def test_snow_rejects_bad_arguments(hm):
    z = np.zeros((8, 8), np.float32)
    with pytest.raises(hm.HmgError):
        hm.snow_simulation(z, 0.1, 1.0, 0.0, 0.01, -1)
    with pytest.raises(hm.HmgError):
        hm.snow_simulation(np.zeros((2, 8), np.float32), 0.1, 1.0, 0.0, 0.01, 3)
    with pytest.raises(ValueError):
        hm.snow_simulation(z, 0.1, np.ones((8, 9), np.float32), 0.0, 0.01, 3)
