"""flow_simulation on the GPU against the FP32 oracle, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(orc, nx, ny, seed):
    z = orc.fbm_terrain(nx, ny, seed) if min(nx, ny) > 2 else \
        np.random.default_rng(seed).random((ny, nx), dtype=np.float32)
    dm = np.random.default_rng(seed + 7).random((ny, nx), dtype=np.float32)
    return z, dm


@pytest.mark.parametrize("shape", [(5, 3), (1, 9), (64, 64), (257, 130), (300, 301)])
@pytest.mark.parametrize("iters,kw", [
    (0, {}), (1, {}), (23, {}), (40, dict(flux_diffusion=False)),
    (30, dict(dry_out_ratio=0.2)), (17, dict(dt=0.25, flux_diffusion_strength=0.05))])
def test_flow_simulation_matches_oracle(hm, orc, shape, iters, kw):
    nx, ny = shape
    z, dm = _case(orc, nx, ny, 40 + nx)
    z0, dm0 = z.copy(), dm.copy()
    got = hm.flow_simulation(z, 0.01, dm, iters, **kw)
    want = orc.flow_simulation(z, 0.01, dm, iters, **kw)
    assert np.array_equal(got, want), (shape, iters, kw)
    assert np.array_equal(z, z0) and np.array_equal(dm, dm0)


def test_flow_simulation_large_grid(hm, orc):
    z, dm = _case(orc, 1500, 1100, 3)
    got = hm.flow_simulation(z, 0.02, dm, 12, dry_out_ratio=0.05)
    want = orc.flow_simulation(z, 0.02, dm, 12, dry_out_ratio=0.05)
    assert np.array_equal(got, want)


def test_flow_simulation_flat_water(hm):
    z = np.full((200, 333), 0.25, np.float32)
    d = hm.flow_simulation(z, 0.05, np.ones_like(z), 50)
    assert np.array_equal(d, np.full_like(z, np.float32(0.05)))


def test_flow_simulation_rejects_bad_arguments(hm):
    z = np.zeros((8, 8), np.float32)
    with pytest.raises(hm.HmgError):
        hm.flow_simulation(z, 0.01, z, -1)
    with pytest.raises(hm.HmgError):
        hm.flow_simulation(z, 0.01, z, 3, dt=0.0)
    with pytest.raises(ValueError):
        hm.flow_simulation(z, 0.01, np.zeros((8, 9), np.float32), 3)
