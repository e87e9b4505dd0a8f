"""Every hmap::NoiseType through gpu::noise_fbm and gpu::noise on the GPU
against the FP32 oracle, bit for bit (types without a kernel branch: zeros)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL_TYPES = list(range(13))


@pytest.mark.parametrize("noise_type", ALL_TYPES)
def test_noise_fbm_every_type(hm, orc, noise_type):
    nx, ny = 193, 67
    g = hm.noise_fbm(noise_type, (nx, ny), (5.0, 3.0), 77)
    o = orc.noise_fbm(nx, ny, (5.0, 3.0), 77, noise_type=noise_type)
    assert np.array_equal(g, o), (noise_type, np.abs(g - o).max())


@pytest.mark.parametrize("noise_type", ALL_TYPES)
def test_noise_single_octave_every_type(hm, orc, noise_type):
    nx, ny = 160, 90
    g = hm.noise(noise_type, (nx, ny), (7.0, 4.0), 5)
    o = orc.noise(nx, ny, (7.0, 4.0), 5, noise_type=noise_type)
    assert np.array_equal(g, o), (noise_type, np.abs(g - o).max())


@pytest.mark.parametrize("noise_type", [2, 3, 4, 6, 7, 9, 10])
def test_noise_fbm_options(hm, orc, noise_type):
    nx, ny = 130, 77
    rng = np.random.default_rng(noise_type)
    ctrl = rng.random((ny, nx), dtype=np.float32)
    wx = (rng.random((ny, nx), dtype=np.float32) - 0.5) * 0.1
    wy = (rng.random((ny, nx), dtype=np.float32) - 0.5) * 0.1
    kw = dict(octaves=5, weight=0.3, persistence=0.6, lacunarity=2.0,
              bbox=(-0.5, 1.5, 0.25, 2.0), period=(4, 6))
    g = hm.noise_fbm(noise_type, (nx, ny), (4.0, 6.0), 42, ctrl_param=ctrl, noise_x=wx,
                     noise_y=wy, **kw)
    o = orc.noise_fbm(nx, ny, (4.0, 6.0), 42, ctrl=ctrl, noise_x=wx, noise_y=wy,
                      noise_type=noise_type, **kw)
    assert np.array_equal(g, o)
    g1 = hm.noise(noise_type, (nx, ny), (4.0, 6.0), 42, noise_x=wx, noise_y=wy,
                  bbox=kw["bbox"], period=kw["period"])
    o1 = orc.noise(nx, ny, (4.0, 6.0), 42, noise_type=noise_type, noise_x=wx, noise_y=wy,
                   bbox=kw["bbox"], period=kw["period"])
    assert np.array_equal(g1, o1)


def test_noise_fbm_large_simplex_and_worley(hm, orc):
    for t in (4, 10):
        g = hm.noise_fbm(t, (1024, 512), (16.0, 8.0), 123456789)
        o = orc.noise_fbm(1024, 512, (16.0, 8.0), 123456789, noise_type=t)
        assert np.array_equal(g, o), t
