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


# Without warping planes every type shares each wavefront's lattice values
# (64 cells of one row; Perlin and VALUE / VALUE_LINEAR at most 31 lattice
# columns, VALUE_CUBIC 13, WORLEY 19, SIMPLEX2 15 columns and 3 rows of its
# skewed lattice; |index| < 2^23); other waves compute per cell.  Both must give the oracle's bits: partial
# wavefronts at the right edge, periods, a reversed bounding box and a
# negative wavenumber (lattice index decreasing along the row), frequencies
# where only some octaves fit the shared form, spans on either side of the
# VALUE_CUBIC and WORLEY limits, huge coordinates (hash arguments past
# 2^31 pi/2, where sin's quadrant must still be k mod 4), periods that
# saturate past 2^31 after an octave, and a negative lacunarity.
@pytest.mark.parametrize("noise_type", [1, 2, 3, 4, 6, 7, 9, 10])
@pytest.mark.parametrize("case", [
    dict(shape=(130, 77), kw=(4.0, 6.0), octaves=5, bbox=(-0.5, 1.5, 0.25, 2.0), period=(4, 6)),
    dict(shape=(193, 5), kw=(5.0, 3.0), octaves=8, bbox=(1.0, 0.0, 1.0, 0.0), period=(0, 0)),
    dict(shape=(64, 3), kw=(-7.0, 2.0), octaves=6, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)),
    dict(shape=(65, 1), kw=(3.0, 3.0), octaves=3, bbox=(0.0, 1.0, 0.0, 1.0), period=(3, 0)),
    dict(shape=(256, 33), kw=(96.0, 8.0), octaves=4, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)),
    dict(shape=(300, 9), kw=(2.0, 2.0), octaves=4, bbox=(4194290.0, 4194310.0, 0.0, 1.0),
         period=(0, 0)),
    dict(shape=(150, 20), kw=(4.0, 4.0), octaves=4, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0),
         lacunarity=-2.0),
    dict(shape=(64, 4), kw=(13.0, 2.0), octaves=2, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)),
    dict(shape=(64, 4), kw=(14.0, 2.0), octaves=1, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)),
    dict(shape=(128, 6), kw=(19.0, 3.0), octaves=2, bbox=(0.25, 1.25, 0.0, 1.0), period=(5, 3)),
    dict(shape=(64, 4), kw=(19.0, 2.0), octaves=1, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)),
    dict(shape=(64, 4), kw=(19.5, 2.0), octaves=1, bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)),
    dict(shape=(130, 7), kw=(4.0, 4.0), octaves=3, bbox=(-2.0, 1.0, -1.0, 1.0),
         period=(2 ** 30, 2 ** 29 + 1)),
])
def test_noise_fbm_shared_lattice(hm, orc, noise_type, case):
    nx, ny = case["shape"]
    rng = np.random.default_rng(nx * ny)
    ctrl = rng.random((ny, nx), dtype=np.float32)
    kw = dict(octaves=case["octaves"], weight=0.4, persistence=0.5,
              lacunarity=case.get("lacunarity", 2.0), bbox=case["bbox"], period=case["period"])
    g = hm.noise_fbm(noise_type, (nx, ny), case["kw"], 9, ctrl_param=ctrl, **kw)
    o = orc.noise_fbm(nx, ny, case["kw"], 9, ctrl=ctrl, noise_type=noise_type, **kw)
    assert np.array_equal(g, o, equal_nan=True), (case, np.nanmax(np.abs(g - o)))
    # gpu::noise (one octave, the same wavefront forms)
    g1 = hm.noise(noise_type, (nx, ny), case["kw"], 9, bbox=case["bbox"], period=case["period"])
    o1 = orc.noise(nx, ny, case["kw"], 9, noise_type=noise_type, bbox=case["bbox"],
                   period=case["period"])
    assert np.array_equal(g1, o1, equal_nan=True), (case, "noise")
