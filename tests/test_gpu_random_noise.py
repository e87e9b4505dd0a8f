"""Seeded random sweep of hmap::gpu::noise_fbm and hmap::gpu::noise against
the FP32 oracle, bit for bit: every implemented type, grid shapes from 1 x 1
up (widths on and off the 64-cell wavefront), wavenumbers from 0.5 to 200
of either sign, bounding boxes shifted, reversed and far from the origin,
periods (zero, small, large), 1 to 10 octaves, lacunarities 1.5 to 3 (and
negative), control planes, and warping planes for a quarter of the cases.
Without warping planes the wavefronts share their lattice values when the
lattice spans allow it and compute per cell otherwise, so the sweep mixes
both forms within one field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TYPES = (1, 2, 3, 4, 6, 7, 9, 10)


def _case(k):
    rng = np.random.default_rng(7000 + k)
    t = int(TYPES[k % len(TYPES)])
    nx = int(rng.choice([1, 2, 5, 63, 64, 65, 127, 130, 200, 333]))
    ny = int(rng.integers(1, 40))
    kw = tuple(float(np.float32(rng.choice([-1.0, 1.0]) * 10 ** rng.uniform(-0.3, 2.3)))
               for _ in range(2))
    lo = float(np.float32(rng.choice([0.0, -3.5, 1000.25, 123456.5])))
    width = float(np.float32(rng.choice([1.0, -1.0, 0.25, 3.0])))
    bbox = (lo, lo + width, float(np.float32(rng.uniform(-2, 2))),
            float(np.float32(rng.uniform(-2, 2))))
    period = tuple(int(rng.choice([0, 0, 1, 3, 7, 1000])) for _ in range(2))
    kw_fbm = dict(octaves=int(rng.integers(1, 11)), weight=float(np.float32(rng.uniform(0, 1))),
                  persistence=float(np.float32(rng.uniform(0.3, 0.8))),
                  lacunarity=float(np.float32(rng.choice([2.0, 2.0, 1.5, 3.0, -2.0]))),
                  bbox=bbox, period=period)
    ctrl = rng.random((ny, nx), dtype=np.float32) if rng.integers(0, 2) else None
    warp = None
    if rng.integers(0, 4) == 0:
        warp = tuple((rng.random((ny, nx), dtype=np.float32) - 0.5) * np.float32(0.2)
                     for _ in range(2))
    return t, nx, ny, kw, int(rng.integers(0, 2**31)), kw_fbm, ctrl, warp


@pytest.mark.parametrize("k", range(64))
def test_random_noise_equals_oracle(hm, orc, k):
    t, nx, ny, kw, seed, kw_fbm, ctrl, warp = _case(k)
    wk = dict(noise_x=warp[0], noise_y=warp[1]) if warp else {}
    g = hm.noise_fbm(t, (nx, ny), kw, seed, ctrl_param=ctrl, **wk, **kw_fbm)
    o = orc.noise_fbm(nx, ny, kw, seed, noise_type=t, ctrl=ctrl, **wk, **kw_fbm)
    same = (g.view(np.uint32) == o.view(np.uint32)) | (np.isnan(g) & np.isnan(o))
    assert same.all(), "case %d type %d %dx%d kw %s %s: %d cells differ" % (
        k, t, nx, ny, kw, kw_fbm, int((~same).sum()))
    g1 = hm.noise(t, (nx, ny), kw, seed, bbox=kw_fbm["bbox"], period=kw_fbm["period"], **wk)
    o1 = orc.noise(nx, ny, kw, seed, noise_type=t, bbox=kw_fbm["bbox"], period=kw_fbm["period"],
                   **wk)
    same = (g1.view(np.uint32) == o1.view(np.uint32)) | (np.isnan(g1) & np.isnan(o1))
    assert same.all(), "case %d noise type %d: %d cells differ" % (k, t, int((~same).sum()))
