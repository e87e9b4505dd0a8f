"""Seeded random sweep of hmap::gpu::thermal and thermal_auto_bedrock
against the FP32 oracle, bit for bit: shapes from 3 x 3 to grids above
2^18 cells (the LDS-blocked kernel below, the streaming one- and two-step
kernels above), step counts from 0 to 11 (odd and even, so the two-step
launches end with a one-step launch or with the border launches), uniform
talus (zero, negative, small and large against the terrain's slopes) and
talus planes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(3, 3), (4, 7), (33, 9), (62, 5), (63, 64), (124, 31), (200, 150), (257, 255),
          (512, 512), (600, 437), (1024, 300), (3, 90000)]


def _case(k):
    rng = np.random.default_rng(5000 + k)
    nx, ny = SHAPES[int(rng.integers(0, len(SHAPES)))]
    iters = int(rng.integers(0, 12))
    kind = int(rng.integers(0, 4))
    scale = 1.0 / max(nx, ny)
    if kind == 0:
        talus = float(np.float32(rng.choice([0.0, -1e-3, 0.5 * scale, 2.0 * scale, 40.0 * scale])))
    else:
        talus = (rng.random((ny, nx), dtype=np.float32) * np.float32(3.0 * scale)).astype(np.float32)
    auto = bool(rng.integers(0, 3) == 0)
    return nx, ny, iters, talus, auto, int(rng.integers(0, 1000))


@pytest.mark.parametrize("k", range(60))
def test_random_thermal_equals_oracle(hm, orc, k):
    nx, ny, iters, talus, auto, seed = _case(k)
    z0 = orc.fbm_terrain(nx, ny, seed)
    z = z0.copy()
    if auto:
        hm.thermal_auto_bedrock(z, talus, iters)
        want = orc.thermal_auto_bedrock(z0, talus, iters)
    else:
        hm.thermal(z, talus, iters)
        want = orc.thermal(z0, talus, iters)
    assert np.array_equal(z, want), "case %d: %d cells differ (%d x %d, %d steps, %s talus, auto %s)" % (
        k, int((z != want).sum()), nx, ny, iters, "scalar" if np.isscalar(talus) else "plane", auto)
