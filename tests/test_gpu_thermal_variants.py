"""thermal_auto_bedrock and sediment_deposition on the GPU against the FP32
oracle, bit for bit, including the masked overloads and deposition maps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _terrain(orc, nx, ny, seed):
    return orc.fbm_terrain(nx, ny, seed)


@pytest.mark.parametrize("shape,iters", [((3, 3), 4), ((64, 48), 0), ((130, 257), 33),
                                         ((300, 300), 120)])
@pytest.mark.parametrize("plane", [False, True])
def test_auto_bedrock_matches_oracle(hm, orc, shape, iters, plane):
    ny, nx = shape
    z0 = _terrain(orc, nx, ny, 5) if min(shape) > 3 else \
        np.random.default_rng(1).random(shape, dtype=np.float32)
    talus = (np.random.default_rng(2).random(shape, dtype=np.float32) * np.float32(0.02)
             if plane else np.float32(1.0) / np.float32(nx))
    z = z0.copy()
    dep = np.empty_like(z)
    hm.thermal_auto_bedrock(z, talus, iters, deposition_map=dep)
    want = orc.thermal_auto_bedrock(z0, talus, iters)
    assert np.array_equal(z, want)
    d = want - z0
    assert np.array_equal(dep, np.where(d < 0, np.float32(0), d))


def test_auto_bedrock_mask(hm, orc):
    z0 = _terrain(orc, 200, 120, 9)
    mask = np.tile(np.linspace(0, 1, 200, dtype=np.float32), (120, 1))
    z = z0.copy()
    hm.thermal_auto_bedrock(z, 1.0 / 200, 50, mask=mask)
    r = orc.thermal_auto_bedrock(z0, np.float32(1.0) / np.float32(200), 50)
    assert np.array_equal(z, z0 * (np.float32(1) - mask) + r * mask)


def test_thermal_device_post_keeps_results(hm, orc):
    # deposition and mask are formed on the device for thermal too
    z0 = _terrain(orc, 160, 96, 4)
    mask = np.random.default_rng(3).random(z0.shape, dtype=np.float32)
    z = z0.copy()
    dep = np.empty_like(z)
    hm.thermal(z, 1.0 / 160, 12, deposition_map=dep, mask=mask)
    r = orc.thermal(z0, np.float32(1.0) / np.float32(160), 12)
    d = r - z0
    assert np.array_equal(dep, np.where(d < 0, np.float32(0), d))
    assert np.array_equal(z, z0 * (np.float32(1) - mask) + r * mask)


@pytest.mark.parametrize("shape,kw", [((96, 64), dict()),
                                      ((257, 130), dict(max_deposition=0.05, iterations=7,
                                                        thermal_subiterations=3)),
                                      ((128, 128), dict(iterations=0)),
                                      ((64, 64), dict(thermal_subiterations=0))])
@pytest.mark.parametrize("plane", [False, True])
def test_sediment_deposition_matches_oracle(hm, orc, shape, kw, plane):
    ny, nx = shape
    z0 = _terrain(orc, nx, ny, 7)
    talus = (np.full(shape, np.float32(2.0) / np.float32(nx), np.float32) if plane
             else np.float32(2.0) / np.float32(nx))
    z = z0.copy()
    dep = np.empty_like(z)
    hm.sediment_deposition(z, talus, deposition_map=dep, **kw)
    wz, wd = orc.sediment_deposition(z0, talus, kw.get("max_deposition", 0.01),
                                     kw.get("iterations", 5), kw.get("thermal_subiterations", 10))
    assert np.array_equal(z, wz)
    assert np.array_equal(dep, wd)


def test_sediment_deposition_mask(hm, orc):
    z0 = _terrain(orc, 100, 80, 2)
    mask = np.random.default_rng(5).random(z0.shape, dtype=np.float32)
    z = z0.copy()
    hm.sediment_deposition(z, 0.02, mask=mask)
    wz, _ = orc.sediment_deposition(z0, np.float32(0.02))
    assert np.array_equal(z, z0 * (np.float32(1) - mask) + wz * mask)


def test_variants_reject_bad_arguments(hm):
    z = np.zeros((8, 8), np.float32)
    with pytest.raises(hm.HmgError):
        hm.thermal_auto_bedrock(z, 0.1, -1)
    with pytest.raises(hm.HmgError):
        hm.sediment_deposition(z, 0.1, iterations=-2)
    with pytest.raises(hm.HmgError):
        hm.sediment_deposition(np.zeros((2, 8), np.float32), 0.1)
