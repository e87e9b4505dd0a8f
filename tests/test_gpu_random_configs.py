"""Seeded random sweep of the vpipes session against the FP32 oracle, bit
for bit: grid shapes from 3 x 3 up (widths on and off the 60-column strip
grid), every parameter of hmg_vpipes_params drawn over a range around
HighMap's defaults (flux diffusion on or off, volume top-up on or off, unit
and non-unit discharge exponents, gravity and per-iteration rain, a rain
plane), FP32 / BF16 / FP16 storage, fused and plain launches, a launch
shape drawn over every chunk length and variant, and thermal steps between
the hydraulic calls.  Each case is small; together they reach parameter
combinations the fixed-case tests do not."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("z", "d", "s", "u", "v")


def _case(k):
    rng = np.random.default_rng(1000 + k)
    nx = int(rng.choice([3, 4, 5, 17, 59, 60, 61, 64, 119, 120, 121, 128, 200, 257]))
    ny = int(rng.integers(3, 160))
    p = dict(water_height=float(10 ** rng.uniform(-4, -1)),
             evap_rate=float(rng.uniform(0.0, 0.1)),
             dt=float(rng.uniform(0.05, 1.0)),
             k_capacity=float(10 ** rng.uniform(-2, 0.3)),
             k_erode=float(rng.uniform(0.0, 0.1)),
             k_depose=float(rng.uniform(0.0, 0.1)),
             k_discharge_exp=float(rng.choice([1.0, 1.0, 0.8, 1.6])),
             downcutting_max_depth_ratio=float(rng.uniform(0.5, 20.0)),
             flux_diffusion=int(rng.integers(0, 2)),
             flux_diffusion_strength=float(rng.uniform(0.0, 0.2)),
             gravity=float(rng.choice([1.0, 9.81])),
             rain_per_iter=float(rng.choice([0.0, 0.0, 1e-3])),
             maintain_water_volume=int(rng.integers(0, 2)))
    storage = int(rng.integers(0, 3))
    mode = str(rng.choice(["fused", "plain"]))
    shape = (int(rng.integers(1, 100)), int(rng.integers(0, 3)))
    rain = bool(rng.integers(0, 4) == 0)
    steps = [int(rng.integers(1, 8)) for _ in range(int(rng.integers(1, 4)))]
    return nx, ny, p, storage, mode, shape, rain, steps, int(rng.integers(0, 2**31))


@pytest.mark.parametrize("k", range(120))
def test_random_session_equals_oracle(hm, orc, k):
    nx, ny, p, storage, mode, shape, rain, steps, seed = _case(k)
    z0 = orc.fbm_terrain(nx, ny, seed % 1000)
    r = None
    if rain:
        r = np.random.default_rng(seed).random((ny, nx), dtype=np.float32)
    s = hm.Session(z0, rain_map=r, mode=mode, storage=storage, **p)
    if mode == "fused":
        s.set_tuning(*shape)
    o = orc.VpState(z0, orc.vp_params(storage=storage, **p), r)
    talus = 1.0 / max(nx, ny)
    for n in steps:
        s.vpipes(n).thermal(talus, 1)
        o.run(n).thermal_step(talus, 1)
    g = s.export()
    s.close()
    for name, a, b in zip(NAMES, g, (o.z, o.d, o.s, o.u, o.v)):
        assert np.array_equal(a, b), "case %d %s: %d cells differ (nx %d ny %d storage %d %s %s)" % (
            k, name, int((a != b).sum()), nx, ny, storage, mode, p)
