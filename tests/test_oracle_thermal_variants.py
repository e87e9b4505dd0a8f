"""Pins of the oracle's thermal_auto_bedrock and sediment_deposition
restatements (CPU only).

No reference test covers these functions.  thermal_auto_bedrock is pinned by
an independent vectorised numpy restatement (bit for bit in FP32), by what its
bedrock rule implies (an interior cell never goes below its previous value,
flat fields and slopes under the talus are fixed points) and by a closed-form
first step on a spike.  sediment_deposition is pinned by closed forms on
fields that thermal leaves alone (flat, and a ramp under the talus, which
extrapolate_borders also reproduces) and by its composition from the pinned
thermal restatement.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from test_oracle_thermal import np_extrapolate, np_thermal_step

NB = [((-1, 0), 1.0), ((0, 1), 1.0), ((0, -1), 1.0), ((1, 0), 1.0),
      ((-1, -1), 1.414), ((-1, 1), 1.414), ((1, -1), 1.414), ((1, 1), 1.414)]


def np_auto_bedrock(z, talus, iterations):
    """thermal.cl:102-145 as vectorised Jacobi steps, then extrapolate_borders
    (thermal_gpu.cpp:98-126)."""
    dt = z.dtype.type
    z0 = z.copy()
    ny, nx = z.shape
    bed = np.zeros_like(z)
    cur = z.copy()
    for it in range(iterations):
        nxt = cur.copy()
        if it == 0:
            bed[1:-1, 1:-1] = -dt(np.finfo(np.float32).max)
        c = cur[1:-1, 1:-1]
        t = dt(talus) if np.isscalar(talus) else talus[1:-1, 1:-1]
        amount = np.zeros_like(c)
        for (di, dj), dist in NB:
            o = cur[1 + dj:ny - 1 + dj, 1 + di:nx - 1 + di]
            dist = dt(np.float32(dist))
            md = dist * t
            e = np.where(c > o,
                         np.where(c - o > md, -dt(np.float32(0.2)) * ((c - o) - md) / dist, dt(0)),
                         np.where(o - c > md, dt(np.float32(0.2)) * ((o - c) - md) / dist, dt(0)))
            amount = amount + e
        zb = bed[1:-1, 1:-1]
        z0c = z0[1:-1, 1:-1]
        act = c >= zb
        v = np.maximum(np.maximum(c + amount, z0c), zb)
        nxt[1:-1, 1:-1] = np.where(act, v, c)
        bed[1:-1, 1:-1] = np.where(act, np.maximum(z0c, v), zb)
        nxt[:, 0] = cur[:, 1]
        nxt[:, -1] = cur[:, -2]
        nxt[0, 1:-1] = cur[1, 1:-1]
        nxt[-1, 1:-1] = cur[-2, 1:-1]
        cur = nxt
    return np_extrapolate(cur)


@pytest.mark.parametrize("shape,iters", [((23, 31), 7), ((5, 5), 3), ((40, 9), 25)])
@pytest.mark.parametrize("plane", [False, True])
def test_auto_bedrock_numpy_restatement_fp32(shape, iters, plane):
    rng = np.random.default_rng(shape[0] + iters)
    z = rng.random(shape).astype(np.float32)
    talus = (rng.random(shape) * 0.05).astype(np.float32) if plane else np.float32(0.02)
    got = O.thermal_auto_bedrock(z, talus, iters)
    want = np_auto_bedrock(z, talus, iters)
    assert np.array_equal(got, want)


def test_auto_bedrock_never_lowers_the_interior():
    z = O.fbm_terrain(48, 40, 4)
    prev = z.copy()
    for k in range(1, 12):
        cur = O.thermal_auto_bedrock(z, 1.0 / 48, k)
        assert (cur[1:-1, 1:-1] >= prev[1:-1, 1:-1]).all()
        prev = O.thermal_auto_bedrock(z, 1.0 / 48, k)
    # while plain thermal does lower some cells
    plain = O.thermal(z, 1.0 / 48, 11)
    assert (plain[1:-1, 1:-1] < z[1:-1, 1:-1]).any()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_auto_bedrock_fixed_points(dtype):
    flat = np.full((9, 12), 0.4, dtype)
    assert np.array_equal(O.thermal_auto_bedrock(flat, 0.01, 20, dtype), flat)
    ramp = np.tile(0.005 * np.arange(12, dtype=dtype), (9, 1)).astype(dtype)
    got = O.thermal_auto_bedrock(ramp, 0.01, 20, dtype)
    assert np.allclose(got, ramp, rtol=0, atol=1e-6 if dtype == np.float32 else 1e-15)


def test_auto_bedrock_spike_first_step_closed_form():
    # a spike of height h on a flat zero field: the spike cannot go below its
    # own start, each neighbour k gains 0.2 * (h - c_k T) / c_k
    h, T = 1.0, 0.05
    z = np.zeros((7, 7), np.float64)
    z[3, 3] = h
    steps = z.copy()
    O.lib().orc_thermal_auto_bedrock_steps_f64(steps, 7, 7, None, T, 1)
    assert steps[3, 3] == h
    rate = float(np.float32(0.2))  # the kernel's float constants
    for (di, dj), c in NB:
        c = float(np.float32(c))
        assert abs(steps[3 + dj, 3 + di] - rate * (h - c * T) / c) < 1e-12
    assert steps[1, 1] == 0.0  # two cells away: unchanged


def test_sediment_deposition_flat_closed_form():
    z = np.full((10, 14), 0.25, np.float64)
    zz, dep = O.sediment_deposition(z, 0.01, 0.02, 5, 10, dtype=np.float64)
    step = float(np.float32(0.02) / np.float32(5))
    assert np.allclose(dep, 5 * step, rtol=0, atol=1e-15)
    assert np.allclose(zz, z + dep, rtol=0, atol=0)


@pytest.mark.parametrize("sub", [0, 10])
def test_sediment_deposition_ramp_under_talus(sub):
    # thermal leaves a ramp under the talus alone and extrapolate_borders
    # reproduces it, so every round adds exactly one step everywhere
    ny, nx = 12, 16
    z = (0.003 * np.arange(nx)[None, :] + 0.002 * np.arange(ny)[:, None]).astype(np.float64)
    zz, dep = O.sediment_deposition(z, 0.01, 0.03, 6, sub, dtype=np.float64)
    assert np.allclose(dep, float(np.float32(0.03) / np.float32(6)) * 6, rtol=0, atol=1e-12)


def test_sediment_deposition_composition_fp32():
    # the same rounds written with the pinned thermal restatement
    z = O.fbm_terrain(33, 21, 8)
    talus = np.float32(2.0) / np.float32(33)
    step = np.float32(0.01) / np.float32(4)
    smap = np.zeros_like(z)
    for _ in range(4):
        smap = smap + step
        zt = z + smap
        for _ in range(3):
            zt = np_thermal_step(zt, talus)
        zt = np_extrapolate(zt)
        d = zt - z
        smap = np.where(d < 0, np.float32(0), d)
    zz, dep = O.sediment_deposition(z, talus, 0.01, 4, 3)
    assert np.array_equal(dep, smap)
    assert np.array_equal(zz, z + smap)
