"""Pins of the snow_simulation restatement (oracle/hm_oracle_erosion.inc,
HighMap/src/hydrology/snow_simulation.cpp:81-154 and
HighMap/src/gpu_opencl/kernels/snow_simulation.cl) and of the FP32
contract's exp, by what the mathematics fixes: libm's exp, rates set to
zero, single-term configurations with closed forms (a snow spike on flat
ground, a uniform ramp), the fall schedule, and the post-filter composed
from the pinned thermal restatement and numpy's gradient."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O

F = np.float32
DIAG = float(F(0.70710678))


# ------------------------------------------------------------------ exp

def test_exp_within_one_ulp_of_libm():
    xs = np.linspace(-87.0, 88.7, 40001, dtype=np.float32)
    xs = np.concatenate([xs, np.float32([0.0, 1e-8, -1e-8, 0.5, -0.5, 1.0, -1.0, 10.0, -10.0])])
    worst = 0.0
    for x in xs:
        got = O.expf(float(x))
        want = math.exp(float(x))
        worst = max(worst, abs(got - want) / float(np.spacing(F(want))))
    assert worst <= 1.0, worst


def test_exp_special_values():
    assert O.expf(0.0) == 1.0
    assert O.expf(-0.0) == 1.0
    assert math.isnan(O.expf(float("nan")))
    assert O.expf(float("inf")) == float("inf")
    assert O.expf(float("-inf")) == 0.0
    assert O.expf(89.5) == float("inf")
    assert O.expf(88.72) == pytest.approx(math.exp(float(F(88.72))), rel=2e-7)
    assert O.expf(-105.0) == 0.0
    # subnormal results, rounded once
    for x in (-88.0, -95.0, -100.0, -103.0):
        got, want = O.expf(x), math.exp(x)
        assert abs(got - want) <= 2 ** -149, x


def test_exp_monotonic():
    xs = np.linspace(-20.0, 20.0, 20001, dtype=np.float32)
    ys = np.array([O.expf(float(x)) for x in xs])
    assert np.all(np.diff(ys) >= 0)


# ------------------------------------------------------------- one step

def _field(n=24, seed=3):
    r = np.random.default_rng(seed)
    z = r.random((n, n), dtype=np.float32)
    s = (r.random((n, n), dtype=np.float32) * F(0.05)).astype(np.float32)
    return z, s


def test_step_identity_when_all_rates_are_zero():
    z, s = _field()
    got = O.snow_step(z, s, F(0.01), F(0.7), k_snow=0.0, k_melt=0.0, k_visc=0.0)
    assert np.array_equal(got, s)


def test_step_melting_alone():
    z, s = _field()
    melt = np.random.default_rng(9).random(z.shape, dtype=np.float32)
    k_melt, dt = 0.05, 0.5
    got = O.snow_step(z, s, F(0.01), melt, dt=dt, k_snow=0.0, k_melt=k_melt, k_visc=0.0)
    m = k_melt * melt.astype(np.float64) * dt
    want = np.maximum(s - np.minimum(m, s), 0.0)
    assert np.allclose(got, want, rtol=0, atol=2e-8)
    # melting more than is there leaves exactly nothing
    got = O.snow_step(z, s, F(0.01), 1.0, dt=dt, k_snow=0.0, k_melt=1.0, k_visc=0.0)
    assert np.all(got == 0.0)


def test_step_viscosity_alone():
    z, s = _field()
    kv = 0.3
    got = O.snow_step(z, s, F(0.01), 0.0, k_snow=0.0, k_visc=kv).astype(np.float64)
    sp = np.pad(s.astype(np.float64), 1, mode="edge")
    ax = sp[1:-1, :-2] + sp[1:-1, 2:] + sp[:-2, 1:-1] + sp[2:, 1:-1]
    dg = sp[:-2, :-2] + sp[:-2, 2:] + sp[2:, :-2] + sp[2:, 2:]
    avg = (ax + DIAG * dg) / (4 + 4 * DIAG)
    want = s + (avg - s) * kv
    assert np.allclose(got, want, rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("kv", [0.0, 0.25])
def test_step_spike_on_flat_ground(kv):
    """A snow spike S on flat ground: it keeps all its snow (its own slope
    is zero by symmetry); each axis neighbour receives k dt (S - T) /
    (1 + 2 diag) (the spike's drops to the neighbour, weight 1, and to the
    two cells beside the spike, weight diag each); diagonal neighbours and
    cells further away receive nothing, then viscosity blends towards the
    weighted neighbour mean."""
    n, c = 15, 7
    S, T, k, dt = 0.2, 0.01, 0.5, 0.5
    z = np.zeros((n, n), np.float32)
    s = np.zeros((n, n), np.float32)
    s[c, c] = F(S)
    got = O.snow_step(z, s, F(T), 0.0, dt=dt, snow_depth=0.1, k_snow=k, k_visc=kv)
    S64 = float(F(S))
    inflow = k * dt * (S64 - float(F(T))) / (1 + 2 * DIAG)
    axis = inflow + (S64 / (4 + 4 * DIAG) - inflow) * kv
    diag = kv * DIAG * S64 / (4 + 4 * DIAG)
    centre = S64 + (0.0 - S64) * kv
    assert got[c, c] == pytest.approx(centre, rel=1e-6)
    for dj, di in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        assert got[c + dj, c + di] == pytest.approx(axis, rel=2e-6)
    axis_vals = {got[c, c - 1], got[c, c + 1], got[c - 1, c], got[c + 1, c]}
    assert len(axis_vals) == 1  # the same operations in the same order
    for dj, di in ((-1, -1), (-1, 1), (1, -1), (1, 1)):
        assert got[c + dj, c + di] == pytest.approx(diag, rel=2e-6, abs=0)
    far = got.copy()
    far[c - 1:c + 2, c - 1:c + 2] = 0
    assert np.all(far == 0)


def test_step_uniform_ramp():
    """z = a i with uniform snow c: every interior cell loses
    min(k (a - T') dt, c) exp(-c / sc0) downhill and receives
    k (a - T') dt / 2 from its uphill neighbour, T' = T (1 + r
    smoothstep(0, sc0, c))."""
    n = 16
    a, c, T, k, dt, sd, kr, ksr = 0.02, 0.03, 0.005, 0.5, 0.5, 0.1, 1.0, 1.0
    z = (np.arange(n, dtype=np.float32) * F(a))[None, :].repeat(n, axis=0)
    s = np.full((n, n), F(c), np.float32)
    got = O.snow_step(z, s, F(T), 0.0, dt=dt, snow_depth=sd, k_snow=k, k_visc=0.0,
                      k_depth_ratio=kr, k_depth_slope_ratio=ksr)
    sc0 = float(F(kr) * F(sd))
    t = min(max(float(F(c)) / sc0, 0.0), 1.0)
    Te = float(F(T)) * (1 + ksr * t * t * (3 - 2 * t))
    A = float(F(a))
    out = min(k * (A - Te) * dt, float(F(c))) * math.exp(-float(F(c)) / sc0)
    want = float(F(c)) - out + k * (A - Te) * dt / 2
    inner = got[1:-1, 1:-1].astype(np.float64)
    assert np.allclose(inner, want, rtol=2e-5, atol=0)


# ---------------------------------------------------------- full call

def _sum_f32(c, times):
    v = F(0.0)
    for _ in range(times):
        v = F(v + c)
    return v


@pytest.mark.parametrize("iterations,ratio,n_fall", [(10, 1.0, 10), (10, 0.3, 3), (7, 0.0, 1),
                                                     (4, 2.0, 8)])
def test_fall_schedule_with_zero_rates(iterations, ratio, n_fall):
    """With every rate zero the snow is the fall alone: n_fall =
    max(1, int(ratio * iterations)) additions of snow_depth / n_fall (at
    most `iterations` of them), the same in every cell for a uniform fall
    map (fill and extrapolation keep a uniform field)."""
    z = np.random.default_rng(1).random((12, 10), dtype=np.float32)
    sd = 0.1
    got = O.snow_simulation(z, sd, 1.0, 0.0, 0.01, iterations, fall_iterations_ratio=ratio,
                            k_snow=0.0, k_visc=0.0, k_melt_factor=0.0, post_filter=False)
    c = F(F(sd) / F(n_fall))
    want = _sum_f32(c, min(n_fall, iterations))
    assert np.all(got == want)


def test_fall_map_interior_with_zero_rates():
    r = np.random.default_rng(2)
    z = r.random((12, 14), dtype=np.float32)
    fall = r.random((12, 14), dtype=np.float32)
    got = O.snow_simulation(z, 0.1, fall, 0.0, 0.01, 5, k_snow=0.0, k_visc=0.0,
                            k_melt_factor=0.0, post_filter=False)
    c = F(F(0.1) / F(5))
    want = np.zeros_like(fall)
    for _ in range(5):
        want = (want + c * fall).astype(np.float32)
    # border cells are filled from their inward neighbours, then extrapolated
    assert np.array_equal(got[1:-1, 1:-1], want[1:-1, 1:-1])


def test_post_filter_with_full_mask_is_the_thermal_run():
    """Snow of twice snow_depth makes the mask exactly 1, so the result is
    thermal(z + s, ratio * |grad(z + s)|, iterations) - z, clamped at 0; the
    gradient is numpy's (central inside, one-sided at the edges) and the
    thermal run the pinned thermal restatement."""
    r = np.random.default_rng(4)
    ny, nx, it, sd, ratio = 20, 24, 6, 0.05, 0.2
    z = (r.random((ny, nx), dtype=np.float32) * F(0.2)).astype(np.float32)
    got = O.snow_simulation(z, sd, 2.0, 0.0, 0.01, it, k_snow=0.0, k_visc=0.0,
                            k_melt_factor=0.0, post_filter=True, thermal_talus_ratio=ratio)
    s = np.full((ny, nx), _sum_f32(F(F(sd) / F(it)) * F(2.0), it), np.float32)
    assert np.all(s / F(sd) >= F(1.2))
    zw = (z + s).astype(np.float32)
    gy, gx = np.gradient(zw)
    talus = (F(ratio) * np.sqrt(gx * gx + gy * gy)).astype(np.float32)
    zf = O.thermal(zw, talus, it)
    want = (zf - z).astype(np.float32)
    want = np.where(want > 0, want, F(0))
    assert np.array_equal(got, want)


def test_fp32_contract_close_to_fp64():
    z = O.fbm_terrain(64, 64, 2)
    s32 = O.snow_simulation(z, 0.1, 1.0, 0.0, 2.5 / 64, 60)
    s64 = O.snow_simulation(z, 0.1, 1.0, 0.0, 2.5 / 64, 60, dtype=np.float64)
    assert np.abs(s32 - s64).max() <= 1e-5 * (s64.max() - s64.min() + 1e-12)


def test_snow_is_non_negative_and_melting_reduces_it():
    z = O.fbm_terrain(48, 40, 5)
    a = O.snow_simulation(z, 0.1, 1.0, 0.0, 2.5 / 48, 40, post_filter=False)
    b = O.snow_simulation(z, 0.1, 1.0, 1.0, 2.5 / 48, 40, post_filter=False)
    assert a.min() >= 0 and b.min() >= 0
    assert b.sum() < a.sum()
