"""Pins of the oracle's flow_simulation restatement (CPU only).

No reference test covers gpu::flow_simulation.  The oracle is pinned by an
independent vectorised numpy restatement (bit for bit in FP32, 1e-12 in
FP64), closed forms (flat water stays put, the interior of a uniform slope
keeps its depth, zero iterations, the dry-out formula) and the reference's
boundary behaviour (clamp-to-edge reads turn a border cell's own outflow
into inflow, so the volume is not conserved at the borders).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from test_oracle_vpipes import nb_b, nb_l, nb_r, nb_t


def np_flow_simulation(z, wh, depth_map, iterations, dt=0.5, flux_diffusion=True,
                       strength=0.01, dry_out_ratio=0.0, dtype=np.float32):
    """flow_simulation.cpp:26-110 with the flow and water passes of
    hydraulic_vpipes.cl:5-116 and water_depth.cpp:40-64, vectorised."""
    f = lambda v: dtype(np.float32(v))
    z = z.astype(dtype)
    d = f(wh) * depth_map.astype(dtype)
    fl, fr, ft, fb = (np.zeros_like(z) for _ in range(4))
    dt_ = f(dt)
    c0 = f(strength)
    c1 = f(1) - f(4) * c0
    for _ in range(iterations):
        h0 = z + d
        l = np.maximum(f(0), fl + dt_ * ((h0 - nb_l(z)) - nb_l(d)))
        r = np.maximum(f(0), fr + dt_ * ((h0 - nb_r(z)) - nb_r(d)))
        t = np.maximum(f(0), ft + dt_ * ((h0 - nb_t(z)) - nb_t(d)))
        b = np.maximum(f(0), fb + dt_ * ((h0 - nb_b(z)) - nb_b(d)))
        if flux_diffusion:
            fd = c0 * (((l + r) + t) + b)
            l, r, t, b = c1 * l + fd, c1 * r + fd, c1 * t + fd, c1 * b + fd
        sm = ((l + r) + t) + b
        with np.errstate(divide="ignore", invalid="ignore"):
            k = np.where(sm > f(1e-5), d / (sm * dt_), f(0))
        k = np.minimum(np.maximum(k, f(0)), f(1))
        fl, fr, ft, fb = l * k, r * k, t * k, b * k
        frl, ftb, flr, fbt = nb_l(fr), nb_b(ft), nb_r(fl), nb_t(fb)
        dv = dt_ * (((((((frl + ftb) + flr) + fbt) - fl) - fr) - ft) - fb)
        d = np.maximum(f(0), d + dv)
    if dry_out_ratio != 0:
        d = np.maximum(f(0), d - f(dry_out_ratio) * d.max())
    return d


def _case(nx, ny, seed):
    z = O.fbm_terrain(nx, ny, seed) if min(nx, ny) > 2 else \
        np.random.default_rng(seed).random((ny, nx), dtype=np.float32)
    dm = np.random.default_rng(seed + 1).random((ny, nx), dtype=np.float32)
    return z, dm


@pytest.mark.parametrize("shape", [(5, 7), (8, 8), (1, 9), (9, 1), (33, 20)])
@pytest.mark.parametrize("iters,kw", [
    (0, {}), (1, {}), (7, dict(flux_diffusion=False)), (40, dict(dry_out_ratio=0.25)),
    (25, dict(dt=0.2, strength=0.05))])
def test_numpy_restatement_fp32_bitexact(shape, iters, kw):
    nx, ny = shape
    z, dm = _case(nx, ny, 11)
    kw_o = dict(kw)
    if "strength" in kw_o:
        kw_o["flux_diffusion_strength"] = kw_o.pop("strength")
    got = O.flow_simulation(z, 0.02, dm, iters, **kw_o)
    want = np_flow_simulation(z, 0.02, dm, iters, **kw)
    assert np.array_equal(got, want)


def test_numpy_restatement_fp64():
    z, dm = _case(40, 24, 5)
    got = O.flow_simulation(z, 0.01, dm, 30, dry_out_ratio=0.1, dtype=np.float64)
    want = np_flow_simulation(z, 0.01, dm, 30, dry_out_ratio=0.1, dtype=np.float64)
    assert np.allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_flat_water_is_a_fixed_point(dtype):
    z = np.full((12, 17), 0.3, np.float32)
    d = O.flow_simulation(z, 0.05, np.ones_like(z), 25, dtype=dtype)
    assert np.array_equal(d, np.full(z.shape, dtype(np.float32(0.05)) * dtype(1)))


def test_uniform_slope_interior_keeps_its_depth():
    # z = a*i: every cell but the first sends the same flux right-to-left and
    # gets as much from its right neighbour (the last one from itself, through
    # the clamped read), so the depth does not move from column 2 on
    nx, ny, a = 24, 10, 0.01
    z = np.tile(a * np.arange(nx, dtype=np.float64), (ny, 1))
    d = O.flow_simulation(z, 0.02, np.ones_like(z), 1, dtype=np.float64)
    wh = float(np.float32(0.02))
    assert np.allclose(d[:, 2:], wh, rtol=0, atol=1e-15)
    # column 0 has no lower neighbour: it keeps its water and gains column
    # 1's outflow, which column 1 loses
    assert (d[:, 0] > wh + 1e-3).all() and (d[:, 1] < wh - 1e-6).all()


def test_zero_iterations_and_dry_out_closed_form():
    z, dm = _case(16, 9, 3)
    d0 = O.flow_simulation(z, 0.5, dm, 0)
    assert np.array_equal(d0, np.float32(0.5) * dm)
    r = np.float32(0.3)
    d1 = O.flow_simulation(z, 0.5, dm, 0, dry_out_ratio=float(r))
    assert np.array_equal(d1, np.maximum(np.float32(0), d0 - r * d0.max()))


def test_volume_conserved_away_from_the_borders():
    # water moves at most one cell per iteration: a patch six cells from the
    # borders keeps its volume for three iterations; on the slope above the
    # border column gains water that the whole grid did not have
    z = np.zeros((16, 16), np.float64)
    dm = np.zeros_like(z)
    dm[6:10, 6:10] = 1.0
    dm[7, 8] = 3.0
    d = O.flow_simulation(z, 0.1, dm, 3, dtype=np.float64)
    v0 = float(np.float32(0.1)) * dm.sum()  # water_height is a float
    assert d[6:10, 6:10].sum() < v0 - 1e-3  # it did spread
    assert abs(d.sum() - v0) < 1e-12
