"""Pins of the oracle's thermal restatement (CPU only).

No reference test covers gpu::thermal.  The oracle is pinned by fixed points,
a closed-form single step, conservation and symmetry properties the exchange
rule implies, and an independent vectorised numpy restatement.
"""
import numpy as np
import pytest

from oracle import pyoracle as O


def np_thermal_step(z, talus):
    """Vectorised Jacobi step written from thermal.cl:5-58 and
    _common_index.cl:106-132 (independent of the C loops)."""
    dt = z.dtype.type
    out = z.copy()
    c = z[1:-1, 1:-1]
    t = talus if np.isscalar(talus) else talus[1:-1, 1:-1]
    t = dt(t) if np.isscalar(t) else t
    nb = [((-1, 0), 1.0), ((0, 1), 1.0), ((0, -1), 1.0), ((1, 0), 1.0),
          ((-1, -1), 1.414), ((-1, 1), 1.414), ((1, -1), 1.414), ((1, 1), 1.414)]
    ny, nx = z.shape
    amount = np.zeros_like(c)
    for (di, dj), dist in nb:
        o = z[1 + dj:ny - 1 + dj, 1 + di:nx - 1 + di]
        dist = dt(np.float32(dist))
        md = dist * t
        e = np.where(c > o,
                     np.where(c - o > md, -dt(np.float32(0.2)) * ((c - o) - md) / dist, dt(0)),
                     np.where(o - c > md, dt(np.float32(0.2)) * ((o - c) - md) / dist, dt(0)))
        amount = amount + e
    out[1:-1, 1:-1] = c + amount
    # border copy rule, in the reference's test order
    out[:, 0] = z[:, 1]
    out[:, -1] = z[:, -2]
    out[0, 1:-1] = z[1, 1:-1]
    out[-1, 1:-1] = z[-2, 1:-1]
    return out


def np_extrapolate(z):
    z = z.copy()
    two = z.dtype.type(2)
    z[:, 0] = two * z[:, 1] - z[:, 2]
    z[:, -1] = two * z[:, -2] - z[:, -3]
    z[0, :] = two * z[1, :] - z[2, :]
    z[-1, :] = two * z[-2, :] - z[-3, :]
    return z


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_numpy_restatement_matches(dtype):
    rng = np.random.default_rng(0)
    z = rng.random((23, 31)).astype(dtype)
    talus = (rng.random((23, 31)) * 0.05).astype(dtype)
    a = z.copy()
    for _ in range(7):
        a = np_thermal_step(a, talus)
    a = np_extrapolate(a)
    b = O.thermal(z, talus, 7, dtype)
    if dtype == np.float32:
        assert np.array_equal(a, b)
    else:
        assert np.abs(a - b).max() < 1e-14


def test_flat_field_is_fixed_point():
    z = np.full((16, 20), 0.375, np.float32)
    assert np.array_equal(O.thermal(z, 0.01, 25), z)


def test_ramp_below_talus_is_fixed_point():
    # slope 2^-8 per cell < talus 2^-7: no exchange; borders restored exactly
    i = np.arange(32, dtype=np.float32)
    z = np.tile(i * np.float32(2.0 ** -8), (24, 1))
    z = z + np.arange(24, dtype=np.float32)[:, None] * np.float32(2.0 ** -9)
    assert np.array_equal(O.thermal(z, 2.0 ** -7, 10), z)


def test_single_spike_first_step_closed_form():
    h, T = 1.0, 0.05
    z = np.zeros((9, 9))
    z[4, 4] = h
    out = O.thermal_steps(z, T, 1, np.float64)
    c1, c2 = np.float64(np.float32(1.414)), np.float64(np.float32(0.2))
    T32 = T
    axial = c2 * (h - T32)
    diag = c2 * (h - c1 * T32) / c1
    assert abs(out[4, 4] - (h - 4 * axial - 4 * diag)) < 1e-14
    for dj, di in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        assert abs(out[4 + dj, 4 + di] - axial) < 1e-15
    for dj, di in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
        assert abs(out[4 + dj, 4 + di] - diag) < 1e-15
    out[3:6, 3:6] = 0
    assert np.all(out == 0)


def test_interior_mass_conserved_for_uniform_talus():
    rng = np.random.default_rng(4)
    z = np.zeros((64, 64))
    z[20:44, 20:44] = rng.random((24, 24))
    s0 = z.sum()
    out = O.thermal_steps(z, 0.01, 12, np.float64)  # disturbance stays > 2 cells from borders
    assert abs(out.sum() - s0) < 1e-12 * s0


def test_rotation_and_mirror_symmetry():
    # the border copy rule tests i before j, so corners break the symmetry:
    # keep the field zero within 8 cells of the border (6 steps)
    rng = np.random.default_rng(7)
    z = np.zeros((40, 40))
    z[8:32, 8:32] = rng.random((24, 24))
    a = O.thermal_steps(z, 0.02, 6, np.float64)
    for k in (1, 2, 3):
        b = O.thermal_steps(np.rot90(z, k).copy(), 0.02, 6, np.float64)
        assert np.abs(np.rot90(a, k) - b).max() < 1e-13
    b = O.thermal_steps(z[:, ::-1].copy(), 0.02, 6, np.float64)
    assert np.abs(a[:, ::-1] - b).max() < 1e-13


def test_extrapolate_borders_reproduces_linear_fields():
    jj, ii = np.mgrid[0:13, 0:17].astype(np.float32)
    z = ii * np.float32(0.25) - jj * np.float32(0.5) + np.float32(3)
    w = z.copy()
    w[0, :] = w[-1, :] = 99.0
    w[:, 0] = w[:, -1] = -99.0
    assert np.array_equal(O.extrapolate_borders(w), z)


def test_talus_plane_equals_scalar_when_uniform():
    z = O.fbm_terrain(48, 40, 3)
    t = np.full(z.shape, 1.0 / 48, np.float32)
    assert np.array_equal(O.thermal(z, t, 9), O.thermal(z, 1.0 / 48, 9))


def test_fp32_contract_against_fp64_oracle_on_256_workload():
    """256^2 fBm (seed 1), talus 2/256, 50 iterations: the FP32 contract stays
    within 1e-4 of the height range of the FP64 oracle."""
    z = O.fbm_terrain(256, 256, 1)
    a = O.thermal(z, 2.0 / 256, 50)
    b = O.thermal(z.astype(np.float64), 2.0 / 256, 50, np.float64)
    assert np.abs(a - b).max() <= 1e-4 * (b.max() - b.min())
