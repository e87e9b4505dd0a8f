"""Pins of the oracle's noise path (CPU only).

The reference has no test, fixture or golden vector for gpu::noise_fbm, so
the oracle is pinned by values derived by hand, by properties the algorithm
fixes, and against numpy's float64 sin/cos for the project-written routines.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O


def test_wang_hash_known_answer_by_hand():
    # wang_hash(0): (0 ^ 61) ^ 0 = 61; *9 = 549; 549 ^ (549 >> 4) = 549 ^ 34 = 519;
    # 519 * 0x27d4eb2d mod 2^32 = 0xC0A8C83B; ^ (>> 15) = 0xC0A8C83B ^ 0x18151
    assert O.lib().orc_wang_hash(0) == 0xC0A9496A


def test_wang_hash_is_a_bijection_on_a_sample():
    # every step of the hash is invertible on uint32, so no two of these
    # 2^20 consecutive seeds may collide
    L = O.lib()
    outs = {L.orc_wang_hash(s) for s in range(1 << 16)}
    assert len(outs) == 1 << 16


def test_xorshift32_full_period():
    """(13, 17, 5) is one of Marsaglia's full-period triples: every nonzero
    state returns to itself after exactly 2^32 - 1 steps."""
    assert O.lib().orc_xorshift_period(1) == (1 << 32) - 1


def test_xorshift_step_known_answer_by_hand():
    # x = 1: x ^= x << 13 -> 0x2001; x ^= x >> 17 -> 0x2001; x ^= x << 5 -> 0x42021
    assert O.lib().orc_xorshift32(1) == 0x42021


def test_fseed_is_uint_times_2_pow_minus_32():
    L = O.lib()
    for seed in (0, 1, 2, 7, 123456789, 0xFFFFFFFF):
        st = L.orc_xorshift32(L.orc_wang_hash(seed))
        assert L.orc_seed_to_fseed(seed) == np.float32(np.float32(st) * np.float32(2.0 ** -32))


@pytest.mark.parametrize("lim,tol", [(2 * math.pi, 1.2e-7), (4096.0, 1.5e-7), (2.0 ** 18, 4e-7)])
def test_project_sin_cos_error(lim, tol):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-lim, lim, 20000), np.linspace(-lim, lim, 20001)])
    xs = xs.astype(np.float32)
    L = O.lib()
    s = np.array([L.orc_sinf(float(x)) for x in xs])
    c = np.array([L.orc_cosf(float(x)) for x in xs])
    x64 = xs.astype(np.float64)
    assert np.abs(s - np.sin(x64)).max() < tol
    assert np.abs(c - np.cos(x64)).max() < tol


def test_project_sin_odd_cos_even_exactly():
    L = O.lib()
    for x in np.random.default_rng(1).uniform(-1e4, 1e4, 2000).astype(np.float32):
        assert L.orc_sinf(float(-x)) == -L.orc_sinf(float(x))
        assert L.orc_cosf(float(-x)) == L.orc_cosf(float(x))


def test_project_sin_cos_symmetric_at_every_magnitude():
    """Beyond 2^22 the reduction is no longer accurate (the hashes only need
    it deterministic), but its quadrant is k mod 4 for every k, so sin stays
    odd and cos even bit for bit over the whole float range: a quadrant
    taken by a saturating or wrapping float-to-int conversion (3 for k at
    or past 2^31, 0 for -k) breaks the symmetry from 2^31 pi/2 up.
    Every 997th float pattern from the smallest subnormal (zero aside: the
    reduction's fma turns -0 into +0), and a dense band over [3e9, 1e10]."""
    pat = np.arange(1, 0x7f800000, 997, dtype=np.uint64).astype(np.uint32).view(np.float32)
    lo, hi = np.array([3e9, 1e10], dtype=np.float32).view(np.uint32)
    band = np.arange(lo, hi, 7, dtype=np.uint64).astype(np.uint32).view(np.float32)
    x = np.concatenate([pat, band])
    s, c = O.sincos_many(x)
    sn, cn = O.sincos_many(-x)
    ok = lambda a, b: (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    bad = np.flatnonzero(~(ok(sn, -s) & ok(cn, c)))
    assert bad.size == 0, (x[bad[:5]], s[bad[:5]], sn[bad[:5]])


def test_fract_matches_opencl_definition():
    L = O.lib()
    assert L.orc_fractf(-0.25) == 0.75
    assert L.orc_fractf(3.5) == 0.5
    # x - floor(x) rounds to 1.0 here; OpenCL clamps to the largest float < 1
    assert L.orc_fractf(-1e-10) == np.float32(np.nextafter(np.float32(1), np.float32(0)))


def test_perlin_zero_at_lattice_points():
    L = O.lib()
    fseed = L.orc_seed_to_fseed(3)
    for i in range(-5, 40, 3):
        for j in range(-7, 30, 4):
            assert L.orc_base_perlin(float(i), float(j), fseed, 0, 0) == 0.0


def test_perlin_bounded_and_continuous():
    L = O.lib()
    fseed = L.orc_seed_to_fseed(5)
    rng = np.random.default_rng(2)
    pts = rng.uniform(-50, 50, (20000, 2)).astype(np.float32)
    v = np.array([L.orc_base_perlin(float(x), float(y), fseed, 0, 0) for x, y in pts])
    # unit gradients: |P| <= sqrt(2)/2 before the 1.42857 scale
    assert np.abs(v).max() <= 1.42857 * math.sqrt(0.5) + 1e-5
    assert v.std() > 0.1
    # continuity across a lattice edge
    for k in range(-3, 4):
        for y in (0.3, 1.7, 5.25):
            a = L.orc_base_perlin(k - 1e-4, y, fseed, 0, 0)
            b = L.orc_base_perlin(k + 1e-4, y, fseed, 0, 0)
            assert abs(a - b) < 5e-3


def test_perlin_tiles_with_period():
    L = O.lib()
    fseed = L.orc_seed_to_fseed(9)
    for x in (0.25, 1.5, 2.75):
        for y in (0.125, 3.5):
            base = L.orc_base_perlin(x, y, fseed, 4, 3)
            assert L.orc_base_perlin(x + 4.0, y, fseed, 4, 3) == base
            assert L.orc_base_perlin(x, y + 3.0, fseed, 4, 3) == base
            assert L.orc_base_perlin(x + 8.0, y + 6.0, fseed, 4, 3) == base


def _positions(nx, ny, kx, ky):
    i = np.arange(nx, dtype=np.float32)
    j = np.arange(ny, dtype=np.float32)
    x = np.float32(kx) * ((i / np.float32(nx)) * np.float32(1) + np.float32(0)) + np.float32(kx) * np.float32(0)
    y = np.float32(ky) * ((j / np.float32(ny)) * np.float32(1) + np.float32(0)) + np.float32(ky) * np.float32(0)
    return x, y


def test_fbm_one_octave_is_0_6_perlin():
    nx, ny, seed = 24, 17, 11
    out = O.noise_fbm(nx, ny, (4.0, 3.0), seed, octaves=1)
    L = O.lib()
    fseed = L.orc_seed_to_fseed(seed)
    x, y = _positions(nx, ny, 4.0, 3.0)
    for jj in range(ny):
        for ii in range(nx):
            p = L.orc_base_perlin(float(x[ii]), float(y[jj]), fseed, 0, 0)
            assert out[jj, ii] == np.float32(np.float32(p) * np.float32(0.6))


def test_fbm_two_octaves_closed_form():
    """n = 0.6 P(p) + 0.6 * ((1 - w) + w * min(P(p) + 1, 2) / 2) * pers * P(2p)."""
    nx, ny, seed, w, pers = 16, 16, 4, 0.3, 0.6
    out = O.noise_fbm(nx, ny, (4.0, 4.0), seed, octaves=2, weight=w, persistence=pers)
    L = O.lib()
    fseed = L.orc_seed_to_fseed(seed)
    x, y = _positions(nx, ny, 4.0, 4.0)
    for jj in range(ny):
        for ii in range(nx):
            p1 = L.orc_base_perlin(float(x[ii]), float(y[jj]), fseed, 0, 0)
            p2 = L.orc_base_perlin(float(x[ii] * 2), float(y[jj] * 2), fseed, 0, 0)
            ref = 0.6 * p1 + 0.6 * ((1 - w) + w * min(p1 + 1, 2) * 0.5) * pers * p2
            assert abs(out[jj, ii] - ref) < 1e-6


def test_fbm_weight_zero_amplitudes():
    nx, ny, seed = 12, 9, 21
    out = O.noise_fbm(nx, ny, (4.0, 4.0), seed, octaves=4, weight=0.0, persistence=0.5)
    L = O.lib()
    fseed = L.orc_seed_to_fseed(seed)
    x, y = _positions(nx, ny, 4.0, 4.0)
    for jj in range(ny):
        for ii in range(nx):
            ref = sum(0.6 * 0.5 ** k * L.orc_base_perlin(float(x[ii] * 2 ** k),
                                                        float(y[jj] * 2 ** k), fseed, 0, 0)
                      for k in range(4))
            assert abs(out[jj, ii] - ref) < 1e-6


def test_fbm_period_makes_the_map_tile():
    # period = kw tiles the full domain: column nx would equal column 0
    nx, ny = 32, 32
    a = O.noise_fbm(nx, ny, (4.0, 4.0), 3, octaves=3, period=(4, 4),
                    bbox=(0.0, 2.0, 0.0, 1.0))
    # x in [0, 8): the second half repeats the first half exactly (x + 4 has
    # the same fraction for these dyadic positions)
    assert np.array_equal(a[:, : nx // 2], a[:, nx // 2:])


@pytest.mark.parametrize("noise_type", [0, 5, 8, 11, 12])
def test_types_without_a_kernel_give_zeros(noise_type):
    # PARBERRY, SIMPLEX2S, VALUE_DELAUNAY, WORLEY_DOUBLE, WORLEY_VALUE have no
    # branch in the noise kernels (noise.cl:415-452, 494-576): zeros
    assert not O.lib().orc_noise_type_supported(noise_type)
    assert np.all(O.noise_fbm(8, 8, (4.0, 4.0), 3, noise_type=noise_type) == 0)
    assert np.all(O.noise(8, 8, (4.0, 4.0), 3, noise_type=noise_type) == 0)


def test_remap_range_and_flat():
    a = O.remap(np.array([[2.0, 4.0, 3.0]], np.float32))
    assert a.min() == 0.0 and a.max() == 1.0 and a[0, 2] == 0.5
    assert np.all(O.remap(np.full((3, 3), 7.0, np.float32)) == 0.0)
