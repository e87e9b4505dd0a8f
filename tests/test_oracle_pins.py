"""Independent pins of oracle functions whose other checks are only
invariants: the erosion/deposition pass of hydraulic_vpipes, the noise
hashes hash12f / grad22f, and the contract's pow (end of the file).

Erosion pass (HighMap/src/gpu_opencl/kernels/hydraulic_vpipes.cl:149-183):
on a tilted plane every interior cell sees the same gradient, so the pass
has a closed form.  The planes below are built from dyadic numbers, so each
intermediate quantity is a short exact expression (3-4-5 triangles for the
slope and the speed); the expected values are written from those numbers,
not from the oracle's code.  k_erode != k_depose, nx != ny and
k_discharge_exp = 1.5 make a swapped branch, the wrong grid side or a
dropped exponent fail.

Hashes (HighMap/src/gpu_opencl/kernels/_common_rand.cl:12-17,53-57):
hash12f(p) = fract(sin(dot(p, (127.1, 311.7)) + fseed) * 43758.5453123) and
grad22f(p) = (cos a, sin a), a = 6.2831853 * hash12f(p), checked against
Python's float64 math.sin / math.cos at lattice points.  The dot product is
formed in float32 exactly as the kernel does (two rounded products, one
rounded sum); the float32 sin has an error of an ulp or two, which the
factor 43758 magnifies to a few 1e-3 in the fractional part: the tolerance
is 0.012 on the circle [0, 1).  A wrong constant moves the hash to an
unrelated value in [0, 1), so with 16 points a mistake passes with
probability below 1e-20.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O

F = np.float32

# ----------------------------------------------------------- erosion pass

NX, NY = 8, 5  # nx != ny: dzn scales the slope by nx, not ny


def _params(**kw):
    # water_height 1/128, k_capacity 0.5, k_erode 0.25, k_depose 0.125,
    # k_discharge_exp 1.5, downcutting_max_depth_ratio 10 (all exact floats)
    base = dict(water_height=1.0 / 128, k_capacity=0.5, k_erode=0.25, k_depose=0.125,
                k_discharge_exp=1.5, downcutting_max_depth_ratio=10.0)
    base.update(kw)
    return O.vp_params(**base)


def _planes(zx, zy, s0, sx, sy, d2, u, v):
    """z = (zx i + zy j) / 1024, s = s0 + (sx i + sy j) / 1024, uniform d2,
    u, v, as float64 planes of shape (NY, NX)."""
    j, i = np.mgrid[0:NY, 0:NX].astype(np.float64)
    z = (zx * i + zy * j) / 1024.0
    s = s0 + (sx * i + sy * j) / 1024.0
    full = lambda c: np.full((NY, NX), c, np.float64)  # noqa: E731
    return z, s, full(d2), full(u), full(v)


def _interior(a):
    return a[1:-1, 1:-1]


def _check(planes, params, expect_amount, erodes, rel64=1e-13, rel32=2e-6):
    """expect_amount: closed-form |z_out - z| on the interior; erodes: which
    branch (z falls and s rises, or the reverse)."""
    z, s = planes[0], planes[1]
    sign = -1.0 if erodes else 1.0
    for dtype, rel in ((np.float64, rel64), (np.float32, rel32)):
        pl = [p.astype(dtype) for p in planes]
        zo, so = O.vp_erosion(*pl, params, dtype=dtype)
        dz = _interior(zo.astype(np.float64) - pl[0].astype(np.float64))
        ds = _interior(so.astype(np.float64) - pl[1].astype(np.float64))
        np.testing.assert_allclose(dz, sign * expect_amount, rtol=rel, atol=0,
                                   err_msg="z change, %s" % dtype.__name__)
        np.testing.assert_allclose(ds, -sign * expect_amount, rtol=rel, atol=0,
                                   err_msg="s change, %s" % dtype.__name__)
    return z, s


def test_erosion_branch_closed_form():
    """Slope 3/1024 by 4/1024 (|grad| = 5/1024), dzn = nx * 5/1024 =
    0.0390625, sin(alpha) = dzn / sqrt(1 + dzn^2); speed |(0.375, 0.5)| =
    0.625; depth = (3/128) / (1/128) = 3; discharge 1.875; capacity
    0.5 * 1.875^1.5 * sin(alpha) > s = 0: erosion by k_erode * capacity."""
    salpha = 0.0390625 / math.sqrt(1.0 + 0.0390625 ** 2)
    capa = 0.5 * 1.875 ** 1.5 * salpha
    assert 0.05 < capa < 0.0502
    _check(_planes(3, 4, 0.0, 0, 0, 3.0 / 128, 0.375, 0.5), _params(), 0.25 * capa, True)


def test_deposition_branch_closed_form():
    """Same slope and flow with s = 0.25 above the capacity: deposition by
    k_depose * (s - capacity)."""
    salpha = 0.0390625 / math.sqrt(1.0 + 0.0390625 ** 2)
    capa = 0.5 * 1.875 ** 1.5 * salpha
    _check(_planes(3, 4, 0.25, 0, 0, 3.0 / 128, 0.375, 0.5), _params(),
           0.125 * (0.25 - capa), False)


def test_depth_ratio_clamp_closed_form():
    """d2 = 1 gives d2 / water_height = 128, clamped to the ratio 10:
    discharge 10 * 0.625 = 6.25, 6.25^1.5 = 15.625; capacity
    0.5 * 15.625 * sin(alpha) ~ 0.305 > s = 0.25: erosion."""
    salpha = 0.0390625 / math.sqrt(1.0 + 0.0390625 ** 2)
    capa = 0.5 * 15.625 * salpha
    assert capa > 0.25
    _check(_planes(3, 4, 0.25, 0, 0, 1.0, 0.375, 0.5), _params(), 0.25 * (capa - 0.25), True)
    # with a larger ratio the clamp moves: depth 128 * 0.625 = 80, 80^1.5
    capa2 = 0.5 * 80.0 ** 1.5 * salpha
    _check(_planes(3, 4, 0.25, 0, 0, 1.0, 0.375, 0.5),
           _params(downcutting_max_depth_ratio=1000.0), 0.25 * (capa2 - 0.25), True)


def test_sediment_enters_the_slope():
    """The slope is that of z + s: z slope (3, 4)/1024 plus s slope (1, -2)
    /1024 = (4, 2)/1024, |grad| = sqrt(20)/1024, dzn = 8 sqrt(20)/1024.
    Capacity is below s ~ 0.3: deposition, per cell (s varies)."""
    dzn = 8.0 * math.sqrt(20.0) / 1024.0
    salpha = dzn / math.sqrt(1.0 + dzn * dzn)
    capa = 0.5 * 1.875 ** 1.5 * salpha
    planes = _planes(3, 4, 0.3, 1, -2, 3.0 / 128, 0.375, 0.5)
    amount = 0.125 * (_interior(planes[1]) - capa)
    _check(planes, _params(), amount, False)


def test_flat_ground_uses_the_minimum_slope_and_sin_alpha():
    """Flat z and s: |grad| = 0, dzn = max(1e-3, 0) = 1e-3 and
    sin(alpha) = max(0.01, 1e-3 / sqrt(1 + 1e-6)) = 0.01 (both literals are
    float32 in the kernel; the FP64 oracle promotes the float32 values)."""
    salpha = float(F(0.01))
    capa = 0.5 * 1.875 ** 1.5 * salpha
    _check(_planes(0, 0, 0.0, 0, 0, 3.0 / 128, 0.375, 0.5), _params(), 0.25 * capa, True)


def test_unit_exponent_is_plain_discharge():
    """k_discharge_exp = 1: capacity 0.5 * 1.875 * sin(alpha)."""
    salpha = 0.0390625 / math.sqrt(1.0 + 0.0390625 ** 2)
    capa = 0.5 * 1.875 * salpha
    _check(_planes(3, 4, 0.0, 0, 0, 3.0 / 128, 0.375, 0.5), _params(k_discharge_exp=1.0),
           0.25 * capa, True)


# ------------------------------------------------------------------ hashes

# lattice points as the fBm uses them: small and large (octave 8 of kw 4
# reaches ~512), negative (bbox offsets), with the fseeds of seeds 1 and 2
POINTS = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (3.0, 7.0), (-2.0, 5.0), (17.0, -9.0),
          (64.0, 128.0), (255.0, 511.0), (300.0, 17.0), (-123.0, -77.0), (12.0, 401.0),
          (499.0, 3.0), (5.0, 5.0), (40.0, 2.0), (2.0, 40.0), (100.0, 100.0)]


def _circ(a, b):
    d = abs(a - b) % 1.0
    return min(d, 1.0 - d)


def _hash_expected(px, py, fseed):
    dp = F(F(F(px) * F(127.1)) + F(F(py) * F(311.7)))  # two products, one sum
    arg = F(dp + F(fseed))
    x = math.sin(float(arg)) * 43758.5453123
    return x - math.floor(x)


@pytest.mark.parametrize("seed", [1, 2])
def test_hash12f_known_answers(seed):
    fseed = O.lib().orc_seed_to_fseed(seed)
    bad = []
    for px, py in POINTS:
        got = O.hash12f(px, py, fseed)
        want = _hash_expected(px, py, fseed)
        assert 0.0 <= got < 1.0
        if _circ(got, want) > 0.012:
            bad.append((px, py, got, want))
    assert not bad, bad


def test_hash12f_fails_for_a_mistyped_constant():
    """The check above is sharp: a hash with 127.2 instead of 127.1 (or
    43758.5 instead of 43758.5453123) disagrees with it at most points."""
    fseed = O.lib().orc_seed_to_fseed(1)

    def wrong(px, py, a=127.1, m=43758.5453123):
        dp = F(F(F(px) * F(a)) + F(F(py) * F(311.7)))
        x = math.sin(float(F(dp + F(fseed)))) * m
        return x - math.floor(x)

    for kw in (dict(a=127.2), dict(m=43758.5)):
        miss = sum(_circ(wrong(px, py, **kw), O.hash12f(px, py, fseed)) > 0.012
                   for px, py in POINTS[1:])
        assert miss >= len(POINTS) - 4, (kw, miss)


@pytest.mark.parametrize("seed", [1, 2])
def test_grad22f_known_answers(seed):
    """grad22f = (cos a, sin a), a = float32(6.2831853 * hash12f): unit
    length, cosine first, the full turn (not pi)."""
    fseed = O.lib().orc_seed_to_fseed(seed)
    for px, py in POINTS:
        h = O.hash12f(px, py, fseed)
        a = float(F(F(6.2831853) * F(h)))
        gx, gy = O.grad22f(px, py, fseed)
        assert abs(gx - math.cos(a)) <= 4e-7 and abs(gy - math.sin(a)) <= 4e-7, (px, py)
        assert abs(gx * gx + gy * gy - 1.0) <= 1e-6


@pytest.mark.parametrize("seed", [1, 2])
def test_perlin_slope_at_a_lattice_point_is_its_gradient(seed):
    """base_perlin near lattice point c: the fade weights of the other three
    corners are O(f^3), so P(c + e x) = 1.42857 e gx(c) + O(e^2) (and the
    same with y): the corner's own gradient, scaled by the kernel's
    1.42857, enters with the offset from that corner."""
    fseed = O.lib().orc_seed_to_fseed(seed)
    eps = 1.0 / 1024
    for cx, cy in [(3.0, 7.0), (17.0, 9.0), (64.0, 128.0), (5.0, 5.0)]:
        gx, gy = O.grad22f(cx, cy, fseed)
        px = O.lib().orc_base_perlin(cx + eps, cy, fseed, 0, 0)
        py = O.lib().orc_base_perlin(cx, cy + eps, fseed, 0, 0)
        assert abs(px / eps - 1.42857 * gx) <= 0.01, (cx, cy, px / eps, gx)
        assert abs(py / eps - 1.42857 * gy) <= 0.01, (cx, cy, py / eps, gy)


# -------------------------------------------------------------------- pow
#
# The contract's pow (orc_powf, shared with the kernels as hm_powf) for
# k_discharge_exp != 1 in the erosion pass (hydraulic_vpipes.cl:167, OpenCL
# pow under -cl-fast-relaxed-math there).  Pinned against Python's float64
# math.pow, the C99 special cases (F.9.4.4) and exact powers.


def _ulps(got, want):
    """|got - want| in units of the float32 spacing at float32(want)."""
    w32 = np.float32(want)
    if w32 == 0.0:
        return abs(got) / 2.0 ** -149
    return abs(got - want) / float(np.spacing(np.abs(w32)))


@pytest.mark.parametrize("y", [0.5, 1.3, 1.5, 2.0, 2.5, 3.0, 0.1, 0.7, -0.7, -1.0, -2.5, 10.0])
def test_powf_within_one_ulp_of_pow(y):
    """Log-uniform x over every float exponent (subnormals included) and
    uniform x in [0, 4): the result is within 0.9 ulp of the float64 pow (a
    coefficient of either series or of log2(e) with a wrong digit or
    exponent moves results by many ulps at some x; that the kernels' copy
    gives the same bits is tests/test_gpu_arith_check.py's)."""
    y = float(F(y))
    rng = np.random.default_rng(int(abs(y) * 1000) + 7)
    xs = np.concatenate([F(2.0) ** rng.uniform(-149, 127, 4000).astype(F),
                         rng.uniform(0, 4, 4000).astype(F)])
    xs = xs[(xs > 0) & np.isfinite(xs)]
    got = O.powf_many(xs, np.full_like(xs, y))
    worst = 0.0
    for x, g in zip(xs.tolist(), got.tolist()):
        try:
            want = math.pow(x, y)
        except OverflowError:
            want = math.inf
        if want > float(np.finfo(F).max):
            assert g == math.inf, (x, y, g)
            continue
        worst = max(worst, _ulps(g, want))
    assert worst <= 0.9, (y, worst)


def test_powf_large_exponents():
    """|y| up to 128: the error grows with |y| (the log's error is scaled by
    it) and stays within 3 ulp."""
    rng = np.random.default_rng(3)
    for y in (40.0, -40.0, 127.9):
        y = float(F(y))
        xs = rng.uniform(0.5, 2.0, 3000).astype(F)
        got = O.powf_many(xs, np.full_like(xs, y))
        for x, g in zip(xs.tolist(), got.tolist()):
            want = math.pow(x, y)
            if want > float(np.finfo(F).max):
                assert g == math.inf
            else:
                assert _ulps(g, want) <= 3.0, (x, y, g, want)


def test_powf_exact_powers():
    for x, y, want in [(2.0, 3.0, 8.0), (2.0, -1.0, 0.5), (4.0, 0.5, 2.0), (0.25, 0.5, 0.5),
                       (2.0, 10.0, 1024.0), (2.0 ** -100, 0.5, 2.0 ** -50), (8.0, 1.0 / 3.0, None),
                       (3.0, 2.0, 9.0), (10.0, 2.0, 100.0), (1.5, 2.0, 2.25), (2.0, 127.0, 2.0 ** 127),
                       (2.0, -149.0, 2.0 ** -149), (0.5, 149.0, 2.0 ** -149)]:
        g = O.powf(x, y)
        if want is None:
            assert abs(g - 2.0) <= float(np.spacing(F(2.0))), g
        else:
            assert g == want, (x, y, g)


def test_powf_special_cases():
    inf, nan = math.inf, math.nan
    P = O.powf
    assert P(nan, 0.0) == 1.0 and P(nan, -0.0) == 1.0
    assert P(1.0, nan) == 1.0 and P(1.0, inf) == 1.0 and P(-1.0, inf) == 1.0
    assert math.isnan(P(nan, 1.5)) and math.isnan(P(2.0, nan))
    assert math.isnan(P(-2.0, 0.5)) and math.isnan(P(-0.3, 1.3))
    assert P(-2.0, 3.0) == -8.0 and P(-2.0, 2.0) == 4.0 and P(-2.0, -1.0) == -0.5
    assert P(0.0, 1.3) == 0.0 and P(0.0, -1.3) == inf
    assert math.copysign(1.0, P(-0.0, 3.0)) == -1.0 and P(-0.0, -3.0) == -inf
    assert math.copysign(1.0, P(-0.0, 2.0)) == 1.0 and P(-0.0, -2.0) == inf
    assert P(inf, 0.5) == inf and P(inf, -0.5) == 0.0
    assert P(-inf, 3.0) == -inf and P(-inf, 2.0) == inf and math.copysign(1.0, P(-inf, -3.0)) == -1.0
    assert P(0.5, inf) == 0.0 and P(0.5, -inf) == inf and P(2.0, inf) == inf and P(2.0, -inf) == 0.0
    assert P(2.0, 128.0) == inf and P(2.0, -150.0) == 0.0 and P(2.0, -149.5) == 2.0 ** -149
    assert P(1e30, 2.0) == inf and P(1e-30, 2.0) == 0.0
