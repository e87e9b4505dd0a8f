"""Pins of the oracle's other noise types (CPU only): PERLIN_BILLOW,
PERLIN_HALF, SIMPLEX2, VALUE, VALUE_CUBIC, VALUE_LINEAR and WORLEY
(noise.cl:65-190, 385-454).

No reference test covers gpu::noise or these branches of gpu::noise_fbm.
Each base is pinned by a scalar float32 Python restatement written from the
kernel text (bit for bit at random points), by what the construction fixes
at lattice points (the value types interpolate their lattice hashes, cubic
rows pass through their middle samples, the Perlin variants are functions of
Perlin), by continuity across cell edges, bounds, and by the fBm closed form
octaves=1 -> 0.6 * base for every type.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O

F = np.float32
TYPES = {"perlin_billow": 2, "perlin_half": 3, "simplex2": 4, "value": 6, "value_cubic": 7,
         "value_linear": 9, "worley": 10}


def sinf(x):
    return F(O.lib().orc_sinf(F(x)))


def cosf(x):
    return F(O.lib().orc_cosf(F(x)))


ONE_MINUS = F(float.fromhex("0x1.fffffep-1"))


def fr(x):
    x = F(x)
    return min(F(x - F(math.floor(x))), ONE_MINUS)


def hash12(x, y, fseed):
    d = F(F(F(x) * F(127.1)) + F(F(y) * F(311.7)))
    return fr(F(sinf(F(d + F(fseed))) * F(43758.5453123)))


def hash22(x, y, fseed):
    a = F(F(F(x) * F(127.1)) + F(F(y) * F(311.7)))
    b = F(F(F(x) * F(269.5)) + F(F(y) * F(183.3)))
    return (fr(F(sinf(F(a + F(fseed))) * F(43758.5453123))),
            fr(F(sinf(F(b + F(fseed))) * F(43758.5453123))))


def grad(x, y, fseed):
    ang = F(F(6.2831853) * hash12(x, y, fseed))
    return cosf(ang), sinf(ang)


def lerp(a, b, t):
    return F(F(F(F(1) - F(t)) * F(a)) + F(F(t) * F(b)))


def py_simplex2(px, py, fseed):
    K1, K2 = F(0.366025403), F(0.211324865)
    x, y = F(F(0.5) * F(px)), F(F(0.5) * F(py))
    s = F(F(x * K1) + F(y * K1))
    ix, iy = F(math.floor(F(x + s))), F(math.floor(F(y + s)))
    t = F(F(ix * K2) + F(iy * K2))
    ax, ay = F(F(x - ix) + t), F(F(y - iy) + t)
    m = F(0) if ax < ay else F(1)
    ox, oy = m, F(F(1) - m)
    pts = [(ax, ay, ix, iy),
           (F(F(ax - ox) + K2), F(F(ay - oy) + K2), F(ix + ox), F(iy + oy)),
           (F(F(ax - F(1)) + F(F(2) * K2)), F(F(ay - F(1)) + F(F(2) * K2)), F(ix + F(1)),
            F(iy + F(1)))]
    tot = []
    for vx, vy, lx, ly in pts:
        w = max(F(F(0.5) - F(F(vx * vx) + F(vy * vy))), F(0))
        gx, gy = grad(lx, ly, fseed)
        tot.append(F(F(F(F(w * w) * w) * w) * F(F(gx * vx) + F(gy * vy))))
    return F(F(100) * F(F(tot[0] + tot[1]) + tot[2]))


def py_value(px, py, fseed, cubic_fade):
    ix, iy = F(math.floor(F(px))), F(math.floor(F(py)))
    fx, fy = fr(px), fr(py)
    v00, v10 = hash12(ix, iy, fseed), hash12(F(ix + F(1)), iy, fseed)
    v01, v11 = hash12(ix, F(iy + F(1)), fseed), hash12(F(ix + F(1)), F(iy + F(1)), fseed)
    ux, uy = fx, fy
    if cubic_fade:
        ux = F(F(fx * fx) * F(F(3) - F(F(2) * fx)))
        uy = F(F(fy * fy) * F(F(3) - F(F(2) * fy)))
    return F(F(F(2) * lerp(lerp(v00, v10, ux), lerp(v01, v11, ux), uy)) - F(1))


def cubic(p0, p1, p2, p3, t):
    a = F(F(F(p3 - p2) - p0) + p1)
    b = F(F(p0 - p1) - a)
    c = F(p2 - p0)
    return F(F(F(F(F(F(a * t) * t) * t) + F(F(b * t) * t)) + F(c * t)) + p1)


def py_value_cubic(px, py, fseed):
    ix, iy = F(math.floor(F(px))), F(math.floor(F(py)))
    fx, fy = fr(px), fr(py)
    rows = []
    for dy in range(-1, 3):
        v = [hash12(F(ix + F(dx)), F(iy + F(dy)), fseed) for dx in range(-1, 3)]
        rows.append(cubic(*v, fx))
    return F(F(1.43) * F(cubic(*rows, fy) - F(0.5)))


def py_worley(px, py, fseed):
    ix, iy = F(math.floor(F(px))), F(math.floor(F(py)))
    fx, fy = fr(px), fr(py)
    best = F(np.finfo(np.float32).max)
    for dx in range(-1, 2):
        for dy in range(-1, 2):
            hx, hy = hash22(F(ix + F(dx)), F(iy + F(dy)), fseed)
            ex, ey = F(fx - F(F(dx) + hx)), F(fy - F(F(dy) + hy))
            best = min(best, F(F(ex * ex) + F(ey * ey)))
    return F(F(F(1.66) * best) - F(1))


def py_max_smooth(a, b, k):
    h = F(max(F(F(k) - F(abs(F(a - b)))), F(0)) / F(k))
    return F(max(F(a), F(b)) + F(F(F(F(F(h * h) * h)) * F(k)) / F(6)))


def py_base(t, px, py, fseed):
    p = F(O.lib().orc_base_perlin(F(px), F(py), F(fseed), 0, 0))
    return {2: lambda: F(F(F(2) * F(abs(p))) - F(1)),
            3: lambda: py_max_smooth(p, F(0), F(0.5)),
            4: lambda: py_simplex2(px, py, fseed),
            6: lambda: py_value(px, py, fseed, True),
            7: lambda: py_value_cubic(px, py, fseed),
            9: lambda: py_value(px, py, fseed, False),
            10: lambda: py_worley(px, py, fseed)}[t]()


def _points(n=120, seed=0, scale=40.0):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 2)) * scale - scale / 4).astype(np.float32)


@pytest.mark.parametrize("name", sorted(TYPES))
def test_scalar_restatement_bitexact(name):
    t = TYPES[name]
    fseed = F(O.lib().orc_seed_to_fseed(7))
    for px, py in _points(seed=t):
        want = py_base(t, px, py, fseed)
        got = F(O.base_noise(t, px, py, fseed))
        assert got == want or (np.isnan(got) and np.isnan(want)), (name, px, py, got, want)


@pytest.mark.parametrize("name", ["value", "value_linear", "value_cubic"])
def test_value_types_hit_their_lattice_hashes(name):
    t = TYPES[name]
    fseed = F(0.3)
    for i in range(-3, 4):
        for j in range(-2, 3):
            h = hash12(F(i), F(j), fseed)
            v = F(O.base_noise(t, F(i), F(j), fseed))
            want = F(F(F(1.43) * F(h - F(0.5)))) if t == 7 else F(F(F(2) * h) - F(1))
            assert v == want


@pytest.mark.parametrize("name", sorted(TYPES))
def test_continuous_across_cell_edges(name):
    t = TYPES[name]
    fseed = F(0.77)
    eps = F(1e-3)
    jumps = []
    for e in range(-3, 4):
        for y in np.linspace(0.05, 0.95, 7, dtype=np.float32):
            a = O.base_noise(t, F(e) - eps, y, fseed)
            b = O.base_noise(t, F(e) + eps, y, fseed)
            jumps.append(abs(a - b))
    # every type here is at least C0 (worley: a min of continuous distances)
    assert max(jumps) < 0.05


def test_bounds():
    fseed = F(0.11)
    pts = _points(400, 3)
    # |perlin| <= 1.42857 sqrt(2)/2; each simplex term is at most
    # 100 max_r (0.5 - r^2)^4 r < 0.93
    pb = 2 * 1.42857 * math.sqrt(2) / 2 - 1 + 1e-6
    for name, (lo, hi) in {"perlin_billow": (-1, pb), "value": (-1, 1),
                           "value_linear": (-1, 1), "worley": (-1, None),
                           "simplex2": (-2.79, 2.79)}.items():
        vals = [O.base_noise(TYPES[name], x, y, fseed) for x, y in pts]
        assert min(vals) >= lo, name
        if hi is not None:
            assert max(vals) <= hi, name


def test_perlin_half_closed_form():
    # max_smooth(v, 0, 0.5): v for v >= 0.5, 0 for v <= -0.5, k/6 at v = 0
    assert py_max_smooth(F(0.7), F(0), F(0.5)) == F(0.7)
    assert py_max_smooth(F(-0.6), F(0), F(0.5)) == F(0)
    assert abs(float(py_max_smooth(F(0), F(0), F(0.5))) - 0.5 / 6) < 1e-7


@pytest.mark.parametrize("name", sorted(TYPES))
def test_fbm_one_octave_is_0_6_base_and_single_octave_kernel(name):
    t = TYPES[name]
    n1 = O.noise_fbm(24, 16, (3.0, 5.0), 9, octaves=1, noise_type=t)
    b = O.noise(24, 16, (3.0, 5.0), 9, noise_type=t)
    assert np.array_equal(n1, (F(0) + b * F(0.6)).astype(np.float32))


def test_simplex_ignores_the_period():
    a = O.noise_fbm(16, 16, (4.0, 4.0), 2, noise_type=4, period=(0, 0))
    b = O.noise_fbm(16, 16, (4.0, 4.0), 2, noise_type=4, period=(2, 2))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["value", "value_cubic", "value_linear", "worley",
                                  "perlin_billow"])
def test_period_makes_the_map_tile(name):
    # with kw equal to the period the lattice repeats across the map edge
    t = TYPES[name]
    n = 32
    a = O.noise_fbm(n, n, (4.0, 4.0), 5, octaves=3, noise_type=t, period=(4, 4),
                    bbox=(0.0, 2.0, 0.0, 1.0))
    assert np.allclose(a[:, :n // 2], a[:, n // 2:], atol=0)
