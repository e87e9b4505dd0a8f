"""Pins of the oracle's hydraulic_vpipes restatement (CPU only).

No reference test covers gpu::hydraulic_vpipes.  The oracle is pinned by
closed forms (flat water, one step on a tilted plane, zero-velocity
transport), invariants (per-cell z+s conservation of the erosion pass, the
maintained water volume), an independent mirror-repeat formulation, exact
integer checks of the fixed-point volume sum, and an independent vectorised
numpy restatement compared bit for bit (FP32) and to 1e-12 (FP64).
"""
from fractions import Fraction

import numpy as np
import pytest

from conftest import VP_WORKLOAD
from oracle import pyoracle as O

# --------------------------------------------------------------- helpers


def fix_int(x):
    """round-half-even(x * 2^64) for a float32 x, as a Python int."""
    return round(Fraction(float(np.float32(x))) * (1 << 64))


def fix_to_double(S):
    S &= (1 << 128) - 1
    lo = S & ((1 << 64) - 1)
    hi = S >> 64
    if hi >= 1 << 63:
        hi -= 1 << 64
    hid = float(hi >> 32) * 4294967296.0 + float(hi & 0xFFFFFFFF)
    lod = float(lo >> 32) * 4294967296.0 + float(lo & 0xFFFFFFFF)
    return hid + lod * 2.0 ** -64


def fixsum_py(a):
    return fix_to_double(sum(fix_int(x) for x in np.asarray(a).reshape(-1)))


# clamp-to-edge neighbour views
def nb_l(a):
    return np.concatenate([a[:, :1], a[:, :-1]], axis=1)


def nb_r(a):
    return np.concatenate([a[:, 1:], a[:, -1:]], axis=1)


def nb_t(a):
    return np.concatenate([a[1:], a[-1:]], axis=0)


def nb_b(a):
    return np.concatenate([a[:1], a[:-1]], axis=0)


def mirror_axis(p, w, dt):
    fw = dt(w)
    sg = p / fw
    sp = np.abs(sg - dt(2) * np.rint(dt(0.5) * sg))
    q = sp * fw - dt(0.5)
    q0 = np.floor(q)
    a = q - q0
    k = q0.astype(np.int64)
    return np.clip(k, 0, w - 1), np.clip(k + 1, 0, w - 1), a


def np_vpipes(z, p, iters, dtype, rain=None):
    """Vectorised restatement of hydraulic_vpipes_gpu.cpp:34-177 with the
    kernels of hydraulic_vpipes.cl, written independently of the C loops."""
    dt_ = dtype
    f = lambda v: dt_(np.float32(v))  # reference constants are floats
    ny, nx = z.shape
    z = z.astype(dtype).copy()
    rain_a = np.ones_like(z) if rain is None else rain.astype(dtype)
    d = f(p["water_height"]) * rain_a
    s = np.zeros_like(z)
    fl, fr, ft, fb = (np.zeros_like(z) for _ in range(4))
    u = v = np.zeros_like(z)
    if dtype == np.float32:
        W0, R = fixsum_py(d), fixsum_py(rain_a)
    else:
        W0, R = d.sum(), rain_a.sum()
    dt, g = f(p["dt"]), f(p.get("gravity", 1.0))
    dtg = dt * g
    c0 = f(p["flux_diffusion_strength"])
    c1 = f(1) - f(4) * c0
    wh = f(p["water_height"])
    ce = f(1) - dt * f(p["evap_rate"])
    ii = np.arange(nx, dtype=dtype)[None, :]
    jj = np.arange(ny, dtype=dtype)[:, None]
    for _ in range(iters):
        d1 = d.copy()
        # flow pass
        h0 = z + d1
        l = np.maximum(f(0), fl + dtg * ((h0 - nb_l(z)) - nb_l(d1)))
        r = np.maximum(f(0), fr + dtg * ((h0 - nb_r(z)) - nb_r(d1)))
        t = np.maximum(f(0), ft + dtg * ((h0 - nb_t(z)) - nb_t(d1)))
        b = np.maximum(f(0), fb + dtg * ((h0 - nb_b(z)) - nb_b(d1)))
        if p.get("flux_diffusion", 1):
            fd = c0 * (((l + r) + t) + b)
            l, r, t, b = c1 * l + fd, c1 * r + fd, c1 * t + fd, c1 * b + fd
        sm = ((l + r) + t) + b
        with np.errstate(divide="ignore", invalid="ignore"):
            k = np.where(sm > f(1e-5), d1 / (sm * dt), f(0))
        k = np.minimum(np.maximum(k, f(0)), f(1))
        fl, fr, ft, fb = l * k, r * k, t * k, b * k
        # water pass
        frl, ftb, flr, fbt = nb_l(fr), nb_b(ft), nb_r(fl), nb_t(fb)
        dv = dt * (((((((frl + ftb) + flr) + fbt) - fl) - fr) - ft) - fb)
        d2 = np.maximum(f(0), d1 + dv)
        dmean = np.maximum(f(0.001) * wh / dt, d2)
        u = (f(0.5) * (((frl - fl) + fr) - flr)) / dmean
        v = (f(0.5) * (((ftb - fb) + ft) - fbt)) / dmean
        # erosion pass
        dzx = f(0.5) * (((nb_r(z) - nb_l(z)) + nb_r(s)) - nb_l(s))
        dzy = f(0.5) * (((nb_t(z) - nb_b(z)) + nb_t(s)) - nb_b(s))
        talus = np.sqrt(dzx * dzx + dzy * dzy)
        dzn = np.maximum(f(1e-3), dt_(np.float32(nx)) * talus)
        salpha = np.maximum(f(0.01), dzn / np.sqrt(f(1) * f(1) + dzn * dzn))
        speed = np.sqrt(u * u + v * v)
        depth = np.minimum(d2 / wh, f(p["downcutting_max_depth_ratio"]))
        disc = depth * speed
        capa = (f(p["k_capacity"]) * disc) * salpha
        ero = capa > s
        amt = np.where(ero, f(p["k_erode"]) * (capa - s), f(p["k_depose"]) * (s - capa))
        z, s = np.where(ero, z - amt, z + amt), np.where(ero, s + amt, s - amt)
        # transport pass
        i0, i1, a = mirror_axis(ii - dt * u, nx, dtype)
        j0, j1, bb = mirror_axis(jj - dt * v, ny, dtype)
        one = f(1)
        s = ((((one - a) * (one - bb)) * s[j0, i0] + (a * (one - bb)) * s[j0, i1])
             + ((one - a) * bb) * s[j1, i0]) + (a * bb) * s[j1, i1]
        # evaporation and top-up
        d = d2 * ce
        if p.get("maintain_water_volume", 1):
            if dtype == np.float32:
                wta = np.float32(W0 - fixsum_py(d))
                Rf = np.float32(R)
            else:
                wta, Rf = W0 - d.sum(), R
            d = d + (wta * rain_a) / Rf
    return z, d, s, u, v


def _params(**kw):
    p = dict(O.VP_DEFAULTS)
    p.update(kw)
    return p


# ------------------------------------------------------------------ tests


def test_fixsum_exact_against_python_integers():
    rng = np.random.default_rng(1)
    cases = [
        rng.random(1000, dtype=np.float32) * 1e-3,
        (rng.standard_normal(500) * 10).astype(np.float32),
        np.array([0.0, -0.0, 1e-45, -1e-45, 2.0 ** -41, 2.0 ** -42, 3.0 * 2.0 ** -66,
                  2.0 ** 30, -2.0 ** 30, 1.5, 7.0 * 2.0 ** -65], np.float32),
    ]
    for a in cases:
        exact = sum(fix_int(x) for x in a)
        assert O.fixsum_parts(a) == exact
        assert O.fixsum(a) == fix_to_double(exact)


def test_fixsum_order_independent():
    a = (np.random.default_rng(2).random(4096, dtype=np.float32) * 3e-3).astype(np.float32)
    ref = O.fixsum_parts(a)
    for seed in range(3):
        assert O.fixsum_parts(np.random.default_rng(seed).permutation(a)) == ref


@pytest.mark.parametrize("shape,iters,kw", [((7, 9), 4, VP_WORKLOAD), ((16, 16), 3, {}),
                                            ((6, 1), 5, VP_WORKLOAD),
                                            ((8, 11), 3, dict(flux_diffusion=0,
                                                              maintain_water_volume=0,
                                                              gravity=9.81))])
def test_numpy_restatement_fp32_bitexact(shape, iters, kw):
    ny, nx = shape
    z = O.fbm_terrain(nx, ny, 6) if nx > 1 else np.linspace(0, 1, ny, dtype=np.float32
                                                            ).reshape(ny, 1)
    p = _params(**kw)
    a = np_vpipes(z, p, iters, np.float32)
    st = O.VpState(z, O.vp_params(**kw)).run(iters)
    for name, x, y in zip("zdsuv", a, (st.z, st.d, st.s, st.u, st.v)):
        assert np.array_equal(x, y), name


def test_numpy_restatement_fp64_and_rain_map():
    z = O.fbm_terrain(12, 10, 8).astype(np.float64)
    rain = np.random.default_rng(3).random((10, 12)) + 0.5
    p = _params(**VP_WORKLOAD)
    a = np_vpipes(z, p, 3, np.float64, rain)
    st = O.VpState(z, O.vp_params(**VP_WORKLOAD), rain, np.float64).run(3)
    for x, y in zip(a, (st.z, st.d, st.s, st.u, st.v)):
        assert np.abs(x - y).max() <= 1e-12 * max(1.0, np.abs(y).max())


def test_flat_terrain_uniform_water_closed_form():
    z = np.full((20, 24), 0.5, np.float32)
    st = O.VpState(z, O.vp_params()).run(10)
    assert np.array_equal(st.z, z)
    for pl in (st.s, st.fl, st.fr, st.ft, st.fb, st.u, st.v):
        assert np.all(pl == 0)
    wh = np.float32(1e-2)
    assert np.abs(st.d - wh).max() <= 4 * np.spacing(wh)


def test_tilted_plane_first_step_closed_form():
    """z = a*i: interior cells send dt*a to the left, flux diffusion spreads
    c0*dt*a to every pipe, K = d1/(dt*a*dt) clamped to 1, no net inflow."""
    nx, ny, a = 32, 20, 1e-3
    z = np.tile(np.arange(nx) * a, (ny, 1))
    p = dict(water_height=1e-2, flux_diffusion_strength=0.01, dt=0.5)
    st = O.VpState(z, O.vp_params(**p), dtype=np.float64).run(1)
    dt, c0, wh = 0.5, np.float64(np.float32(0.01)), np.float64(np.float32(1e-2))
    K = min(wh / (dt * a * dt), 1.0)
    sl = (slice(3, ny - 3), slice(3, nx - 3))
    assert np.allclose(st.fl[sl], (1 - 3 * c0) * dt * a * K, rtol=1e-12, atol=0)
    for pl in (st.fr, st.ft, st.fb):
        assert np.allclose(pl[sl], c0 * dt * a * K, rtol=1e-12, atol=0)
    assert np.allclose(st.v[sl], 0.0, atol=1e-15)
    # u = 0.5*(fr(i-1) - fl + fr - fl(i+1)) / max(1e-3*wh/dt, d2), d2 = wh
    u = 0.5 * (2 * c0 - 2 * (1 - 3 * c0)) * dt * a * K / wh
    assert np.allclose(st.u[sl], u, rtol=1e-10)


def test_erosion_pass_conserves_z_plus_s_per_cell():
    rng = np.random.default_rng(5)
    ny, nx = 30, 40
    z = rng.random((ny, nx), dtype=np.float32)
    s = (rng.random((ny, nx), dtype=np.float32) * 0.01).astype(np.float32)
    d2 = (rng.random((ny, nx), dtype=np.float32) * 0.02).astype(np.float32)
    u = rng.standard_normal((ny, nx)).astype(np.float32)
    v = rng.standard_normal((ny, nx)).astype(np.float32)
    zo, so = np.zeros_like(z), np.zeros_like(z)
    p = O.vp_params(**VP_WORKLOAD)
    O.lib().orc_vp_erosion_f32(z, s, d2, u, v, zo, so, nx, ny, p)
    tot0 = z.astype(np.float64) + s
    tot1 = zo.astype(np.float64) + so
    assert np.abs(tot1 - tot0).max() <= 2 * np.spacing(np.float32(1.0))
    # erosion where capacity exceeds the sediment, deposition elsewhere
    assert np.any(zo < z) and np.any(zo > z)


def test_transport_zero_velocity_is_box_average():
    rng = np.random.default_rng(6)
    ny, nx = 16, 32
    s = rng.random((ny, nx), dtype=np.float32)
    zero = np.zeros_like(s)
    out = np.zeros_like(s)
    O.lib().orc_vp_transport_f32(zero, zero, s, out, nx, ny, 0.5)
    box = 0.25 * (s + nb_l(s) + nb_b(s) + nb_b(nb_l(s)))
    assert np.abs(out - box).max() <= 2e-7


def test_transport_mirrored_repeat_against_reflection():
    """Mirroring by reflection of the coordinate into [0, w] with period 2w
    (an independent formulation of CLK_ADDRESS_MIRRORED_REPEAT)."""
    rng = np.random.default_rng(8)
    ny, nx, dt = 12, 20, 0.5
    s = rng.random((ny, nx), dtype=np.float32)
    u = (rng.standard_normal((ny, nx)) * 30).astype(np.float32)
    v = (rng.standard_normal((ny, nx)) * 20).astype(np.float32)
    out = np.zeros_like(s)
    O.lib().orc_vp_transport_f32(u, v, s, out, nx, ny, dt)

    def refl(p, w):
        m = np.mod(p, 2 * w)
        m = np.where(m > w, 2 * w - m, m)
        q = m - 0.5
        q0 = np.floor(q)
        return (np.clip(q0, 0, w - 1).astype(int), np.clip(q0 + 1, 0, w - 1).astype(int),
                q - q0)

    jj, ii = np.mgrid[0:ny, 0:nx]
    i0, i1, a = refl(ii - dt * u.astype(np.float64), nx)
    j0, j1, b = refl(jj - dt * v.astype(np.float64), ny)
    ref = ((1 - a) * (1 - b) * s[j0, i0] + a * (1 - b) * s[j0, i1] + (1 - a) * b * s[j1, i0]
           + a * b * s[j1, i1])
    assert np.abs(out - ref).max() < 2e-5


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_water_volume_maintained(dtype):
    z = O.fbm_terrain(64, 48, 2).astype(dtype)
    st = O.VpState(z, O.vp_params(**VP_WORKLOAD), dtype=dtype)
    for _ in range(5):
        st.run(4)
        vol = O.fixsum(st.d) if dtype == np.float32 else st.d.sum()
        tol = 1e-5 if dtype == np.float32 else 1e-12
        assert abs(vol - st.W0) <= tol * st.W0


def _f32_patterns(step=4099):
    bits = np.arange(0, 1 << 32, step, dtype=np.uint64).astype(np.uint32)
    edge = np.array([0x477fe000, 0x477fefff, 0x477ff000, 0x38800000, 0x387fffff, 0x33800000,
                     0x33000000, 0x33000001, 0x7f7fffff, 0x00000001, 0x3f800000, 0x3f801000,
                     0x3f803000, 0x3f808000, 0x3f818000], np.uint32)
    bits = np.concatenate([bits, edge, edge | 0x80000000])
    x = bits.view(np.float32)
    return x[np.isfinite(x)]


def test_round_f16_matches_ieee_binary16():
    L = O.lib()
    x = _f32_patterns()
    got = np.array([L.orc_round_f16(float(v)) for v in x], np.float32)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(got, ref)


def test_round_bf16_matches_bit_rounding():
    L = O.lib()
    x = _f32_patterns()
    got = np.array([L.orc_round_bf16(float(v)) for v in x], np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)
    # values that round past the largest finite bfloat16 become infinity
    assert np.array_equal(got, r)


@pytest.mark.parametrize("storage", [1, 2])
def test_storage_rounding_keeps_planes_representable(storage):
    z = O.fbm_terrain(40, 32, 2)
    st = O.VpState(z, O.vp_params(storage=storage, **VP_WORKLOAD)).run(6)
    dt = np.float16 if storage == 2 else None
    for plane in (st.fl, st.fr, st.ft, st.fb):
        if dt is not None:
            assert np.array_equal(plane.astype(dt).astype(np.float32), plane)
        else:
            assert np.all((plane.view(np.uint32) & 0xFFFF) == 0)
    # FP32 storage (storage 0) is the FP32 contract itself
    a = O.VpState(z, O.vp_params(storage=0, **VP_WORKLOAD)).run(6)
    b = O.VpState(z, O.vp_params(**VP_WORKLOAD)).run(6)
    assert np.array_equal(a.z, b.z)


def test_fp32_contract_tracks_fp64_at_reference_defaults():
    """At hmap::gpu::hydraulic_vpipes' default parameters the FP32 contract
    stays within 1e-4 of the height range of the FP64 oracle after 50
    iterations (256^2, seed 2)."""
    z = O.fbm_terrain(256, 256, 2)
    a = O.VpState(z, O.vp_params()).run(50)
    b = O.VpState(z.astype(np.float64), O.vp_params(), dtype=np.float64).run(50)
    assert np.abs(a.z - b.z).max() <= 1e-4 * (b.z.max() - b.z.min())
