"""GPU parity: libhmg (HIP, gfx950) against the FP32 oracle, bit for bit.

The FP32 oracle restates the reference kernels with the same operation order
as the HIP code, so every comparison below is exact equality
(np.array_equal), except where a tolerance is written next to it.  Inputs
are generated once by the oracle's noise and fed to both sides.
"""
import numpy as np
import pytest

from conftest import VP_WORKLOAD

pytestmark = pytest.mark.gpu


def _terrain(orc, nx, ny, seed):
    return orc.fbm_terrain(nx, ny, seed)


# ------------------------------------------------------------------ noise


@pytest.mark.parametrize("shape,seed,kw", [((256, 256), 1, (4.0, 4.0)),
                                           ((97, 61), 7, (3.0, 5.5)),
                                           ((1024, 64), 123456789, (16.0, 2.0))])
def test_noise_fbm_bitexact(hm, orc, shape, seed, kw):
    nx, ny = shape
    g = hm.noise_fbm(hm.NoiseType.PERLIN, shape, kw, seed)
    o = orc.noise_fbm(nx, ny, kw, seed)
    assert g.shape == (ny, nx)
    assert np.array_equal(g, o), np.abs(g - o).max()


def test_noise_fbm_options_bitexact(hm, orc):
    nx, ny = 130, 77
    rng = np.random.default_rng(3)
    ctrl = rng.random((ny, nx), dtype=np.float32)
    wx = (rng.random((ny, nx), dtype=np.float32) - 0.5) * 0.1
    wy = (rng.random((ny, nx), dtype=np.float32) - 0.5) * 0.1
    kw = dict(octaves=5, weight=0.3, persistence=0.6, lacunarity=2.0,
              bbox=(-0.5, 1.5, 0.25, 2.0), period=(4, 6))
    g = hm.noise_fbm(hm.NoiseType.PERLIN, (nx, ny), (4.0, 6.0), 42, ctrl_param=ctrl,
                     noise_x=wx, noise_y=wy, **kw)
    o = orc.noise_fbm(nx, ny, (4.0, 6.0), 42, ctrl=ctrl, noise_x=wx, noise_y=wy, **kw)
    assert np.array_equal(g, o)


def test_noise_type_outside_the_enum_raises(hm):
    with pytest.raises(hm.HmgError):
        hm.noise_fbm(13, (16, 16), (4, 4), 0)
    with pytest.raises(hm.HmgError):
        hm.noise(-1, (16, 16), (4, 4), 0)


# ---------------------------------------------------------------- thermal


def test_thermal_256_talus_2_over_256_50_iterations(hm, orc):
    """hmap::gpu::thermal on 256^2 fBm (seed 1), talus 2/256, 50 iterations."""
    z0 = _terrain(orc, 256, 256, 1)
    z = z0.copy()
    hm.thermal(z, 2.0 / 256, 50)
    o32 = orc.thermal(z0, 2.0 / 256, 50)
    assert np.array_equal(z, o32)
    o64 = orc.thermal(z0.astype(np.float64), 2.0 / 256, 50, np.float64)
    rng = o64.max() - o64.min()
    assert np.abs(z - o64).max() <= 1e-4 * rng


@pytest.mark.parametrize("nx,ny,iters", [(3, 3, 4), (5, 7, 3), (130, 67, 0), (130, 67, 1),
                                         (130, 67, 8), (131, 70, 9), (64, 16, 17),
                                         (300, 200, 25), (520, 512, 10)])
def test_thermal_shapes_and_step_counts(hm, orc, nx, ny, iters):
    z0 = _terrain(orc, nx, ny, 5) if nx > 4 else np.random.default_rng(1).random(
        (ny, nx), dtype=np.float32)
    z = z0.copy()
    hm.thermal(z, 1.0 / nx, iters)
    assert np.array_equal(z, orc.thermal(z0, 1.0 / nx, iters))


def test_thermal_talus_plane_bedrock_deposition_mask(hm, orc):
    nx, ny = 150, 90
    rng = np.random.default_rng(11)
    z0 = _terrain(orc, nx, ny, 9)
    talus = (rng.random((ny, nx), dtype=np.float32) * 3.0 / nx).astype(np.float32)
    z = z0.copy()
    dep = np.zeros_like(z)
    hm.thermal(z, talus, 12, deposition_map=dep)
    o = orc.thermal(z0, talus, 12)
    assert np.array_equal(z, o)
    assert np.array_equal(dep, np.where((o - z0) < 0, np.float32(0), o - z0))

    bed = (z0 - 0.01).astype(np.float32)
    bed[::3] = z0[::3] + 0.05
    z = z0.copy()
    hm.thermal(z, 1.0 / nx, 5, bedrock=bed)
    ob = z0.copy()
    for _ in range(5):  # interior: max(old, bedrock); borders: inward copy
        nxt = ob.copy()
        nxt[1:-1, 1:-1] = np.maximum(ob[1:-1, 1:-1], bed[1:-1, 1:-1])
        nxt[:, 0] = ob[:, 1]
        nxt[:, -1] = ob[:, -2]
        nxt[0, 1:-1] = ob[1, 1:-1]
        nxt[-1, 1:-1] = ob[-2, 1:-1]
        ob = nxt
    assert np.array_equal(z, orc.extrapolate_borders(ob))

    mask = rng.random((ny, nx), dtype=np.float32)
    z = z0.copy()
    hm.thermal(z, 1.0 / nx, 7, mask=mask)
    r = orc.thermal(z0, 1.0 / nx, 7)
    assert np.array_equal(z, z0 * (np.float32(1) - mask) + r * mask)


# ----------------------------------------------------------------- vpipes


def _vp_oracle(orc, z0, iters, rain=None, **kw):
    st = orc.VpState(z0, orc.vp_params(**kw), rain).run(iters)
    return st.z, st.d, st.s, st.u, st.v


def _vp_gpu(hm, z0, iters, mode, rain=None, **kw):
    z = z0.copy()
    outs = [np.zeros_like(z0) for _ in range(4)]
    hm.hydraulic_vpipes(z, iterations=iters, rain_map=rain, water_depth=outs[0],
                        sediment=outs[1], vel_u=outs[2], vel_v=outs[3], mode=mode, **kw)
    return (z, *outs)


NAMES = ("z", "d", "s", "u", "v")


def _assert_same(g, o):
    for name, a, b in zip(NAMES, g, o):
        assert np.array_equal(a, b), "%s differs: max |diff| %g" % (name, np.abs(a - b).max())


@pytest.mark.parametrize("mode", ["plain", "fused"])
@pytest.mark.parametrize("nx,ny,iters", [(64, 48, 1), (64, 48, 2), (97, 130, 7),
                                         (256, 256, 30)])
def test_vpipes_bitexact_workload_params(hm, orc, mode, nx, ny, iters):
    z0 = _terrain(orc, nx, ny, 2)
    _assert_same(_vp_gpu(hm, z0, iters, mode, **VP_WORKLOAD),
                 _vp_oracle(orc, z0, iters, **VP_WORKLOAD))


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_vpipes_bitexact_reference_defaults(hm, orc, mode):
    z0 = _terrain(orc, 200, 150, 4)
    _assert_same(_vp_gpu(hm, z0, 50, mode), _vp_oracle(orc, z0, 50))


@pytest.mark.parametrize("mode", ["plain", "fused"])
@pytest.mark.parametrize("variant", ["rain_map", "no_volume", "no_diffusion", "gravity_rain",
                                     "tiny"])
def test_vpipes_bitexact_variants(hm, orc, mode, variant):
    nx, ny, iters = 120, 90, 12
    kw = dict(VP_WORKLOAD)
    rain = None
    if variant == "rain_map":
        rain = (np.random.default_rng(5).random((ny, nx), dtype=np.float32) + 0.5)
    elif variant == "no_volume":
        kw["maintain_water_volume"] = 0
    elif variant == "no_diffusion":
        kw["flux_diffusion"] = 0
    elif variant == "gravity_rain":
        kw.update(gravity=9.81, rain_per_iter=1e-3, maintain_water_volume=0)
    elif variant == "tiny":
        nx, ny = 1, 5
    z0 = _terrain(orc, nx, ny, 8) if nx > 1 else np.linspace(0, 1, ny, dtype=np.float32
                                                             ).reshape(ny, 1)
    o = _vp_oracle(orc, z0, iters, rain, **kw)
    g = _vp_gpu(hm, z0, iters, mode, rain, **kw)
    _assert_same(g, o)


@pytest.mark.parametrize("kexp", [1.3, 0.7, 2.5])
@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_vpipes_discharge_exponent(hm, orc, mode, kexp):
    """k_discharge_exp != 1 goes through the contract's pow, shared by the
    kernels and the oracle: bit for bit like every other path."""
    z0 = _terrain(orc, 96, 80, 3)
    g = _vp_gpu(hm, z0, 12, mode, k_discharge_exp=kexp, **VP_WORKLOAD)
    o = _vp_oracle(orc, z0, 12, k_discharge_exp=kexp, **VP_WORKLOAD)
    _assert_same(g, o)


# --------------------------------------------------------------- sessions


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_session_vpipes_thermal_interleaved(hm, orc, mode):
    """One continuous state: 10 hydraulic iterations then one thermal step
    (talus 1/n) with border extrapolation, 6 times."""
    n = 192
    z0 = _terrain(orc, n, n, 2)
    s = hm.Session(z0, mode=mode, **VP_WORKLOAD)
    o = orc.VpState(z0, orc.vp_params(**VP_WORKLOAD))
    for _ in range(6):
        s.vpipes(10).thermal(1.0 / n, 1)
        o.run(10).thermal_step(1.0 / n, 1)
    g = s.export()
    s.close()
    _assert_same(g, (o.z, o.d, o.s, o.u, o.v))


@pytest.mark.parametrize("mode", ["plain", "fused"])
@pytest.mark.parametrize("storage", [1, 2])
def test_session_16bit_storage_bitexact(hm, orc, mode, storage):
    """BF16 (1) / FP16 (2) storage of water, sediment and fluxes with FP32
    arithmetic: identical to the oracle with the same storage rounding."""
    nx, ny = 150, 110
    z0 = _terrain(orc, nx, ny, 12)
    s = hm.Session(z0, mode=mode, storage=storage, **VP_WORKLOAD)
    o = orc.VpState(z0, orc.vp_params(storage=storage, **VP_WORKLOAD))
    for _ in range(3):
        s.vpipes(7).thermal(1.0 / nx, 1)
        o.run(7).thermal_step(1.0 / nx, 1)
    g = s.export()
    s.close()
    _assert_same(g, (o.z, o.d, o.s, o.u, o.v))


def test_fused_launch_shapes_do_not_change_results(hm, orc):
    z0 = _terrain(orc, 300, 170, 13)
    outs = []
    for ch, pf in ((1, 0), (7, 1), (32, 0), (64, 1), (200, 1), (1, 2), (5, 2), (72, 2), (200, 2)):
        s = hm.Session(z0, **VP_WORKLOAD)
        s.set_tuning(ch, pf)
        s.vpipes(9)
        outs.append(s.export())
        s.close()
    for o in outs[1:]:
        _assert_same(o, outs[0])


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_gathers_ahead_variant_matches_oracle(hm, orc, storage):
    """The gathers-ahead form of the fused kernel (launch variant 2) with a
    rain plane, thermal steps between calls and 16-bit storage: the same
    state as the oracle."""
    nx, ny = 203, 141
    z0 = _terrain(orc, nx, ny, 21)
    rain = np.random.default_rng(4).random((ny, nx), dtype=np.float32)
    kw = dict(VP_WORKLOAD, rain_per_iter=1e-4)
    s = hm.Session(z0, rain_map=rain, storage=storage, **kw)
    s.set_tuning(17, 2)
    o = orc.VpState(z0, orc.vp_params(storage=storage, **kw), rain=rain)
    for _ in range(2):
        s.vpipes(6).thermal(1.0 / nx, 1)
        o.run(6).thermal_step(1.0 / nx, 1)
    g = s.export()
    s.close()
    _assert_same(g, (o.z, o.d, o.s, o.u, o.v))


@pytest.mark.parametrize("storage", [0, 1])
def test_online_shape_tuning_keeps_results(hm, orc, storage):
    """A session of 2^20 cells tunes its launch shape on probe iterations
    from iteration 20 on (across calls); the state equals a fixed-shape
    run."""
    z0 = _terrain(orc, 1024, 1024, 12)
    a = hm.Session(z0, storage=storage, **VP_WORKLOAD)
    for k in (5, 17, 4, 4):  # the probes (iterations 20-26) straddle calls
        a.vpipes(k)
    b = hm.Session(z0, storage=storage, **VP_WORKLOAD)
    b.set_tuning(64, 1)
    b.vpipes(30)
    _assert_same(a.export(), b.export())
    c = hm.Session(z0, storage=storage, **VP_WORKLOAD)  # shape from the cache
    c.vpipes(30)
    _assert_same(c.export(), b.export())
    for s in (a, b, c):
        s.close()


def test_tuning_kept_when_session_ends_mid_probe(hm, orc):
    """A session that ends after one full round of probes (as a drop-in call
    shorter than the probes does) caches the best shape probed so far, so
    the next session of that shape starts tuned.  At 1024 x 1040 on 256 CUs
    three shapes qualify (8, 12 and 16 rows); a warm-up probe and one round
    are iterations 20-23."""
    z0 = _terrain(orc, 1024, 1040, 13)
    a = hm.Session(z0, storage=0, **VP_WORKLOAD)
    a.vpipes(24)
    assert a.tuning()[2], "still probing after 24 iterations"
    a.close()
    b = hm.Session(z0, storage=0, **VP_WORKLOAD)
    ch, pf, probing = b.tuning()
    assert not probing
    assert (ch, pf) in {(8, 2), (12, 1), (16, 2)}
    b.close()


@pytest.mark.parametrize("storage", [0, 1])
def test_session_sums_and_max_backtrace(hm, orc, storage):
    """The device reductions over a session's state: the sums of z, s and d
    (exact fixed point, returned in FP64) equal the exactly rounded FP64
    sums of the exported planes to 1e-12 relative; the max backtrace is
    dt * max(|u|, |v|) of the exported velocities, exactly (an FP64 product
    of two floats); before any iteration it is 0."""
    import math

    z0 = _terrain(orc, 200, 136, 21)
    s = hm.Session(z0, storage=storage, **VP_WORKLOAD)
    assert s.max_backtrace() == 0.0
    s.vpipes(40)
    z, d, sd, u, v = s.export()
    gz, gs, gd = s.sums()
    for g, a in ((gz, z), (gs, sd), (gd, d)):
        ref = math.fsum(a.astype(np.float64).ravel())
        assert abs(g - ref) <= 1e-12 * max(1.0, abs(ref)), (g, ref)
    dt = np.float64(np.float32(VP_WORKLOAD["dt"]))
    want = dt * max(np.abs(u).max(), np.abs(v).max()).astype(np.float64)
    assert s.max_backtrace() == want
    assert want > 0.0
    s.close()


def test_plain_and_fused_agree_4096(hm, orc):
    """At full size: 6 iterations plain vs fused are identical, and the water
    volume stays at W0 (maintain_water_volume) to FP32 rounding."""
    n = 4096
    z0 = _terrain(orc, n, n, 2)
    outs = []
    for mode in ("plain", "fused"):
        s = hm.Session(z0, mode=mode, **VP_WORKLOAD)
        s.vpipes(6)
        outs.append(s.export())
        s.close()
    _assert_same(outs[0], outs[1])
    d = outs[1][1].astype(np.float64)
    W0 = n * n * np.float64(np.float32(1e-3))
    assert abs(d.sum() - W0) <= 1e-5 * W0
