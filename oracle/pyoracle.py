"""ctypes binding of liboracle.so (hm_oracle.c).  TEST INFRASTRUCTURE ONLY.

Arrays follow hmap::Array: shape (ny, nx) in numpy terms, C order, so the
linear index is j*nx + i with i the fastest axis
(HighMap/include/highmap/array.hpp:236-239).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class VpParams(C.Structure):
    """Mirror of orc_vp_params (and of hmg_vpipes_params' leading fields)."""

    _fields_ = [
        ("water_height", C.c_float),
        ("evap_rate", C.c_float),
        ("dt", C.c_float),
        ("k_capacity", C.c_float),
        ("k_erode", C.c_float),
        ("k_depose", C.c_float),
        ("k_discharge_exp", C.c_float),
        ("downcutting_max_depth_ratio", C.c_float),
        ("flux_diffusion_strength", C.c_float),
        ("gravity", C.c_float),
        ("rain_per_iter", C.c_float),
        ("maintain_water_volume", C.c_int),
        ("flux_diffusion", C.c_int),
        ("storage", C.c_int),
    ]


# Defaults of hmap::gpu::hydraulic_vpipes (erosion.hpp:1206-1223).
VP_DEFAULTS = dict(
    water_height=1e-2,
    evap_rate=0.1,
    dt=0.5,
    k_capacity=0.5,
    k_erode=1e-3,
    k_depose=1e-2,
    k_discharge_exp=1.0,
    downcutting_max_depth_ratio=10.0,
    flux_diffusion_strength=0.01,
    gravity=1.0,
    rain_per_iter=0.0,
    maintain_water_volume=1,
    flux_diffusion=1,
    storage=0,
)

_INT_FIELDS = ("maintain_water_volume", "flux_diffusion", "storage")


def vp_params(**kw):
    d = dict(VP_DEFAULTS)
    for k, v in kw.items():
        if k not in d:
            raise KeyError(k)
        d[k] = v
    return VpParams(**{k: (int(v) if k in _INT_FIELDS else float(v)) for k, v in d.items()})


def build():
    """Compile liboracle.so and cpu_bench with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    u32 = C.c_uint32
    L.orc_wang_hash.argtypes = [u32]
    L.orc_wang_hash.restype = u32
    L.orc_xorshift32.argtypes = [u32]
    L.orc_xorshift32.restype = u32
    L.orc_seed_to_fseed.argtypes = [u32]
    L.orc_seed_to_fseed.restype = C.c_float
    for name in ("orc_sinf", "orc_cosf", "orc_fractf"):
        getattr(L, name).argtypes = [C.c_float]
        getattr(L, name).restype = C.c_float
    L.orc_hash12f.argtypes = [C.c_float, C.c_float, C.c_float]
    L.orc_hash12f.restype = C.c_float
    L.orc_base_perlin.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]
    L.orc_base_perlin.restype = C.c_float
    L.orc_base_noise.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int]
    L.orc_base_noise.restype = C.c_float
    L.orc_noise_type_supported.argtypes = [C.c_int]
    L.orc_noise_type_supported.restype = C.c_int
    L.orc_noise_f32.argtypes = [F32P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, u32,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_noise_f32.restype = C.c_int
    L.orc_noise_fbm_f32.argtypes = [F32P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                    u32, C.c_int, C.c_float, C.c_float, C.c_float,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_noise_fbm_f32.restype = C.c_int
    L.orc_remap_f32.argtypes = [F32P, C.c_size_t, C.c_float, C.c_float]
    L.orc_thermal_steps_f32.argtypes = [F32P, C.c_int, C.c_int, C.c_void_p, C.c_float, C.c_int]
    L.orc_thermal_steps_f64.argtypes = [F64P, C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int]
    L.orc_extrapolate_borders_f32.argtypes = [F32P, C.c_int, C.c_int]
    L.orc_extrapolate_borders_f64.argtypes = [F64P, C.c_int, C.c_int]
    L.orc_fixsum_f32.argtypes = [F32P, C.c_size_t]
    L.orc_fixsum_f32.restype = C.c_double
    L.orc_fixsum_parts.argtypes = [F32P, C.c_size_t, C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_int64)]
    L.orc_xorshift_period.argtypes = [u32]
    L.orc_xorshift_period.restype = C.c_uint64
    for name in ("orc_round_bf16", "orc_round_f16"):
        getattr(L, name).argtypes = [C.c_float]
        getattr(L, name).restype = C.c_float
    pp = C.POINTER(VpParams)
    dp = C.POINTER(C.c_double)
    L.orc_vpipes_init_f32.argtypes = [F32P] * 8 + [C.c_int, C.c_int, pp, C.c_void_p, dp, dp]
    L.orc_vpipes_init_f64.argtypes = [F64P] * 8 + [C.c_int, C.c_int, pp, C.c_void_p, dp, dp]
    L.orc_vpipes_run_f32.argtypes = [F32P] * 9 + [C.c_int, C.c_int, pp, C.c_void_p,
                                                  C.c_double, C.c_double, C.c_int]
    L.orc_vpipes_run_f64.argtypes = [F64P] * 9 + [C.c_int, C.c_int, pp, C.c_void_p,
                                                  C.c_double, C.c_double, C.c_int]
    L.orc_vp_flow_f32.argtypes = [F32P] * 10 + [C.c_int, C.c_int, pp]
    L.orc_vp_water_f32.argtypes = [F32P] * 8 + [C.c_int, C.c_int, pp]
    L.orc_vp_erosion_f32.argtypes = [F32P] * 7 + [C.c_int, C.c_int, pp]
    L.orc_vp_transport_f32.argtypes = [F32P] * 4 + [C.c_int, C.c_int, C.c_float]
    L.orc_vp_erosion_f64.argtypes = [F64P] * 7 + [C.c_int, C.c_int, pp]
    L.orc_grad22f.argtypes = [C.c_float] * 3 + [C.POINTER(C.c_float)] * 2
    L.orc_thermal_auto_bedrock_steps_f32.argtypes = [F32P, C.c_int, C.c_int, C.c_void_p,
                                                     C.c_float, C.c_int]
    L.orc_thermal_auto_bedrock_steps_f64.argtypes = [F64P, C.c_int, C.c_int, C.c_void_p,
                                                     C.c_double, C.c_int]
    L.orc_sediment_deposition_f32.argtypes = [F32P, C.c_int, C.c_int, C.c_void_p, C.c_float,
                                              C.c_void_p, C.c_float, C.c_int, C.c_int]
    L.orc_sediment_deposition_f64.argtypes = [F64P, C.c_int, C.c_int, C.c_void_p, C.c_double,
                                              C.c_void_p, C.c_float, C.c_int, C.c_int]
    fl = C.c_float
    L.orc_flow_simulation_f32.argtypes = [F32P, C.c_int, C.c_int, fl, F32P, C.c_int, fl, C.c_int,
                                          fl, fl, F32P]
    L.orc_flow_simulation_f64.argtypes = [F64P, C.c_int, C.c_int, fl, F64P, C.c_int, fl, C.c_int,
                                          fl, fl, F64P]
    L.orc_expf.argtypes = [fl]
    L.orc_expf.restype = fl
    L.orc_powf.argtypes = [fl, fl]
    L.orc_powf.restype = fl
    L.orc_powf_many.argtypes = [F32P, F32P, F32P, C.c_long]
    L.orc_sincos_many.argtypes = [F32P, F32P, F32P, C.c_long]
    snow_tail = [C.c_int] + [fl] * 7 + [C.c_int, fl]
    L.orc_snow_simulation_f32.argtypes = ([F32P, C.c_int, C.c_int, fl, F32P, F32P, F32P]
                                          + snow_tail + [F32P])
    L.orc_snow_simulation_f64.argtypes = ([F64P, C.c_int, C.c_int, fl, F64P, F64P, F64P]
                                          + snow_tail + [F64P])
    L.orc_snow_step_f32.argtypes = [F32P] * 5 + [C.c_int, C.c_int] + [fl] * 7
    L.orc_snow_step_f64.argtypes = [F64P] * 5 + [C.c_int, C.c_int] + [C.c_double] * 7
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ----------------------------------------------------------------- noise


def noise_fbm(nx, ny, kw=(4.0, 4.0), seed=0, octaves=8, weight=0.7, persistence=0.5,
              lacunarity=2.0, ctrl=None, noise_x=None, noise_y=None,
              bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0), noise_type=1):
    """FP32 restatement of hmap::gpu::noise_fbm (every hmap::NoiseType the
    kernels implement; others are all zeros)."""
    out = np.zeros((ny, nx), np.float32)
    keep = []

    def opt(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == (ny, nx)
        keep.append(a)
        return _ptr(a)

    bb = np.asarray(bbox, np.float32)
    pe = np.asarray(period, np.int32)
    rc = lib().orc_noise_fbm_f32(out, nx, ny, noise_type, kw[0], kw[1], seed & 0xFFFFFFFF,
                                 octaves, weight, persistence, lacunarity,
                                 opt(ctrl), opt(noise_x), opt(noise_y), _ptr(bb), _ptr(pe))
    if rc not in (0, 1):
        raise ValueError("noise_fbm failed (%d)" % rc)
    return out


def noise(nx, ny, kw=(4.0, 4.0), seed=0, noise_type=1, noise_x=None, noise_y=None,
          bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0)):
    """FP32 restatement of hmap::gpu::noise (single octave)."""
    out = np.zeros((ny, nx), np.float32)
    keep = [np.ascontiguousarray(a, np.float32) if a is not None else None
            for a in (noise_x, noise_y)]
    bb = np.asarray(bbox, np.float32)
    pe = np.asarray(period, np.int32)
    lib().orc_noise_f32(out, nx, ny, noise_type, kw[0], kw[1], seed & 0xFFFFFFFF,
                        _ptr(keep[0]), _ptr(keep[1]), _ptr(bb), _ptr(pe))
    return out


def hash12f(px, py, fseed):
    """hash12f at one point (float32 inputs)."""
    return float(lib().orc_hash12f(px, py, fseed))


def grad22f(px, py, fseed):
    """grad22f at one point: (cos, sin) of 6.2831853 * hash12f."""
    gx, gy = C.c_float(), C.c_float()
    lib().orc_grad22f(px, py, fseed, C.byref(gx), C.byref(gy))
    return gx.value, gy.value


def vp_erosion(z, s, d2, u, v, params, dtype=np.float32):
    """One erosion pass of hydraulic_vpipes on given planes; returns (z, s)."""
    planes = [np.ascontiguousarray(a, dtype) for a in (z, s, d2, u, v)]
    ny, nx = planes[0].shape
    zo, so = np.empty_like(planes[0]), np.empty_like(planes[0])
    fn = lib().orc_vp_erosion_f32 if dtype == np.float32 else lib().orc_vp_erosion_f64
    fn(*planes, zo, so, nx, ny, C.byref(params))
    return zo, so


def base_noise(noise_type, px, py, fseed, period=(0, 0)):
    """Single-octave value at one point (float32 inputs)."""
    return lib().orc_base_noise(noise_type, px, py, fseed, period[0], period[1])


def remap(a, vmin=0.0, vmax=1.0):
    a = np.ascontiguousarray(a, np.float32).copy()
    lib().orc_remap_f32(a.reshape(-1), a.size, vmin, vmax)
    return a


def fbm_terrain(n_x, n_y, seed):
    """remap(noise_fbm(PERLIN, kw 4, seed, 8 octaves, w 0.7, pers 0.5, lac 2)) in [0,1]."""
    return remap(noise_fbm(n_x, n_y, (4.0, 4.0), seed))


# --------------------------------------------------------------- thermal


def thermal_steps(z, talus, iterations, dtype=np.float32):
    z = np.ascontiguousarray(z, dtype).copy()
    ny, nx = z.shape
    L = lib()
    fn = L.orc_thermal_steps_f32 if dtype == np.float32 else L.orc_thermal_steps_f64
    if np.isscalar(talus):
        fn(z, nx, ny, None, float(talus), iterations)
    else:
        t = np.ascontiguousarray(talus, dtype)
        fn(z, nx, ny, _ptr(t), 0.0, iterations)
    return z


def extrapolate_borders(z, dtype=np.float32):
    z = np.ascontiguousarray(z, dtype).copy()
    ny, nx = z.shape
    L = lib()
    (L.orc_extrapolate_borders_f32 if dtype == np.float32 else L.orc_extrapolate_borders_f64)(z, nx, ny)
    return z


def thermal(z, talus, iterations=10, dtype=np.float32):
    """hmap::gpu::thermal restated: Jacobi steps then extrapolate_borders."""
    return extrapolate_borders(thermal_steps(z, talus, iterations, dtype), dtype)


def _talus_args(talus, shape, dtype):
    if np.isscalar(talus):
        return None, float(talus)
    t = np.ascontiguousarray(talus, dtype)
    if t.shape != shape:
        raise ValueError("talus shape mismatch")
    return t, 0.0


def thermal_auto_bedrock(z, talus, iterations=10, dtype=np.float32):
    """hmap::gpu::thermal_auto_bedrock restated: Jacobi steps then
    extrapolate_borders; returns the new z."""
    z = np.ascontiguousarray(z, dtype).copy()
    ny, nx = z.shape
    t, ts = _talus_args(talus, z.shape, dtype)
    fn = (lib().orc_thermal_auto_bedrock_steps_f32 if dtype == np.float32
          else lib().orc_thermal_auto_bedrock_steps_f64)
    fn(z, nx, ny, _ptr(t), ts, int(iterations))
    return extrapolate_borders(z, dtype)


def sediment_deposition(z, talus, max_deposition=0.01, iterations=5, thermal_subiterations=10,
                        dtype=np.float32):
    """hmap::gpu::sediment_deposition restated; returns (z, deposition)."""
    z = np.ascontiguousarray(z, dtype).copy()
    ny, nx = z.shape
    t, ts = _talus_args(talus, z.shape, dtype)
    dep = np.empty_like(z)
    fn = (lib().orc_sediment_deposition_f32 if dtype == np.float32
          else lib().orc_sediment_deposition_f64)
    fn(z, nx, ny, _ptr(t), ts, _ptr(dep), float(max_deposition), int(iterations),
       int(thermal_subiterations))
    return z, dep


def fixsum(a):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    return lib().orc_fixsum_f32(a, a.size)


# ---------------------------------------------------------------- vpipes

PLANES = ("z", "d", "s", "fl", "fr", "ft", "fb", "u", "v")


class VpState:
    """Caller-owned vpipes state planes for the oracle."""

    def __init__(self, z, params, rain=None, dtype=np.float32):
        self.dtype = dtype
        self.z = np.ascontiguousarray(z, dtype).copy()
        ny, nx = self.z.shape
        self.nx, self.ny = nx, ny
        for name in PLANES[1:]:
            setattr(self, name, np.zeros((ny, nx), dtype))
        self.p = params
        self.rain = None if rain is None else np.ascontiguousarray(rain, dtype)
        W0 = C.c_double()
        R = C.c_double()
        L = lib()
        fn = L.orc_vpipes_init_f32 if dtype == np.float32 else L.orc_vpipes_init_f64
        fn(self.d, self.s, self.fl, self.fr, self.ft, self.fb, self.u, self.v, nx, ny,
           C.byref(self.p), _ptr(self.rain), C.byref(W0), C.byref(R))
        self.W0, self.R = W0.value, R.value

    def run(self, iterations):
        L = lib()
        fn = L.orc_vpipes_run_f32 if self.dtype == np.float32 else L.orc_vpipes_run_f64
        fn(self.z, self.d, self.s, self.fl, self.fr, self.ft, self.fb, self.u, self.v,
           self.nx, self.ny, C.byref(self.p), _ptr(self.rain), self.W0, self.R, iterations)
        return self

    def thermal_step(self, talus, iterations=1):
        """One thermal() call on z of the running state (steps + borders)."""
        self.z = thermal(self.z, talus, iterations, self.dtype)
        return self


def hydraulic_vpipes(z, params=None, rain=None, iterations=50, dtype=np.float32):
    """hmap::gpu::hydraulic_vpipes restated; returns (z, d, s, u, v)."""
    st = VpState(z, params or vp_params(), rain, dtype).run(iterations)
    return st.z, st.d, st.s, st.u, st.v


def fixsum_parts(a):
    """Exact 128-bit fixed-point sum of round(a * 2^64) as a Python int."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    lo = C.c_uint64()
    hi = C.c_int64()
    lib().orc_fixsum_parts(a, a.size, C.byref(lo), C.byref(hi))
    return (hi.value << 64) + lo.value


# ------------------------------------------------------- flow simulation


def flow_simulation(z, water_height, depth_map, iterations, dt=0.5, flux_diffusion=True,
                    flux_diffusion_strength=0.01, dry_out_ratio=0.0, dtype=np.float32):
    """hmap::gpu::flow_simulation restated; returns the water depth."""
    z = np.ascontiguousarray(z, dtype)
    ny, nx = z.shape
    dm = np.ascontiguousarray(depth_map, dtype)
    if dm.shape != z.shape:
        raise ValueError("depth_map shape mismatch")
    d = np.empty_like(z)
    fn = lib().orc_flow_simulation_f32 if dtype == np.float32 else lib().orc_flow_simulation_f64
    fn(z, nx, ny, float(water_height), dm, int(iterations), float(dt), int(bool(flux_diffusion)),
       float(flux_diffusion_strength), float(dry_out_ratio), d)
    return d


# ------------------------------------------------------------------ snow

SNOW_DEFAULTS = dict(dt=0.5, fall_iterations_ratio=1.0, k_snow=0.5, k_visc=0.1, k_melt_factor=0.8,
                     k_depth_ratio=1.0, k_depth_slope_ratio=1.0, post_filter=True,
                     thermal_talus_ratio=0.2)


def expf(x):
    """The FP32 contract's exp (shared with the kernels)."""
    return float(lib().orc_expf(float(x)))


def powf(x, y):
    """The FP32 contract's pow (shared with the kernels)."""
    return float(lib().orc_powf(float(x), float(y)))


def sincos_many(x):
    """(orc_sinf, orc_cosf) elementwise over a float32 array."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().orc_sincos_many(x, s, c, x.size)
    return s, c


def powf_many(x, y):
    """powf elementwise over two float32 arrays of one shape."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    assert x.shape == y.shape
    out = np.empty_like(x)
    lib().orc_powf_many(x, y, out, x.size)
    return out


def snow_simulation(z, snow_depth, fall_map, melting_map, talus, iterations, dtype=np.float32,
                    **kw):
    """hmap::gpu::snow_simulation restated; returns the snow depth.  talus,
    fall_map and melting_map are planes of z's shape (scalars broadcast)."""
    p = dict(SNOW_DEFAULTS)
    unknown = set(kw) - set(p)
    if unknown:
        raise TypeError("unknown parameters %s" % sorted(unknown))
    p.update(kw)
    z = np.ascontiguousarray(z, dtype)
    ny, nx = z.shape

    def plane(a):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype), z.shape), dtype)

    f, m, t = plane(fall_map), plane(melting_map), plane(talus)
    s = np.empty_like(z)
    fn = lib().orc_snow_simulation_f32 if dtype == np.float32 else lib().orc_snow_simulation_f64
    fn(z, nx, ny, float(snow_depth), f, m, t, int(iterations), p["dt"],
       p["fall_iterations_ratio"], p["k_snow"], p["k_visc"], p["k_melt_factor"],
       p["k_depth_ratio"], p["k_depth_slope_ratio"], int(bool(p["post_filter"])),
       p["thermal_talus_ratio"], s)
    return s


def snow_step(z, s, talus, melt, dt=0.5, snow_depth=0.1, k_snow=0.5, k_melt=0.0, k_visc=0.1,
              k_depth_ratio=1.0, k_depth_slope_ratio=1.0, dtype=np.float32):
    """One snow_simulation kernel launch (no fall, no border fill)."""
    z = np.ascontiguousarray(z, dtype)
    ny, nx = z.shape

    def plane(a):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype), z.shape), dtype)

    so = np.empty_like(z)
    fn = lib().orc_snow_step_f32 if dtype == np.float32 else lib().orc_snow_step_f64
    fn(z, plane(s), plane(talus), plane(melt), so, nx, ny, dt, snow_depth, k_snow, k_melt, k_visc,
       k_depth_ratio, k_depth_slope_ratio)
    return so
