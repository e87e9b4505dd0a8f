"""highmap_amd -- HighMap's grid-erosion hot path on MI355X (gfx950).

Python mirror of the reference's hmap::gpu operator interface for this path:

    init_gpu()          hmap::gpu::init_opencl   (gpu_opencl.cpp:32-134)
    noise_fbm(...)      hmap::gpu::noise_fbm     (primitives_gpu.cpp:543-591)
    thermal(...)        hmap::gpu::thermal       (thermal_gpu.cpp:17-96)
    hydraulic_vpipes()  hmap::gpu::hydraulic_vpipes (hydraulic_vpipes_gpu.cpp:15-178)

Same argument names, defaults and meaning as the C++ declarations
(HighMap/include/highmap/erosion.hpp:1206-1223,1784-1801,
HighMap/include/highmap/primitives/coherent_noise.hpp:821-834).  Arrays are
numpy float32 of shape (ny, nx), C order, which is hmap::Array's layout
(linear index j*nx + i); `shape` arguments are (nx, ny) like glm::ivec2.
Like the reference, erosion functions update `z` in place and fill the
optional output arrays that are passed.

Everything runs in libhmg.so (HIP, gfx950) through its C ABI (include/hmg.h);
there is no CPU fallback: without the library or a GPU the calls raise.
"""
import ctypes as C
import os
import threading

import numpy as np

__all__ = [
    "HmgError", "NoiseType", "Storage", "init_gpu", "noise_fbm", "thermal",
    "hydraulic_vpipes", "Session", "VpipesParams", "lib", "remap",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhmg.so")

HMG_OK = 0
_STATUS = {-1: "invalid argument", -2: "out of device memory", -3: "device error",
           -4: "non-finite value", -5: "unsupported", -6: "not ready"}


class HmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_STATUS.get(code, "error"), code, msg))
        self.code = code


class NoiseType:
    """hmap::NoiseType (HighMap/include/highmap/functions.hpp:66-81)."""
    PARBERRY, PERLIN, PERLIN_BILLOW, PERLIN_HALF, SIMPLEX2, SIMPLEX2S, VALUE, \
        VALUE_CUBIC, VALUE_DELAUNAY, VALUE_LINEAR, WORLEY, WORLEY_DOUBLE, \
        WORLEY_VALUE = range(13)


class Storage:
    F32, BF16, F16 = 0, 1, 2


class VpipesParams(C.Structure):
    """hmg_vpipes_params (include/hmg.h)."""

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
        ("iterations", C.c_int),
        ("mode", C.c_int),
    ]

    MODES = {"fused": 0, "plain": 1}

    @classmethod
    def make(cls, **kw):
        p = cls()
        lib().hmg_vpipes_defaults(C.byref(p))
        for k, v in kw.items():
            if k == "mode" and isinstance(v, str):
                v = cls.MODES[v]
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, type(getattr(p, k))(v))
        return p


_lib = None
_lib_lock = threading.Lock()


def lib():
    """Loads libhmg.so (never builds it: run highmap_amd.build first)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise HmgError(-3, "%s is missing; build it with `python -m highmap_amd.build`"
                           % LIB_PATH)
        # One HIP runtime per process: torch ships its own libamdhip64 with the
        # same soname; loading torch first makes libhmg bind to that copy, so
        # torch tensors and libhmg share the runtime (a second runtime loaded
        # after ours would find no GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i32, f32, u32, sz = C.c_void_p, C.c_int, C.c_float, C.c_uint32, C.c_size_t
        pp = C.POINTER(VpipesParams)
        sig = {
            "hmg_init": ([i32, C.POINTER(vp)], i32),
            "hmg_destroy": ([vp], i32),
            "hmg_last_error": ([], C.c_char_p),
            "hmg_abi_version": ([], i32),
            "hmg_vpipes_defaults": ([pp], None),
            "hmg_sync": ([vp], i32),
            "hmg_stream": ([vp], vp),
            "hmg_noise_fbm": ([vp, vp, i32, i32, i32, f32, f32, u32, i32, f32, f32, f32,
                               vp, vp, vp, vp, vp], i32),
            "hmg_noise_fbm_dev": ([vp, vp, i32, i32, i32, f32, f32, u32, i32, f32, f32, f32,
                                   vp, vp, vp, vp, vp], i32),
            "hmg_remap_dev": ([vp, vp, sz, f32, f32], i32),
            "hmg_noise": ([vp, vp, i32, i32, i32, f32, f32, u32, vp, vp, vp, vp], i32),
            "hmg_noise_dev": ([vp, vp, i32, i32, i32, f32, f32, u32, vp, vp, vp, vp], i32),
            "hmg_thermal": ([vp, vp, i32, i32, vp, f32, i32, vp, vp, vp], i32),
            "hmg_thermal_dev": ([vp, vp, i32, i32, vp, f32, i32, vp], i32),
            "hmg_thermal_dev_noalloc": ([vp, vp, i32, i32, vp, f32, i32, vp], i32),
            "hmg_thermal_set_block_steps": ([i32], i32),
            "hmg_vpipes": ([vp, vp, i32, i32, pp, vp, vp, vp, vp, vp], i32),
            "hmg_session_create": ([vp, vp, i32, i32, pp, vp, i32, C.POINTER(vp)], i32),
            "hmg_session_create_dev": ([vp, vp, i32, i32, pp, vp, i32, C.POINTER(vp)], i32),
            "hmg_session_vpipes": ([vp, i32], i32),
            "hmg_session_thermal": ([vp, f32, i32], i32),
            "hmg_session_export": ([vp, vp, vp, vp, vp, vp], i32),
            "hmg_session_sums": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)], i32),
            "hmg_session_max_backtrace": ([vp, C.POINTER(C.c_double)], i32),
            "hmg_session_z_dev": ([vp], vp),
            "hmg_session_set_tuning": ([vp, i32, i32], i32),
            "hmg_session_get_tuning": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)],
                                       i32),
            "hmg_session_sync": ([vp], i32),
            "hmg_session_destroy": ([vp], i32),
            "hmg_session_bytes": ([i32, i32, i32], sz),
            "hmg_batch_create": ([vp, i32, pp, C.POINTER(vp)], i32),
            "hmg_batch_add": ([vp, vp, i32, i32, vp, i32, C.POINTER(i32)], i32),
            "hmg_batch_set_tuning": ([vp, i32, i32], i32),
            "hmg_batch_autotune": ([vp, i32, C.POINTER(i32), C.POINTER(i32)], i32),
            "hmg_batch_run": ([vp, i32, C.POINTER(f32)], i32),
            "hmg_batch_tuned_form": ([vp, C.POINTER(i32)], i32),
            "hmg_batch_export": ([vp, i32, vp, vp, vp, vp, vp], i32),
            "hmg_batch_destroy": ([vp], i32),
            "hmg_thermal_auto_bedrock": ([vp, vp, i32, i32, vp, f32, i32, vp, vp], i32),
            "hmg_thermal_auto_bedrock_dev": ([vp, vp, i32, i32, vp, f32, i32], i32),
            "hmg_sediment_deposition": ([vp, vp, i32, i32, vp, f32, vp, f32, i32, i32, vp], i32),
            "hmg_flow_simulation": ([vp, vp, i32, i32, f32, vp, i32, f32, i32, f32, f32, vp],
                                    i32),
            "hmg_flow_simulation_dev": ([vp, vp, i32, i32, f32, vp, i32, f32, i32, f32, f32,
                                         vp], i32),
            "hmg_snow_simulation": ([vp, vp, i32, i32, f32, vp, vp, vp, i32] + [f32] * 7
                                    + [i32, f32, vp], i32),
            "hmg_snow_simulation_dev": ([vp, vp, i32, i32, f32, vp, vp, vp, i32] + [f32] * 7
                                        + [i32, f32, vp], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = L
        return L


def _check(rc):
    if rc != HMG_OK:
        raise HmgError(rc, lib().hmg_last_error().decode(errors="replace"))


def _ptr(a):
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def _f32(a, shape=None, name="array"):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError("%s has shape %s, expected %s" % (name, a.shape, shape))
    return a


def _inplace(z):
    if not (isinstance(z, np.ndarray) and z.dtype == np.float32 and z.flags.c_contiguous
            and z.ndim == 2):
        raise TypeError("z must be a C-contiguous float32 numpy array of shape (ny, nx)")
    return z


class Context:
    """An hmg context on one HIP device (its own stream and memory pool)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().hmg_init(device, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().hmg_destroy(self.handle)
            self.handle = None

    def sync(self):
        _check(lib().hmg_sync(self.handle))

    @property
    def stream(self):
        return lib().hmg_stream(self.handle)


_default_ctx = None
_ctx_lock = threading.Lock()


def init_gpu(device=0):
    """hmap::gpu::init_opencl(): returns True when a device is ready."""
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(device)
    return True


def _ctx(ctx):
    if ctx is not None:
        return ctx
    init_gpu()
    return _default_ctx


# ------------------------------------------------------------------ noise


def noise_fbm(noise_type, shape, kw, seed, octaves=8, weight=0.7, persistence=0.5,
              lacunarity=2.0, ctrl_param=None, noise_x=None, noise_y=None, stretching=None,
              bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0), ctx=None):
    """hmap::gpu::noise_fbm; returns a float32 array of shape (ny, nx).
    `stretching` is accepted and ignored, as in the reference."""
    nx, ny = int(shape[0]), int(shape[1])
    out = np.empty((ny, nx), np.float32)
    c = _f32(ctrl_param, (ny, nx), "ctrl_param")
    gx = _f32(noise_x, (ny, nx), "noise_x")
    gy = _f32(noise_y, (ny, nx), "noise_y")
    bb = np.asarray(bbox, np.float32)
    pe = np.asarray(period, np.int32)
    _check(lib().hmg_noise_fbm(_ctx(ctx).handle, _ptr(out), nx, ny, int(noise_type),
                               float(kw[0]), float(kw[1]), int(seed) & 0xFFFFFFFF, int(octaves),
                               float(weight), float(persistence), float(lacunarity),
                               _ptr(c), _ptr(gx), _ptr(gy), _ptr(bb), _ptr(pe)))
    return out


def noise(noise_type, shape, kw, seed, noise_x=None, noise_y=None, stretching=None,
          bbox=(0.0, 1.0, 0.0, 1.0), period=(0, 0), ctx=None):
    """hmap::gpu::noise (one octave); returns a float32 array of shape
    (ny, nx).  `stretching` is accepted and ignored, as in the reference."""
    nx, ny = int(shape[0]), int(shape[1])
    out = np.empty((ny, nx), np.float32)
    gx = _f32(noise_x, (ny, nx), "noise_x")
    gy = _f32(noise_y, (ny, nx), "noise_y")
    bb = np.asarray(bbox, np.float32)
    pe = np.asarray(period, np.int32)
    _check(lib().hmg_noise(_ctx(ctx).handle, _ptr(out), nx, ny, int(noise_type), float(kw[0]),
                           float(kw[1]), int(seed) & 0xFFFFFFFF, _ptr(gx), _ptr(gy), _ptr(bb),
                           _ptr(pe)))
    return out


# ---------------------------------------------------------------- thermal


def thermal(z, talus, iterations=10, bedrock=None, deposition_map=None, mask=None, ctx=None):
    """hmap::gpu::thermal; z is updated in place.  `talus` is a scalar or an
    array; `mask` gives the masked overload (apply_with_mask)."""
    z = _inplace(z)
    ny, nx = z.shape
    t = None
    ts = 0.0
    if np.isscalar(talus):
        ts = float(talus)
    else:
        t = _f32(talus, (ny, nx), "talus")
    b = _f32(bedrock, (ny, nx), "bedrock")
    m = _f32(mask, (ny, nx), "mask")
    if deposition_map is not None:
        deposition_map = _inplace(deposition_map)
        if deposition_map.shape != z.shape:
            raise ValueError("deposition_map shape mismatch")
    _check(lib().hmg_thermal(_ctx(ctx).handle, _ptr(z), nx, ny, _ptr(t), ts, int(iterations),
                             _ptr(b), _ptr(deposition_map), _ptr(m)))
    return z


def _talus(talus, shape):
    if np.isscalar(talus):
        return None, float(talus)
    return _f32(talus, shape, "talus"), 0.0


def thermal_auto_bedrock(z, talus, iterations=10, deposition_map=None, mask=None, ctx=None):
    """hmap::gpu::thermal_auto_bedrock (erosion/thermal_gpu.cpp:98-146); z is
    updated in place.  `talus` is a scalar or an array; `mask` gives the
    masked overload."""
    z = _inplace(z)
    ny, nx = z.shape
    t, ts = _talus(talus, (ny, nx))
    m = _f32(mask, (ny, nx), "mask")
    if deposition_map is not None:
        deposition_map = _inplace(deposition_map)
        if deposition_map.shape != z.shape:
            raise ValueError("deposition_map shape mismatch")
    _check(lib().hmg_thermal_auto_bedrock(_ctx(ctx).handle, _ptr(z), nx, ny, _ptr(t), ts,
                                          int(iterations), _ptr(deposition_map), _ptr(m)))
    return z


def sediment_deposition(z, talus, deposition_map=None, max_deposition=0.01, iterations=5,
                        thermal_subiterations=10, mask=None, ctx=None):
    """hmap::gpu::sediment_deposition (erosion/deposition.cpp:19-63); z is
    updated in place."""
    z = _inplace(z)
    ny, nx = z.shape
    t, ts = _talus(talus, (ny, nx))
    m = _f32(mask, (ny, nx), "mask")
    if deposition_map is not None:
        deposition_map = _inplace(deposition_map)
        if deposition_map.shape != z.shape:
            raise ValueError("deposition_map shape mismatch")
    _check(lib().hmg_sediment_deposition(_ctx(ctx).handle, _ptr(z), nx, ny, _ptr(t), ts,
                                         _ptr(deposition_map), float(max_deposition),
                                         int(iterations), int(thermal_subiterations), _ptr(m)))
    return z


# ----------------------------------------------------------------- vpipes


def hydraulic_vpipes(z, water_height=1e-2, maintain_water_volume=True, evap_rate=0.1,
                     iterations=50, dt=0.5, k_capacity=0.5, k_erode=1e-3, k_depose=1e-2,
                     k_discharge_exp=1.0, downcutting_max_depth_ratio=10.0,
                     flux_diffusion=True, flux_diffusion_strength=0.01, rain_map=None,
                     water_depth=None, sediment=None, vel_u=None, vel_v=None,
                     gravity=1.0, rain_per_iter=0.0, mode="fused", ctx=None):
    """hmap::gpu::hydraulic_vpipes; z is updated in place and the optional
    output arrays (water_depth, sediment, vel_u, vel_v) are filled.
    gravity, rain_per_iter and mode are extensions (reference: 1, 0)."""
    z = _inplace(z)
    ny, nx = z.shape
    p = VpipesParams.make(water_height=water_height,
                          maintain_water_volume=int(bool(maintain_water_volume)),
                          evap_rate=evap_rate, iterations=iterations, dt=dt,
                          k_capacity=k_capacity, k_erode=k_erode, k_depose=k_depose,
                          k_discharge_exp=k_discharge_exp,
                          downcutting_max_depth_ratio=downcutting_max_depth_ratio,
                          flux_diffusion=int(bool(flux_diffusion)),
                          flux_diffusion_strength=flux_diffusion_strength, gravity=gravity,
                          rain_per_iter=rain_per_iter, mode=mode)
    r = _f32(rain_map, (ny, nx), "rain_map")
    outs = []
    for a in (water_depth, sediment, vel_u, vel_v):
        if a is not None:
            a = _inplace(a)
            if a.shape != z.shape:
                raise ValueError("output shape mismatch")
        outs.append(a)
    _check(lib().hmg_vpipes(_ctx(ctx).handle, _ptr(z), nx, ny, C.byref(p), _ptr(r),
                            *[_ptr(a) for a in outs]))
    return z


def flow_simulation(z, water_height, depth_map, iterations, dt=0.5, flux_diffusion=True,
                    flux_diffusion_strength=0.01, dry_out_ratio=0.0, ctx=None):
    """hmap::gpu::flow_simulation (hydrology/flow_simulation.cpp:26-110):
    returns the water depth after `iterations` flow + water passes starting
    from water_height * depth_map, then the optional dry-out.  z and
    depth_map are not changed."""
    zz = _f32(z, None, "z")
    ny, nx = zz.shape
    dm = _f32(depth_map, (ny, nx), "depth_map")
    d = np.empty((ny, nx), np.float32)
    _check(lib().hmg_flow_simulation(_ctx(ctx).handle, _ptr(zz), nx, ny, float(water_height),
                                     _ptr(dm), int(iterations), float(dt),
                                     int(bool(flux_diffusion)), float(flux_diffusion_strength),
                                     float(dry_out_ratio), _ptr(d)))
    return d


def snow_simulation(z, snow_depth, fall_map, melting_map, talus, iterations, dt=0.5,
                    fall_iterations_ratio=1.0, k_snow=0.5, k_visc=0.1, k_melt_factor=0.8,
                    k_depth_ratio=1.0, k_depth_slope_ratio=1.0, post_filter=True,
                    thermal_talus_ratio=0.2, ctx=None):
    """hmap::gpu::snow_simulation (hydrology/snow_simulation.cpp:81-154):
    returns the snow depth after `iterations` steps of fall, avalanching,
    viscosity and melting, then the optional thermal post-filter.
    fall_map, melting_map and talus are planes of z's shape (scalars are
    broadcast); the inputs are not changed."""
    zz = _f32(z, None, "z")
    ny, nx = zz.shape

    def plane(a, name):
        return _f32(np.broadcast_to(np.asarray(a, np.float32), (ny, nx)), (ny, nx), name)

    f, m, t = plane(fall_map, "fall_map"), plane(melting_map, "melting_map"), plane(talus, "talus")
    out = np.empty((ny, nx), np.float32)
    _check(lib().hmg_snow_simulation(_ctx(ctx).handle, _ptr(zz), nx, ny, float(snow_depth),
                                     _ptr(f), _ptr(m), _ptr(t), int(iterations), float(dt),
                                     float(fall_iterations_ratio), float(k_snow), float(k_visc),
                                     float(k_melt_factor), float(k_depth_ratio),
                                     float(k_depth_slope_ratio), int(bool(post_filter)),
                                     float(thermal_talus_ratio), _ptr(out)))
    return out


class Session:
    """A device-resident erosion state: hydraulic iterations and thermal
    steps interleave on it without host transfers (hmg_session_*)."""

    def __init__(self, z, params=None, rain_map=None, storage=Storage.F32, ctx=None, **kw):
        self.ctx = _ctx(ctx)
        if params is None:
            params = VpipesParams.make(**kw)
        self.params = params
        h = C.c_void_p()
        if hasattr(z, "data_ptr"):  # device tensor (z on the context's device)
            ny, nx = z.shape
            rp = None if rain_map is None else C.c_void_p(rain_map.data_ptr())
            _check(lib().hmg_session_create_dev(self.ctx.handle, C.c_void_p(z.data_ptr()), nx,
                                                ny, C.byref(params), rp, int(storage),
                                                C.byref(h)))
        else:
            z = _f32(z)
            ny, nx = z.shape
            r = _f32(rain_map, (ny, nx), "rain_map")
            _check(lib().hmg_session_create(self.ctx.handle, _ptr(z), nx, ny, C.byref(params),
                                            _ptr(r), int(storage), C.byref(h)))
        self.handle = h
        self.nx, self.ny = nx, ny
        self.storage = storage

    def vpipes(self, iterations):
        _check(lib().hmg_session_vpipes(self.handle, int(iterations)))
        return self

    def thermal(self, talus, iterations=1):
        _check(lib().hmg_session_thermal(self.handle, float(talus), int(iterations)))
        return self

    def sync(self):
        _check(lib().hmg_session_sync(self.handle))

    def export(self):
        """Returns (z, d, s, u, v) as hydraulic_vpipes would."""
        shp = (self.ny, self.nx)
        out = [np.empty(shp, np.float32) for _ in range(5)]
        _check(lib().hmg_session_export(self.handle, *[_ptr(a) for a in out]))
        return tuple(out)

    def sums(self):
        """(sum z, sum s, sum d) of the state hydraulic_vpipes would return,
        summed exactly on the device (fixed point) and given in FP64."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().hmg_session_sums(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def max_backtrace(self):
        """max over the cells of dt * max(|u|, |v|) (FP64): how many cells the
        next sediment backtrace reaches."""
        a = C.c_double()
        _check(lib().hmg_session_max_backtrace(self.handle, C.byref(a)))
        return a.value

    def set_tuning(self, chunk_rows, prefetch=1):
        _check(lib().hmg_session_set_tuning(self.handle, int(chunk_rows), int(prefetch)))
        return self

    def tuning(self):
        """(chunk_rows, prefetch, still_tuning) of the fused kernel."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().hmg_session_get_tuning(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, bool(c.value)

    def z_dev_ptr(self):
        return lib().hmg_session_z_dev(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            lib().hmg_session_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent maps advanced together by batched launches
    (hmg_batch_*); each map's result equals running it alone."""

    def __init__(self, storage=Storage.F32, params=None, ctx=None, **kw):
        self.ctx = _ctx(ctx)
        self.params = params if params is not None else VpipesParams.make(**kw)
        h = C.c_void_p()
        _check(lib().hmg_batch_create(self.ctx.handle, int(storage), C.byref(self.params),
                                      C.byref(h)))
        self.handle = h
        self.shapes = []

    def add(self, z, iterations, rain_map=None):
        z = _f32(z)
        ny, nx = z.shape
        r = _f32(rain_map, (ny, nx), "rain_map")
        mid = C.c_int()
        _check(lib().hmg_batch_add(self.handle, _ptr(z), nx, ny, _ptr(r), int(iterations),
                                   C.byref(mid)))
        self.shapes.append((ny, nx))
        return mid.value

    def set_tuning(self, chunk_rows, prefetch=1):
        _check(lib().hmg_batch_set_tuning(self.handle, int(chunk_rows), int(prefetch)))

    def autotune(self, probe=2):
        a, b = C.c_int(), C.c_int()
        _check(lib().hmg_batch_autotune(self.handle, int(probe), C.byref(a), C.byref(b)))
        return a.value, b.value

    STREAM, GRAPH, PERSISTENT, AUTO = 0, 1, 2, 3  # forms of run (HMG_BATCH_*)

    def run(self, use_graph=True, form=None):
        """Runs every map to its iteration count; returns device ms.  form
        (HMG_BATCH_*) overrides use_graph: STREAM, GRAPH, PERSISTENT (one
        cooperative launch, a grid barrier between iterations) or AUTO (what
        autotune picked)."""
        f = int(form) if form is not None else (self.GRAPH if use_graph else self.STREAM)
        ms = C.c_float()
        _check(lib().hmg_batch_run(self.handle, f, C.byref(ms)))
        return ms.value

    def tuned_form(self):
        f = C.c_int()
        _check(lib().hmg_batch_tuned_form(self.handle, C.byref(f)))
        return f.value

    def export(self, map_id):
        shp = self.shapes[map_id]
        out = [np.empty(shp, np.float32) for _ in range(5)]
        _check(lib().hmg_batch_export(self.handle, int(map_id), *[_ptr(a) for a in out]))
        return tuple(out)

    def close(self):
        if getattr(self, "handle", None):
            lib().hmg_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def remap(a_dev_ptr, n, vmin=0.0, vmax=1.0, ctx=None):
    """hmap::remap on a device buffer of n float32 values."""
    _check(lib().hmg_remap_dev(_ctx(ctx).handle, C.c_void_p(a_dev_ptr), n, vmin, vmax))
