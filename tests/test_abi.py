"""The C ABI library loads and exports every entry point include/*.h declares.
CPU-only: nothing here launches device work."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(hmg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_core_entry_points():
    names = _declared("hmg.h")
    for n in ("hmg_init", "hmg_noise_fbm", "hmg_thermal", "hmg_vpipes", "hmg_session_create",
              "hmg_session_vpipes", "hmg_session_thermal", "hmg_session_export",
              "hmg_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from highmap_amd import LIB_PATH

    assert os.path.exists(LIB_PATH), "libhmg.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (hmg_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in _declared("hmg.h") if n not in exported]
    assert not missing, missing


def test_library_loads_and_answers_without_gpu():
    import highmap_amd

    L = highmap_amd.lib()
    assert L.hmg_abi_version() == 1
    p = highmap_amd.VpipesParams()
    L.hmg_vpipes_defaults(ctypes.byref(p))
    # defaults of hmap::gpu::hydraulic_vpipes (erosion.hpp:1206-1223)
    f32 = lambda x: float(np.float32(x))
    assert p.water_height == f32(1e-2) and p.maintain_water_volume == 1
    assert p.evap_rate == f32(0.1) and p.iterations == 50 and p.dt == 0.5
    assert p.k_capacity == 0.5 and p.k_erode == f32(1e-3) and p.k_depose == f32(1e-2)
    assert p.k_discharge_exp == 1.0 and p.downcutting_max_depth_ratio == 10.0
    assert p.flux_diffusion == 1 and p.flux_diffusion_strength == f32(0.01)
    assert p.gravity == 1.0 and p.rain_per_iter == 0.0 and p.mode == 0


def test_cpp_dropin_exports_reference_signatures():
    """libhmap_gpu.so defines hmap::gpu::* with the reference's parameter
    lists (HighMap/include/highmap/erosion.hpp:1206-1223,1784-1801,
    coherent_noise.hpp:821-834)."""
    lib = os.path.join(ROOT, "highmap_amd", "libhmap_gpu.so")
    assert os.path.exists(lib), "libhmap_gpu.so not built"
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    want = [
        "hmap::gpu::init_opencl()",
        "hmap::gpu::noise_fbm(hmap::NoiseType, glm::ivec2, glm::vec2, unsigned int, int, float, "
        "float, float, hmap::Array const*, hmap::Array const*, hmap::Array const*, "
        "hmap::Array const*, glm::vec4, glm::ivec2)",
        "hmap::gpu::thermal(hmap::Array&, hmap::Array const&, int, hmap::Array*, hmap::Array*)",
        "hmap::gpu::thermal(hmap::Array&, hmap::Array const*, hmap::Array const&, int, "
        "hmap::Array*, hmap::Array*)",
        "hmap::gpu::thermal(hmap::Array&, float, int, hmap::Array*, hmap::Array*)",
        "hmap::gpu::hydraulic_vpipes(hmap::Array&, float, bool, float, int, float, float, float, "
        "float, float, float, bool, float, hmap::Array*, hmap::Array*, hmap::Array*, "
        "hmap::Array*, hmap::Array*)",
    ]
    for w in want:
        assert w in out, w


@pytest.mark.slow
def test_device_code_has_no_single_rounding_mixed_precision_ops():
    """The kernels must round to FP32 before any FP16 conversion; the
    compiler's mixed-precision fma (v_fma_mix*/v_mad_mix*) rounds once and
    would change results.  Checks the gfx950 assembly of every source."""
    import shutil

    from highmap_amd import build as hb

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    assert hb.HIP_SOURCES, "no HIP sources listed"
    for src in hb.HIP_SOURCES:
        path = os.path.join(hb.CSRC, src)
        assert os.path.exists(path), src
        flags = [f for f in hb.HIPCC_FLAGS if f != "-fPIC"]
        asm = subprocess.run([hipcc] + flags + ["-x", "hip", "-S", "--cuda-device-only", path,
                                                "-o", "-"], capture_output=True, text=True,
                             check=True).stdout
        assert not re.search(r"\bv_(fma|mad)_mix", asm), src


def test_session_bytes_accounting():
    import highmap_amd

    L = highmap_amd.lib()
    n = 4096 * 4096
    assert L.hmg_session_bytes(4096, 4096, 0) >= 2 * n * (4 + 6 * 4)
    assert L.hmg_session_bytes(4096, 4096, 1) < L.hmg_session_bytes(4096, 4096, 0)


def test_no_device_is_an_error_not_a_fallback():
    """Without a GPU, init reports a device error; nothing computes on the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import highmap_amd

    with pytest.raises(highmap_amd.HmgError):
        highmap_amd.Context(0)
