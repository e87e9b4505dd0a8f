"""Builds libhmg.so (HIP kernels + C ABI, gfx950) and libhmap_gpu.so (the
C++ hmap::gpu drop-in over it), in-tree.

    python -m highmap_amd.build          # or highmap_amd.build.build()

Every HIP source is compiled with:
  -ffp-contract=off                        one IEEE operation per source op
  -fhip-fp32-correctly-rounded-divide-sqrt FP32 / and sqrt correctly rounded
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed FP32 instructions
so the kernels reproduce the FP32 restatement of the reference bit for bit.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhmg.so")
# HBM stream kernels used only by bench.py's measurements
PROBE_LIB = os.path.join(HERE, "libhmg_probe.so")
SHIM_LIB = os.path.join(HERE, "libhmap_gpu.so")
ARCH = os.environ.get("HMG_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["runtime.hip", "noise.hip", "thermal.hip", "vpipes.hip", "server.hip", "snow.hip"]

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I" + INC,
    "-I" + CSRC,
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def _compile_one(src):
    hipcc = _hipcc()
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([hipcc] + HIPCC_FLAGS + lang + ["-c", path, "-o", obj])
    return obj


def build(verbose=False):
    """Compile every source and link both libraries.  Always recompiles."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    probe = os.path.join(CSRC, "hbm_probe.hip")
    if os.path.exists(probe):
        tmp = PROBE_LIB + ".tmp"
        _run([_hipcc()] + HIPCC_FLAGS + ["-shared", "-x", "hip", probe, "-o", tmp])
        os.replace(tmp, PROBE_LIB)
    shim = os.path.join(CSRC, "hmap_gpu.cpp")
    if os.path.exists(shim):
        tmp = SHIM_LIB + ".tmp"
        va = os.path.join(CSRC, "virtual_array.cpp")
        _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
              "-I" + INC, shim, va, "-o", tmp, "-L" + HERE, "-lhmg",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, SHIM_LIB)
        bench = os.path.join(HERE, "va_bench")
        _run(["g++", "-std=c++20", "-O2", "-pthread", "-I" + INC,
              os.path.join(CSRC, "va_bench.cpp"), "-o", bench + ".tmp", "-L" + HERE,
              "-lhmap_gpu", "-lhmg", "-Wl,-rpath,$ORIGIN"])
        os.replace(bench + ".tmp", bench)
        for name in ("bf16_check", "arith_check", "pow_check"):
            chk = os.path.join(ROOT, "tests", "cpp", name)
            _run([_hipcc()] + HIPCC_FLAGS + ["-x", "hip", chk + ".cpp", "-o", chk + ".tmp"])
            os.replace(chk + ".tmp", chk)
        for name in ("shim_driver", "io_driver", "va_driver"):
            drv = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
            if os.path.exists(drv):
                out = os.path.join(ROOT, "tests", "cpp", name)
                _run(["g++", "-std=c++20", "-O2", "-pthread", "-I" + INC, drv, "-o", out + ".tmp",
                      "-L" + HERE, "-lhmap_gpu", "-lhmg",
                      "-Wl,-rpath,$ORIGIN/../../highmap_amd"])
                os.replace(out + ".tmp", out)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
