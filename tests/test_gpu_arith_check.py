"""The kernels' short arithmetic forms against the full ones over every
float pattern they can receive, on the GPU (tests/cpp/arith_check.cpp):
the square root without small-input scaling (bit-equal to sqrtf from 2^-96
up, below 2^-47 under it), its wave-checked form (bit-equal everywhere), and
the float form of the fixed-point volume term (equal to the integer form on
[0, 1)); and the kernels' pow, sin and cos against the oracle's copies,
bit for bit (tests/cpp/pow_check.cpp)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_short_arithmetic_forms_equal_full_forms_everywhere():
    exe = os.path.join(ROOT, "tests", "cpp", "arith_check")
    assert os.path.exists(exe), "tests/cpp/arith_check not built (run the build first)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout.split()
    counts = {out[i]: int(out[i + 1]) for i in range(0, len(out), 4)}
    assert counts == {"sqrt_big": 0, "sqrt_small": 0, "sqrt_sumsq": 0, "fix": 0}, r.stdout


def _pow_inputs():
    """x: every 4099th float pattern from +0 to +inf (subnormals, the normal
    range, the largest floats), their negatives, and the special values;
    y: exponents the erosion pass can see (k_discharge_exp) and C99's
    special-case exponents.  About 1.6e6 pairs."""
    pat = np.arange(0, 0x7f800001, 4099, dtype=np.uint64).astype(np.uint32)
    x = pat.view(np.float32)
    specials = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 0.5, 2.0, 1e-45, 3.4e38],
                        dtype=np.float32)
    x = np.concatenate([x[::7], -x[::97], specials])
    ys = np.array([1.3, 0.7, 1.5, 2.0, 0.5, 2.5, -0.7, 3.0, 10.0, 1.0000001, 0.0, -0.0, 1.0, -1.0,
                   np.inf, -np.inf, np.nan, 127.9, -40.0, 5.0], dtype=np.float32)
    xx = np.tile(x, ys.size)
    yy = np.repeat(ys, x.size)
    return xx, yy


def test_pow_device_equals_oracle(orc):
    exe = os.path.join(ROOT, "tests", "cpp", "pow_check")
    assert os.path.exists(exe), "tests/cpp/pow_check not built (run the build first)"
    x, y = _pow_inputs()
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([x.size], dtype=np.int32).tofile(f)
            x.tofile(f)
            y.tofile(f)
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        got = np.fromfile(fout, dtype=np.float32)
    want = orc.powf_many(x, y)
    assert got.size == want.size
    both_nan = np.isnan(got) & np.isnan(want)
    same = (got.view(np.uint32) == want.view(np.uint32)) | both_nan
    bad = np.flatnonzero(~same)
    assert bad.size == 0, "%d of %d differ, first x=%r y=%r device %r oracle %r" % (
        bad.size, x.size, x[bad[0]], y[bad[0]], got[bad[0]], want[bad[0]])


def _sincos_inputs():
    """Every 1031st float pattern from +0 to +inf and their negatives (the
    reduction's quadrant beyond 2^31 pi/2 included), every 13th pattern of
    [3e9, 1e10] where a quadrant taken by an out-of-range conversion would
    show, a dense sweep of [-8 pi, 8 pi], and the special values."""
    pat = np.arange(0, 0x7f800001, 1031, dtype=np.uint64).astype(np.uint32).view(np.float32)
    lo, hi = np.array([3e9, 1e10], dtype=np.float32).view(np.uint32)
    band = np.arange(lo, hi, 13, dtype=np.uint64).astype(np.uint32).view(np.float32)
    small = np.linspace(-8 * np.pi, 8 * np.pi, 200001, dtype=np.float32)
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 2.0 ** 25, -2.0 ** 25, 2.0 ** 31,
                         3.4e38], dtype=np.float32)
    return np.concatenate([pat, -pat, band, -band, small, specials])


def test_sincos_device_equals_oracle(orc):
    exe = os.path.join(ROOT, "tests", "cpp", "pow_check")
    assert os.path.exists(exe), "tests/cpp/pow_check not built (run the build first)"
    x = _sincos_inputs()
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([x.size], dtype=np.int32).tofile(f)
            x.tofile(f)
            x.tofile(f)
        r = subprocess.run([exe, fin, fout, "sincos"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        got = np.fromfile(fout, dtype=np.float32).reshape(3, x.size)
    ws, wc = orc.sincos_many(x)
    for name, g, w in (("hm_sin", got[0], ws), ("hm_sincos sin", got[1], ws),
                       ("hm_sincos cos", got[2], wc)):
        same = (g.view(np.uint32) == w.view(np.uint32)) | (np.isnan(g) & np.isnan(w))
        bad = np.flatnonzero(~same)
        assert bad.size == 0, "%s: %d of %d differ, first x=%r device %r oracle %r" % (
            name, bad.size, x.size, x[bad[0]], g[bad[0]], w[bad[0]])
