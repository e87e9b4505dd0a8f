"""The kernels' short arithmetic forms against the full ones over every
float pattern they can receive, on the GPU (tests/cpp/arith_check.cpp):
the square root without small-input scaling (bit-equal to sqrtf from 2^-96
up, below 2^-47 under it), its wave-checked form (bit-equal everywhere), and
the float form of the fixed-point volume term (equal to the integer form on
[0, 1))."""
import os
import subprocess

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
