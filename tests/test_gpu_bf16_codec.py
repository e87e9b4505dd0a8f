"""The kernels' BF16 storage codec (the conversion instruction, with the
integer rule for NaNs) against the integer round-to-nearest-even form, over
all 2^32 float patterns on the GPU (tests/cpp/bf16_check.cpp)."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_bf16_codec_equals_integer_rounding_everywhere():
    exe = os.path.join(ROOT, "tests", "cpp", "bf16_check")
    assert os.path.exists(exe), "tests/cpp/bf16_check not built (run the build first)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("differences 0 "), r.stdout
