"""The thermal kernels divide by the diagonal distance 1.414f with a product
and two fma corrections instead of a full IEEE division.  This checks the
constants shipped in hmg_cell.hpp and, over all 2^32 float inputs on the
CPU, that the sequence equals the correctly rounded quotient wherever the
kernel uses it (2^-100 <= |x| < inf; other inputs take the division)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT


def _constants():
    src = open(os.path.join(ROOT, "highmap_amd", "csrc", "hmg_cell.hpp")).read()
    body = src[src.index("float div_1414(float x)"):]
    b = re.search(r"constexpr float B = ([0-9.]+)f;", body).group(1)
    r = re.search(r"constexpr float R = (0x[0-9a-fA-F.]+p[-+]?[0-9]+)f;", body).group(1)
    assert "0x1p-100f" in body[:800]
    return b, r


def _cpu_has_fma():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return " fma " in flags or " fma\n" in flags


@pytest.mark.skipif(not _cpu_has_fma(), reason="needs a host CPU with FMA instructions")
def test_division_by_1414_exhaustive(tmp_path):
    b, r = _constants()
    exe = tmp_path / "div_const_check"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "cpp", "div_const_check.c"), "-o", str(exe),
                    "-lm"], check=True)
    out = subprocess.run([str(exe), b, r], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches=0" in out.stdout
