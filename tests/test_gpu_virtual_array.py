"""hmap::gpu functions run tile by tile over a hmap::VirtualArray (the C++
drop-in), as examples/ex_virtual_array does, in sequential and distributed
modes, with RAM tiles and with the disk-backed store that keeps two tiles in
memory: every tile bit for bit equal to the FP32 oracle run on that tile.
Thermal in the downscaled single-array mode (half the cells per axis)
equals the oracle's thermal on the restated resampling."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import virtual_array as V

pytestmark = pytest.mark.gpu

F = np.float32
DRIVER = os.path.join(ROOT, "tests", "cpp", "va_driver")


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    d = tmp_path_factory.mktemp("vagpu")
    assert os.path.exists(DRIVER), "tests/cpp/va_driver not built (run the build first)"
    r = subprocess.run([DRIVER, "gpu", str(d)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr

    def load(name):
        return np.fromfile(os.path.join(d, name + ".f32"), dtype=np.float32).reshape(1024, 1024)

    return load


@pytest.fixture(scope="module")
def want(orc):
    va = V.VA((1024, 1024), (0.0, 1.0, 0.0, 1.0), (220, 312), 32)
    res = {}
    for r in va.regions():
        va.tile(r)[...] = orc.noise_fbm(r.shape[0], r.shape[1], (2.0, 2.0), 0, 8, 0.7, 0.5, 2.0,
                                        bbox=r.bbox)
    res["noise"] = va.to_array()
    for r in va.regions():
        va.tile(r)[...] = orc.thermal(va.tile(r), F(2.0) / F(1024.0), 20)
    res["thermal"] = va.to_array()
    p = orc.vp_params(water_height=1e-3, maintain_water_volume=1, evap_rate=0.01)
    for r in va.regions():
        va.tile(r)[...] = orc.VpState(va.tile(r), p).run(20).z
    res["vpipes"] = va.to_array()
    va.smooth_overlap_buffers()
    va.remap(0.0, 1.0)
    res["final"] = va.to_array()
    return res


@pytest.mark.parametrize("stage", ["noise", "thermal", "vpipes", "final"])
@pytest.mark.parametrize("mode", ["seq", "dist", "disk"])
def test_tiles_match_oracle(out, want, stage, mode):
    assert np.array_equal(out(stage + "_" + mode), want[stage]), (stage, mode)


def test_thermal_in_downscaled_mode_matches_oracle(out, orc):
    va = V.VA((1024, 1024), (0.0, 1.0, 0.0, 1.0), (220, 312), 32)
    for r in va.regions():
        va.tile(r)[...] = orc.noise_fbm(r.shape[0], r.shape[1], (2.0, 2.0), 0, 8, 0.7, 0.5, 2.0,
                                        bbox=r.bbox)
    va.single_array(lambda a: orc.thermal(a, F(2.0) / F(512.0), 20), k_cutoff=0.5)
    assert np.array_equal(out("thermal_downscaled"), va.to_array())
