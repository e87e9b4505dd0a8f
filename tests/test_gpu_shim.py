"""The C++ drop-in (include/hmap_gpu.hpp, libhmap_gpu.so) called as HighMap's
own examples call hmap::gpu::*, compared bit for bit with the FP32 oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "cpp", "shim_driver")


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    d = tmp_path_factory.mktemp("shim")
    assert os.path.exists(DRIVER), "tests/cpp/shim_driver not built (run the build first)"
    r = subprocess.run([DRIVER, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    def load(name, shape):
        return np.fromfile(os.path.join(d, name + ".f32"), dtype=np.float32).reshape(shape)

    return load


def test_shim_noise_and_thermal(out, orc):
    n = out("noise256", (256, 256))
    assert np.array_equal(n, orc.noise_fbm(256, 256, (4.0, 4.0), 1))
    z0 = orc.remap(n)
    ref = orc.thermal(z0, np.float32(2.0) / np.float32(256.0), 50)
    assert np.array_equal(out("thermal256", (256, 256)), ref)
    dep = ref - z0
    assert np.array_equal(out("deposition256", (256, 256)),
                          np.where(dep < 0, np.float32(0), dep))


def test_shim_masked_thermal(out, orc):
    z0 = orc.remap(out("noise256", (256, 256)))
    talus = np.full((256, 256), np.float32(1.0) / np.float32(256.0), np.float32)
    r = orc.thermal(z0, talus, 7)
    mask = np.tile(np.arange(256, dtype=np.float32) / np.float32(255), (256, 1))
    assert np.array_equal(out("thermal_masked256", (256, 256)),
                          z0 * (np.float32(1) - mask) + r * mask)


def test_shim_vpipes(out, orc):
    shp = (128, 192)
    zin = out("vp_in", shp)
    ref = orc.remap(orc.noise_fbm(192, 128, (4.0, 4.0), 2))
    assert np.array_equal(zin, ref)
    st = orc.VpState(zin, orc.vp_params(water_height=1e-3, evap_rate=0.01, k_capacity=0.1,
                                        k_erode=0.05, k_depose=0.05)).run(40)
    for name, plane in (("vp_z", st.z), ("vp_d", st.d), ("vp_s", st.s), ("vp_u", st.u),
                        ("vp_v", st.v)):
        assert np.array_equal(out(name, shp), plane), name
    dft = orc.VpState(zin, orc.vp_params()).run(50)
    assert np.array_equal(out("vp_default_z", shp), dft.z)


def test_shim_flow_simulation(out, orc):
    z = orc.remap(orc.noise_fbm(256, 256, (4.0, 4.0), 2))
    want = orc.flow_simulation(z, 1e-2, np.ones_like(z), 400)
    assert np.array_equal(out("flow256", (256, 256)), want)


def test_shim_thermal_auto_bedrock(out, orc):
    zin = out("auto_in", (256, 256))
    assert np.array_equal(zin, orc.noise_fbm(256, 256, (2.0, 2.0), 1))
    want = orc.thermal_auto_bedrock(zin, np.float32(0.5) / np.float32(256), 500)
    assert np.array_equal(out("auto256", (256, 256)), want)


def test_shim_sediment_deposition(out, orc):
    z0 = orc.noise_fbm(256, 256, (4.0, 4.0), 1)
    talus = np.full((256, 256), np.float32(2.0) / np.float32(256.0), np.float32)
    z, dep = orc.sediment_deposition(z0, talus, 0.01, 10, 50)
    assert np.array_equal(out("sed256", (256, 256)), z)
    assert np.array_equal(out("sed_dep256", (256, 256)), dep)


def test_shim_snow_simulation(out, orc):
    zin = out("snow_in", (256, 256))
    assert np.array_equal(zin, orc.remap(orc.noise_fbm(256, 256, (4.0, 4.0), 2)))
    talus = np.float32(2.5) / np.float32(256)
    ref0 = orc.snow_simulation(zin, 0.1, 1.0, 0.0, talus, 500)
    assert np.array_equal(out("snow0", (256, 256)), ref0)
    melt = np.repeat((np.arange(256, dtype=np.float32) / np.float32(255))[:, None], 256, axis=1)
    ref1 = orc.snow_simulation(zin, 0.1, 1.0, melt, talus, 500)
    assert np.array_equal(out("snow1", (256, 256)), ref1)
