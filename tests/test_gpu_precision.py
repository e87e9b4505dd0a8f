"""Precision bounds of the three storage types against the FP64 oracle.

The 4096^2 vpipes + thermal state (fBm seed 2, thermal with talus 1/4096
after every 10th hydraulic iteration, hmap::gpu::hydraulic_vpipes default
parameters) runs on the device with FP32, BF16 and FP16 storage of water,
sediment and fluxes (FP32 arithmetic throughout) and on the host in FP64.
At iterations 10 and 50 each storage type must stay within a bound on
  - max |z - z_fp64| / range(z_fp64), and
  - |sum(z + s) - sum_fp64(z + s)| / sum_fp64(z + s).
The bounds are about 2.5x what these runs give (the device state is bit
for bit the FP32 oracle with the same storage rounding, so the numbers are
reproducible): FP32 3.7e-7 / 1.1e-4 of the range, FP16 2.1e-6 / 3.5e-4,
BF16 2.2e-5 / 1.2e-3; mass 5e-12 / 5e-11 (FP32), 3e-10 / 2.4e-7 (FP16),
2.5e-8 / 3.1e-5 (BF16).  Beyond about 50 iterations these parameters let
one rounding difference grow to O(range) for any FP32 implementation, so
no bound is asserted there.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096
STORAGES = {"f32": 0, "bf16": 1, "f16": 2}
# iteration -> storage -> (z error / range, mass error)
BOUNDS = {
    10: {"f32": (1e-6, 1.5e-11), "f16": (5e-6, 1e-9), "bf16": (5e-5, 7e-8)},
    50: {"f32": (3e-4, 1.5e-10), "f16": (1e-3, 6e-7), "bf16": (3e-3, 8e-5)},
}


@pytest.fixture(scope="module")
def errors(hm, orc):
    z0 = orc.fbm_terrain(N, N, 2)
    o = orc.VpState(z0.astype(np.float64), orc.vp_params(), dtype=np.float64)
    sess = {k: hm.Session(z0, storage=v) for k, v in STORAGES.items()}
    out = {}
    for it in range(10, 51, 10):
        o.run(10).thermal_step(1.0 / N, 1)
        for s in sess.values():
            s.vpipes(10).thermal(1.0 / N, 1)
        if it not in BOUNDS:
            continue
        rng = float(o.z.max() - o.z.min())
        m64 = float((o.z + o.s).sum())
        for k, s in sess.items():
            z, _, sed, _, _ = s.export()
            zerr = float(np.abs(z - o.z).max()) / rng
            merr = abs(float((z.astype(np.float64) + sed).sum()) - m64) / m64
            out[it, k] = (zerr, merr)
            print("iteration %d %s: z error %.3g of range, mass %.3g" % (it, k, zerr, merr),
                  flush=True)
    for s in sess.values():
        s.close()
    return out


@pytest.mark.parametrize("it", sorted(BOUNDS))
@pytest.mark.parametrize("storage", sorted(STORAGES))
def test_storage_error_vs_fp64_is_bounded(errors, it, storage):
    zerr, merr = errors[it, storage]
    zmax, mmax = BOUNDS[it][storage]
    assert zerr <= zmax, (it, storage, zerr)
    assert merr <= mmax, (it, storage, merr)


def test_wider_storage_is_closer(errors):
    """At each checkpoint FP32 < FP16 < BF16 in z error (mantissas of 24,
    11 and 8 bits for water, sediment and fluxes)."""
    for it in BOUNDS:
        e = {k: errors[it, k][0] for k in STORAGES}
        assert e["f32"] < e["f16"] < e["bf16"], (it, e)
