"""hmg_thermal_dev_noalloc (the allocation-free thermal entry point, second
Jacobi plane in a static per-device array) against hmg_thermal_dev (second
plane from the context's pool), through the C ABI, bit for bit.  Two calls
in a row, the second on a smaller grid, so the second sees the array as the
first left it; odd and even step counts (the odd ones end in the array and
are copied back); the LDS-blocked path (below 2^18 cells) and the one-step
path; a talus plane; and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INVALID = -1


def _run(hm, fn, z0, talus, iterations, talus_plane=None):
    import torch

    from highmap_amd import lib

    ny, nx = z0.shape
    z = torch.from_numpy(z0).cuda()
    tp = None if talus_plane is None else torch.from_numpy(talus_plane).cuda()
    ctx = hm._default_ctx
    rc = getattr(lib(), fn)(ctx.handle, C.c_void_p(z.data_ptr()), nx, ny,
                            None if tp is None else C.c_void_p(tp.data_ptr()), float(talus),
                            int(iterations), None)
    assert rc == 0, lib().hmg_last_error()
    assert lib().hmg_sync(ctx.handle) == 0
    torch.cuda.synchronize()
    return z.cpu().numpy()


# (nx, ny, steps) pairs: each case calls twice, the second grid smaller
CASES = [((1024, 1024, 7), (700, 333, 4)),     # one step per launch, odd then even
         ((300, 200, 9), (97, 61, 3)),         # LDS-blocked path (< 2^18 cells)
         ((4096, 4096, 3), (2048, 1000, 5))]   # the largest grid the array holds


@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_equals_pool_form_two_calls(hm, orc, first, second):
    for (nx, ny, it), seed in ((first, 31), (second, 32)):
        z0 = orc.fbm_terrain(nx, ny, seed)
        want = _run(hm, "hmg_thermal_dev", z0, 1.0 / nx, it)
        got = _run(hm, "hmg_thermal_dev_noalloc", z0, 1.0 / nx, it)
        assert np.array_equal(got, want), (nx, ny, it)


def test_noalloc_talus_plane(hm, orc):
    nx, ny = 257, 130
    z0 = orc.fbm_terrain(nx, ny, 33)
    talus = (np.random.default_rng(3).random((ny, nx), dtype=np.float32) * 2.0 / nx)
    for it in (5, 2):
        want = _run(hm, "hmg_thermal_dev", z0, 0.0, it, talus)
        got = _run(hm, "hmg_thermal_dev_noalloc", z0, 0.0, it, talus)
        assert np.array_equal(got, want)
    assert np.array_equal(got, orc.thermal(z0, talus, 2))


def test_noalloc_rejects_grids_above_its_array(hm):
    from highmap_amd import lib

    import torch

    ctx = hm._default_ctx
    # checked on the host before any device work; the buffer has the size
    # the call is given, so nothing could run out of bounds either way
    z = torch.zeros((4096, 4097), dtype=torch.float32, device="cuda")
    rc = lib().hmg_thermal_dev_noalloc(ctx.handle, C.c_void_p(z.data_ptr()), 4097, 4096, None,
                                       0.01, 3, None)
    assert rc == INVALID
    assert b"scratch" in lib().hmg_last_error()
