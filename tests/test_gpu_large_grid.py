"""Grids above 2^28 cells.  The streaming kernels address their planes with
32-bit byte offsets (a 16-byte flux quadruple per cell), so a session that
large runs the plain passes whatever mode it asks for, and the batch, the
server and flow_simulation reject such grids before touching any memory.
The session is checked against the FP32 oracle (2 iterations and a thermal
step on 16385 x 16385: 1 GiB per plane, about 20 GiB of device memory)."""
import ctypes as C

import numpy as np
import pytest

from conftest import VP_WORKLOAD

pytestmark = pytest.mark.gpu

INVALID = -1
N = 16385  # 16385^2 = 2^28 + 32769 cells


def test_batch_server_and_flow_reject_grids_above_2_28(hm):
    import torch

    from highmap_amd import lib, serve

    L = lib()
    ctx = hm._default_ctx
    # the size check comes first; the buffers are real and large enough for
    # what the calls are told, so nothing could read out of bounds anyway
    small = np.zeros(16, np.float32)
    b = hm.Batch(**VP_WORKLOAD)
    mid = C.c_int()
    rc = L.hmg_batch_add(b.handle, small.ctypes.data, N, N, None, 3, C.byref(mid))
    assert rc == INVALID and b"2^28" in L.hmg_last_error()
    b.close()
    sv = serve.Server(storage=hm.Storage.BF16, params=hm.VpipesParams.make(**VP_WORKLOAD))
    with pytest.raises(hm.HmgError) as e:
        sv.submit(N, N, serve.JOB_THERMAL, 3, 0.01, seed=1)
    assert e.value.code == INVALID
    sv.close()
    zd = torch.zeros(N * N, dtype=torch.float32, device="cuda")
    rc = L.hmg_flow_simulation_dev(ctx.handle, C.c_void_p(zd.data_ptr()), N, N, 0.01,
                                   C.c_void_p(zd.data_ptr()), 2, 0.5, 1, 0.01, 0.0,
                                   C.c_void_p(zd.data_ptr()))
    assert rc == INVALID and b"2^28" in L.hmg_last_error()


def test_session_above_2_28_cells_matches_oracle(hm, orc):
    z0 = np.random.default_rng(7).random((N, N), dtype=np.float32)
    s = hm.Session(z0, mode="fused", **VP_WORKLOAD)
    assert not s.tuning()[2]  # no online tuning: the plain passes run
    s.vpipes(2).thermal(1.0 / N, 1)
    g = s.export()
    s.close()
    o = orc.VpState(z0, orc.vp_params(**VP_WORKLOAD))
    o.run(2).thermal_step(1.0 / N, 1)
    for name, a, b in zip("zdsuv", g, (o.z, o.d, o.s, o.u, o.v)):
        assert np.array_equal(a, b), name
