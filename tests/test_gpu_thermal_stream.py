"""The streaming Jacobi steps (uniform talus; one step per launch, or two
per launch with the second step reading the first from registers; the
thermal() call's extrapolate_borders folded into a last single step)
against the FP32 oracle, bit for bit (compared as bit patterns, so the sign
of a zero counts).  With block steps 0 or 2 every size takes the streaming
kernels;
the shapes cover one strip (nx <= 62), the strip edge (62, 63, 64, 124,
125, 126), the moved last strip, grids of 3 to 6 rows, the last chunk
shorter than 3 rows, zero talus, a talus no slope exceeds and signed
zeros.  Talus planes and negative scalars take the streaming kernels with
each cell's own talus (all eight exchanges per cell; one or two steps per
launch as above) at every size from 4 x 4: the same shapes, planes with zeros, tiny, negative and huge values,
NaNs and infinities in the heights, and signed zeros."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 2], ids=["one_step", "two_steps"])
def streaming(hm, request):
    """0: one step per launch at every size; 2: two steps per launch (the
    last odd step alone, the borders after an even count by their own
    launches)."""
    L = hm.lib()
    assert L.hmg_thermal_set_block_steps(request.param) == 0
    yield
    assert L.hmg_thermal_set_block_steps(8) == 0


def _bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                          np.ascontiguousarray(b).view(np.uint32))


def _input(orc, nx, ny, seed):
    if nx >= 8 and ny >= 8:
        return orc.fbm_terrain(nx, ny, seed)
    return np.random.default_rng(seed).random((ny, nx), dtype=np.float32)


SHAPES = [(3, 3), (4, 4), (5, 7), (7, 4), (62, 9), (63, 10), (64, 5), (65, 6), (124, 11),
          (125, 12), (126, 13), (130, 67), (200, 131), (257, 3), (3, 257), (1000, 301)]


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("iters", [1, 2, 4, 5])
def test_stream_steps_match_oracle(hm, orc, streaming, nx, ny, iters):
    z0 = _input(orc, nx, ny, nx * 1000 + ny)
    for talus in (0.5 / nx, 4.0 / nx):
        z = z0.copy()
        hm.thermal(z, talus, iters)
        assert _bits_equal(z, orc.thermal(z0, talus, iters)), talus


@pytest.mark.parametrize("talus", [0.0, 1e30])
def test_stream_zero_and_huge_talus(hm, orc, streaming, talus):
    z0 = _input(orc, 131, 77, 5)
    z = z0.copy()
    hm.thermal(z, talus, 3)
    assert _bits_equal(z, orc.thermal(z0, talus, 3))


def test_stream_signed_zeros(hm, orc, streaming):
    """A field of -0.0 with a spike: where every exchange is zero the
    reference adds +0 to -0 and gets +0."""
    z0 = np.full((40, 70), -0.0, np.float32)
    z0[20, 30] = 1.0
    z0[5, 5:9] = 0.0
    for it in (1, 2):
        z = z0.copy()
        hm.thermal(z, 0.01, it)
        assert _bits_equal(z, orc.thermal(z0, 0.01, it))


def test_stream_default_path_large_grid(hm, orc):
    """Default kernel choice: a grid of 2^18 cells or more streams."""
    z0 = orc.fbm_terrain(1030, 777, 8)
    for it in (1, 4):
        z = z0.copy()
        hm.thermal(z, 1.0 / 1030, it)
        assert _bits_equal(z, orc.thermal(z0, 1.0 / 1030, it))


def test_stream_in_session_matches_oracle(hm, orc, streaming):
    from conftest import VP_WORKLOAD

    nx, ny = 190, 97
    z0 = orc.fbm_terrain(nx, ny, 2)
    s = hm.Session(z0, **VP_WORKLOAD)
    o = orc.VpState(z0, orc.vp_params(**VP_WORKLOAD))
    for k in (3, 10, 4):
        s.vpipes(k).thermal(1.0 / nx, 2)
        o.run(k).thermal_step(1.0 / nx, 2)
    g = s.export()
    s.close()
    for a, b in zip(g, (o.z, o.d, o.s, o.u, o.v)):
        assert _bits_equal(a, b)


@pytest.mark.parametrize("iters", [2, 4, 7])
def test_default_choice_above_2_18_cells(hm, orc, iters):
    """The default kernel choice at 1500 x 1100: two steps per launch, then
    a last odd step with the borders or, after an even count, the border
    launches."""
    z0 = orc.fbm_terrain(1500, 1100, 11)
    z = z0.copy()
    hm.thermal(z, 2.0 / 1500, iters)
    assert _bits_equal(z, orc.thermal(z0, 2.0 / 1500, iters))


def _plane(nx, ny, seed, scale):
    r = np.random.default_rng(seed)
    t = (r.random((ny, nx), dtype=np.float32) * np.float32(scale)).astype(np.float32)
    t[r.random((ny, nx)) < 0.1] = 0.0
    t[r.random((ny, nx)) < 0.05] = -np.float32(scale)
    t[r.random((ny, nx)) < 0.02] = np.float32(1e-35)
    return t


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("iters", [1, 2, 5])
def test_plane_stream_steps_match_oracle(hm, orc, streaming, nx, ny, iters):
    z0 = _input(orc, nx, ny, nx * 1000 + ny + 7)
    t = _plane(nx, ny, nx + ny, 3.0 / nx)
    z = z0.copy()
    hm.thermal(z, t, iters)
    assert _bits_equal(z, orc.thermal(z0, t, iters))
    for neg in (-1e-3, -2.0 / nx):
        z = z0.copy()
        hm.thermal(z, neg, iters)
        assert _bits_equal(z, orc.thermal(z0, neg, iters)), neg


def test_plane_stream_special_values(hm, orc, streaming):
    """Heights with NaN, +-inf, huge opposite values (differences that
    overflow), tiny differences around the talus (quotients below 2^-100)
    and -0.0; a plane holding NaN and +-inf talus.  The call delivers the
    result and reports it as non-finite (-4, HMG_ERR_NONFINITE)."""
    nx, ny = 131, 45
    z0 = orc.fbm_terrain(nx, ny, 11)
    z0[10, 20] = np.nan
    z0[30, 70] = np.inf
    z0[31, 71] = -np.inf
    z0[5, 100] = 3e38
    z0[5, 101] = -3e38
    z0[40, 5:60] = -0.0
    z0[20, 40:44] = np.float32(1e-31) * np.arange(4, dtype=np.float32)
    t = _plane(nx, ny, 3, 2.0 / nx)
    t[20, 40:44] = np.float32(1e-31)
    t[12, 12] = np.nan
    t[13, 13] = np.inf
    t[14, 14] = -np.inf
    for it in (1, 3):
        z = z0.copy()
        with pytest.raises(hm.HmgError) as e:
            hm.thermal(z, t, it)
        assert e.value.code == -4
        want = orc.thermal(z0, t, it)
        same = (z.view(np.uint32) == want.view(np.uint32)) | (np.isnan(z) & np.isnan(want))
        assert same.all(), (it, np.argwhere(~same)[:5])


@pytest.mark.parametrize("nx,ny", SHAPES)
@pytest.mark.parametrize("iters", [1, 2, 5])
def test_auto_bedrock_stream_matches_oracle(hm, orc, nx, ny, iters):
    """thermal_auto_bedrock's streaming step (from 4 x 4; the bedrock plane
    updated in place by the owner of each column, the moved last strip
    writing only the columns past the previous strip's) with a talus plane
    and with a scalar."""
    z0 = _input(orc, nx, ny, nx * 7 + ny)
    t = _plane(nx, ny, nx * ny, 3.0 / nx)
    for talus in (t, 1.5 / nx, -1e-3):
        z = z0.copy()
        hm.thermal_auto_bedrock(z, talus, iters)
        want = orc.thermal_auto_bedrock(z0, talus, iters)
        assert _bits_equal(z, want), (np.isscalar(talus), int((z != want).sum()))


def test_plane_and_auto_bedrock_default_path_large_grid(hm, orc):
    """Default kernel choice at 2^20 cells: a talus plane (two steps per
    launch, an odd last step alone) and thermal_auto_bedrock's stream."""
    nx, ny = 1030, 777
    z0 = orc.fbm_terrain(nx, ny, 12)
    t = _plane(nx, ny, 5, 2.0 / nx)
    for it in (1, 4, 5):
        z = z0.copy()
        hm.thermal(z, t, it)
        assert _bits_equal(z, orc.thermal(z0, t, it)), it
    for it in (1, 3):
        z = z0.copy()
        hm.thermal_auto_bedrock(z, t, it)
        assert _bits_equal(z, orc.thermal_auto_bedrock(z0, t, it)), it
