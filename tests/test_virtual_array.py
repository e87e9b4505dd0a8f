"""hmap::VirtualArray of the C++ drop-in (include/hmap_virtual_array.hpp)
against the restatement in oracle/virtual_array.py, bit for bit, and pins
of the restatement: tiles partition the grid, halos only towards
neighbours, bounding boxes span their cells, same-shape round trips are the
identity, std::lerp as the C++ library computes it, overlap smoothing
keeps agreeing overlaps unchanged, resampling reproduces constant, linear
and (bicubic) quadratic fields at the reference's sample positions, and
the downscaled single-array mode returns the array unchanged when the
operation is the identity.  The tile storages (RAM, disk-backed LRU,
disk-sequential) give the same grid as RAM, and their eviction, reload,
trim and one-tile rules hold.  No GPU call."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import virtual_array as V

F = np.float32
DRIVER = os.path.join(ROOT, "tests", "cpp", "va_driver")

GEOS = [((1024, 1024), (0.0, 1.0, 0.0, 1.0), (220, 312), 32),
        ((1024, 1024), (0.0, 1.0, 0.0, 1.0), (8, 8), 7),
        ((130, 97), (-1.5, 2.25, 0.1, 0.9), (40, 33), 5),
        ((64, 48), (0.0, 1.0, 0.0, 1.0), (64, 48), 3),
        ((50, 70), (0.0, 3.0, -1.0, 1.0), (17, 70), 4)]


def pattern(nx, ny):
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    return (((7 * i + 13 * j) % 101).astype(np.float32) / F(101)).astype(np.float32)


def hexf(s):
    return F(float.fromhex(s))


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("va")
    assert os.path.exists(DRIVER), "tests/cpp/va_driver not built (run the build first)"
    r = subprocess.run([DRIVER, "geom", str(d)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(os.path.join(d, "geom.txt")).read().split("\n")

    def load(name, shape=None):
        a = np.fromfile(os.path.join(d, name + ".f32"), dtype=np.float32)
        return a if shape is None else a.reshape(shape)

    return lines, load


# -------------------------------------------------- pins of the restatement

@pytest.mark.parametrize("geo", GEOS)
def test_tiles_partition_the_grid(geo):
    shape, bbox, tile, halo = geo
    va = V.VA(shape, bbox, tile, halo)
    cover = np.zeros((shape[1], shape[0]), np.int32)
    ntx, nty = va.max_tiles()
    for r in va.regions():
        tx, ty = r.key
        hl, hr, hb, ht = r.halo
        assert (hl == 0) == (tx == 0) and (hr == 0) == (tx == ntx - 1)
        assert (hb == 0) == (ty == 0) and (ht == 0) == (ty == nty - 1)
        assert {hl, hr, hb, ht} <= {0, halo}
        i0, j0 = va.origin(r)
        w, h = r.shape[0] - hl - hr, r.shape[1] - hb - ht
        assert 0 < w <= tile[0] and 0 < h <= tile[1]
        cover[j0 + hb:j0 + hb + h, i0 + hl:i0 + hl + w] += 1
        # the tile (halo included) lies inside the grid
        assert i0 >= 0 and j0 >= 0 and i0 + r.shape[0] <= shape[0] and j0 + r.shape[1] <= shape[1]
        # its bounding box spans its cells
        dx = (bbox[1] - bbox[0]) / shape[0]
        dy = (bbox[3] - bbox[2]) / shape[1]
        assert r.bbox[0] == pytest.approx(bbox[0] + i0 * dx, abs=1e-5)
        assert r.bbox[1] == pytest.approx(bbox[0] + (i0 + r.shape[0]) * dx, abs=1e-5)
        assert r.bbox[2] == pytest.approx(bbox[2] + j0 * dy, abs=1e-5)
        assert r.bbox[3] == pytest.approx(bbox[2] + (j0 + r.shape[1]) * dy, abs=1e-5)
    assert np.all(cover == 1)


@pytest.mark.parametrize("geo", GEOS[2:])
def test_same_shape_round_trip_is_identity(geo):
    shape, bbox, tile, halo = geo
    va = V.VA(shape, bbox, tile, halo)
    a = pattern(*shape)
    va.from_array(a)
    assert np.array_equal(va.to_array(), a)
    for j in (0, shape[1] // 2, shape[1] - 1):
        for i in (0, shape[0] // 3, shape[0] - 1):
            assert va.get(i, j) == a[j, i]
    # halo cells hold the grid's values too
    for r in va.regions():
        i0, j0 = va.origin(r)
        assert np.array_equal(va.tile(r), a[j0:j0 + r.shape[1], i0:i0 + r.shape[0]])


def test_smoothing_keeps_agreeing_overlaps():
    va = V.VA((130, 97), (0, 1, 0, 1), (40, 33), 5)
    a = pattern(130, 97)
    va.from_array(a)
    va.smooth_overlap_buffers()
    assert np.array_equal(va.to_array(), a)


def test_percentiles_and_unique_values_of_a_known_distribution():
    """The pattern takes the 101 values k/101 nearly equally often: its
    p-percentile lies within a bin and a value of p, and its distinct values
    are exactly those 101."""
    va = V.VA((130, 97), (0, 1, 0, 1), (40, 33), 5)
    a = pattern(130, 97)
    va.from_array(a)
    for p_lo, p_hi in ((0.05, 0.95), (0.25, 0.5)):
        lo, hi = va.range_percentile(p_lo, p_hi)
        assert abs(lo - p_lo * 100 / 101) <= 2 / 101 + 1 / 1024
        assert abs(hi - p_hi * 100 / 101) <= 2 / 101 + 1 / 1024
    assert np.array_equal(va.unique_values(), np.unique(a))
    assert va.unique_values().size == 101


def test_lerp_and_smoothstep5_closed_forms():
    assert V.smoothstep5(0.0) == 0 and V.smoothstep5(1.0) == 1 and V.smoothstep5(0.5) == 0.5
    assert V.lerp(2.0, 6.0, 0.0) == 2.0 and V.lerp(2.0, 6.0, 1.0) == 6.0
    assert V.lerp(2.0, 6.0, 0.25) == 3.0
    assert V.lerp(-2.0, 6.0, 0.5) == 2.0


# ------------------------------------------------ the C++ VirtualArray


def test_lerp_matches_the_cpp_library(run):
    lines, _ = run
    rows = [ln.split() for ln in lines if ln.startswith("lerp ")]
    assert len(rows) == 9
    for _, a, b, t, want in rows:
        assert V.lerp(hexf(a), hexf(b), hexf(t)) == hexf(want), (a, b, t)


@pytest.mark.parametrize("k", range(len(GEOS)))
def test_cpp_virtual_array_matches_restatement(run, k):
    lines, load = run
    shape, bbox, tile, halo = GEOS[k]
    nx, ny = shape
    va = V.VA(shape, bbox, tile, halo)
    # tile regions, cell centres and corners
    head = [ln.split() for ln in lines if ln.startswith("va %d " % k)][0]
    assert (int(head[2]), int(head[3])) == va.max_tiles()
    start = lines.index(" ".join(head))
    rows = [ln.split() for ln in lines[start + 1:start + 1 + int(head[2]) * int(head[3])]]
    for row in rows:
        tx, ty = int(row[1]), int(row[2])
        r = va.region(tx, ty)
        assert tuple(hexf(x) for x in row[3:7]) == r.bbox
        assert (int(row[7]), int(row[8])) == r.shape
        assert tuple(int(x) for x in row[9:13]) == r.halo
        b0, b1, b2, b3 = r.bbox
        dxc = F(F(b1 - b0) / F(r.shape[0]))
        dyc = F(F(b3 - b2) / F(r.shape[1]))
        assert hexf(row[13]) == F(b0 + F(F(F(1) + F(0.5)) * dxc))
        assert hexf(row[14]) == F(b2 + F(F(F(2) + F(0.5)) * dyc))
        tcx, tcy = F(F(3) / F(r.shape[0])), F(F(1) / F(r.shape[1]))
        assert hexf(row[15]) == F(F(b0 * F(F(1) - tcx)) + F(b1 * tcx))
        assert hexf(row[16]) == F(F(b2 * F(F(1) - tcy)) + F(b3 * tcy))
    # from_array / to_array
    src = pattern(nx, ny)
    va.from_array(src)
    assert np.array_equal(load("same_%d" % k, (ny, nx)), va.to_array())
    assert np.array_equal(load("half_%d" % k, (ny // 2, nx // 2)), va.to_array((nx // 2, ny // 2)))
    vb = V.VA(shape, bbox, tile, halo)
    vb.from_array(pattern(nx // 3 + 2, ny // 2 + 3))
    assert np.array_equal(load("from_small_%d" % k, (ny, nx)), vb.to_array())
    # reductions
    red = [ln.split() for ln in lines if ln.startswith("red %d " % k)][0]
    assert hexf(red[2]) == va.sum()
    assert hexf(red[3]) == va.min() and hexf(red[4]) == va.max()
    assert hexf(red[5]) == va.mean()
    assert hexf(red[6]) == vb.sum()
    pct = [ln.split() for ln in lines if ln.startswith("pct %d " % k)][0]
    assert (hexf(pct[2]), hexf(pct[3])) == va.range_percentile(0.05, 0.95)
    assert (hexf(pct[4]), hexf(pct[5])) == va.range_percentile(0.2, 0.7, 64)
    uv = va.unique_values()
    assert int(pct[6]) == uv.size and hexf(pct[7]) == uv[0] and hexf(pct[8]) == uv[-1]
    # get / set / interpolation
    va.set(nx - 1, 0, 5.0)
    got = [ln.split() for ln in lines if ln.startswith("get %d " % k)][0]
    assert hexf(got[2]) == va.get(nx // 2, ny // 3)
    assert hexf(got[3]) == F(5.0)
    b = va.bbox
    xq = F(b[0] + F(F(0.3) * F(b[1] - b[0])))
    assert hexf(got[4]) == va.get_nearest(xq, b[3])
    xb = F(b[0] + F(F(0.37) * F(b[1] - b[0])))
    yb = F(b[2] + F(F(0.61) * F(b[3] - b[2])))
    assert hexf(got[5]) == va.get_bilinear(xb, yb)
    # overlap smoothing of disagreeing tiles, then remap
    vs = V.VA(shape, bbox, tile, halo)
    for r in vs.regions():
        ii = np.arange(r.shape[0])[None, :]
        jj = np.arange(r.shape[1])[:, None]
        base = F(r.key[0] * 3 + r.key[1] * 5)
        vs.tile(r)[...] = base + ((ii + 2 * jj) % 7).astype(np.float32) / F(7)
    vs.smooth_overlap_buffers()
    tiles = np.concatenate([vs.tile(r).ravel() for r in vs.regions()])
    assert np.array_equal(load("smooth_tiles_%d" % k), tiles)
    vs.remap(0.0, 1.0)
    assert np.array_equal(load("smooth_remap_%d" % k, (ny, nx)), vs.to_array())


# ------------------------------------------------ resampling and the modes


@pytest.fixture(scope="module")
def modes(tmp_path_factory):
    d = tmp_path_factory.mktemp("vam")
    r = subprocess.run([DRIVER, "modes", str(d)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr

    def load(name, shape):
        return np.fromfile(os.path.join(d, name + ".f32"), dtype=np.float32).reshape(
            shape[1], shape[0])

    return load


def _exact_axis(nt, ns):
    """The reference grid's source coordinates in float64."""
    s = 0.5 / nt
    return (s + np.arange(nt) * (1 - s) / nt) * ns - 0.5


@pytest.mark.parametrize("nt,ns", [(7, 20), (40, 13), (64, 64), (1, 5), (333, 77)])
def test_sample_positions(nt, ns):
    """Pixel-centred start 0.5/nt, end point excluded: the same shape does
    not sample the cell centres (the last cell sits at n - 1.5 + 0.5/n)."""
    c = V.sample_axis(nt, ns)
    assert np.allclose(c, _exact_axis(nt, ns), atol=2e-5 * ns)
    if nt == ns:
        assert c[0] == 0 and abs(c[-1] - (ns - 1.5 + 0.5 / ns)) < 1e-4


def test_resampling_reproduces_polynomials():
    """Constants exactly; a linear field (different slopes along i and j)
    bilinearly at every point inside the source; a quadratic field
    bicubically where the 4x4 stencil stays inside (Catmull-Rom has
    quadratic precision)."""
    ns = (37, 23)
    i = np.arange(ns[0], dtype=np.float64)[None, :]
    j = np.arange(ns[1], dtype=np.float64)[:, None]
    c = np.full((ns[1], ns[0]), 0.3, np.float32)
    for shape in ((50, 31), (12, 9), (37, 23)):
        assert np.all(V.resample_bilinear(c, shape) == F(0.3))
        assert np.all(V.resample_bicubic(c, shape) == F(0.3))
        cx, cy = _exact_axis(shape[0], ns[0]), _exact_axis(shape[1], ns[1])
        lin = (0.25 * i + 1.5 * j).astype(np.float32)
        got = V.resample_bilinear(lin, shape)
        inx = (cx >= 0) & (cx <= ns[0] - 1)
        iny = (cy >= 0) & (cy <= ns[1] - 1)
        want = 0.25 * cx[None, :] + 1.5 * cy[:, None]
        assert np.allclose(got[np.ix_(iny, inx)], want[np.ix_(iny, inx)], atol=2e-4)
        quad = (0.02 * i * i + 0.5 * j).astype(np.float32)
        got = V.resample_bicubic(quad, shape)
        okx = (cx >= 1) & (np.floor(cx) <= ns[0] - 3)
        oky = (cy >= 1) & (np.floor(cy) <= ns[1] - 3)
        want = 0.02 * cx[None, :] ** 2 + 0.5 * cy[:, None]
        assert np.allclose(got[np.ix_(oky, okx)], want[np.ix_(oky, okx)], atol=2e-4)


def test_bicubic_interpolates_at_the_nodes():
    a = pattern(19, 11)
    i = np.arange(19)[None, :]
    j = np.arange(11)[:, None]
    z = np.zeros((1, 1), np.float32)
    assert np.array_equal(V.bicubic_at(a, i, j, z, z), a)
    # and a cubic in x is not reproduced (the pin tells the orders apart)
    x = np.float32(0.25)
    assert V.catmull_rom(-1.0, 0.0, 1.0, 8.0, x) == F(0.109375)  # x - 3x^2 + 3x^3, not x^3
    assert V.catmull_rom(1.0, 0.0, 1.0, 4.0, x) == F(0.0625)     # x^2 exactly


def test_downscaled_identity_keeps_the_array():
    """With an operation that changes nothing, coarse + (full - low) is the
    full array up to the rounding of one subtraction and one addition."""
    va = V.VA((130, 97), (0, 1, 0, 1), (40, 33), 5)
    a = pattern(130, 97)
    va.from_array(a)
    va.single_array(lambda x: x, k_cutoff=0.5)
    assert np.allclose(va.to_array(), a, atol=4e-7)


@pytest.mark.parametrize("k", [0, 2, 3, 4])  # 1 is 0's grid in 8 x 8 tiles
def test_cpp_resampling_and_modes_match_restatement(modes, k):
    shape, bbox, tile, halo = GEOS[k]
    nx, ny = shape
    src = pattern(nx, ny)
    down = (nx // 3 + 1, ny // 2 + 5)
    up = (2 * nx - 3, ny + 7)
    assert np.array_equal(modes("rs_lin_down_%d" % k, down), V.resample_bilinear(src, down))
    assert np.array_equal(modes("rs_lin_up_%d" % k, up), V.resample_bilinear(src, up))
    assert np.array_equal(modes("rs_lin_same_%d" % k, shape), V.resample_bilinear(src, shape))
    assert np.array_equal(modes("rs_cub_down_%d" % k, down), V.resample_bicubic(src, down))
    assert np.array_equal(modes("rs_cub_up_%d" % k, up), V.resample_bicubic(src, up))
    small = pattern(nx // 3 + 2, ny // 2 + 3)
    va = V.VA(shape, bbox, tile, halo)
    va.from_array_bicubic(small)
    assert np.array_equal(modes("from_cub_%d" % k, shape), va.to_array())

    def add_index(a):
        h, w = a.shape
        return a + (np.arange(w)[None, :] * h + np.arange(h)[:, None]).astype(np.float32)

    for name, kw, f in (("single", {}, add_index), ("single_s3", {"stride": 3}, add_index),
                        ("down_05", {"k_cutoff": 0.5}, lambda a: a * F(0.5)),
                        ("down_1", {"k_cutoff": 1.0}, lambda a: a * F(0.5))):
        vc = V.VA(shape, bbox, tile, halo)
        vc.from_array(src)
        vc.single_array(f, **kw)
        assert np.array_equal(modes("%s_%d" % (name, k), shape), vc.to_array()), name


def test_cpp_from_array_bilinear(modes):
    """from_array_bilinear: corner clamped to n - 2, so the 2x2 cells exist;
    on a linear source it lands on the scaled cell centre."""
    nx, ny = 130, 97
    va_lin = modes("from_lin_2", (nx, ny))
    sx, sy = 130 // 3 + 2, 97 // 2 + 3
    assert np.all(np.isfinite(va_lin))
    # the pattern is piecewise linear with period 101 (7 i + 13 j mod 101):
    # check the mean against the bicubic form instead of a closed form
    assert abs(float(va_lin.mean()) - float(modes("from_cub_2", (nx, ny)).mean())) < 0.02
    assert va_lin.shape == (ny, nx) and (sx, sy) == (45, 51)


# ------------------------------------------------------------ tile storages


@pytest.fixture(scope="module")
def storage(tmp_path_factory):
    d = tmp_path_factory.mktemp("vas")
    r = subprocess.run([DRIVER, "storage", str(d)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(os.path.join(d, "storage.txt")).read().split("\n")

    def load(name):
        return np.fromfile(os.path.join(d, name + ".f32"), dtype=np.float32).reshape(97, 130)

    return lines, load


def _storage_sequence(smooth):
    va = V.VA((130, 97), (-1.5, 2.25, 0.1, 0.9), (40, 33), 5)
    va.from_array(pattern(130, 97))
    for r in va.regions():
        ii = np.arange(r.shape[0])[None, :]
        jj = np.arange(r.shape[1])[:, None]
        t = va.tile(r)
        t[...] = (t + F(r.key[0] * 3 + r.key[1] * 5)) + ((ii + 2 * jj) % 7).astype(np.float32) / F(7)
    if smooth:
        va.smooth_overlap_buffers()
    va.remap(0.0, 1.0)
    va.set(7, 9, 2.5)
    a = va.to_array()
    for r in va.regions():
        va.tile(r)[...] = F(2) * va.tile(r) + F(1)
    return a, va.to_array(), va.sum()


MODES = ("VA_RAM", "VA_DISK_LRU", "VA_DISK_LRU_MIN", "VA_DISK_SEQUENTIAL")


@pytest.mark.parametrize("mode", MODES)
def test_storages_give_the_grid_of_the_restatement(storage, mode):
    """Every store gives the RAM grid; the sequential store holds one tile,
    so overlap smoothing is skipped there (as the reference's does)."""
    lines, load = storage
    a, b, bsum = _storage_sequence(smooth=mode != "VA_DISK_SEQUENTIAL")
    assert np.array_equal(load("st_a_" + mode), a)
    assert np.array_equal(load("st_b_" + mode), b)
    row = [ln for ln in lines if ln.startswith("mode %s " % mode)][0]
    head, info = row.split("|")
    _, _, live, got, s = head.split()
    assert hexf(got) == F(2.5) and hexf(s) == bsum
    want_live = {"VA_RAM": 2 ** 64 - 1, "VA_DISK_LRU": 12, "VA_DISK_LRU_MIN": 2,
                 "VA_DISK_SEQUENTIAL": 1}[mode]
    assert int(live) == want_live
    assert info == {"VA_RAM": "RAM", "VA_DISK_LRU": "DiskLRU (capacity=12)",
                    "VA_DISK_LRU_MIN": "DiskLRU (capacity=2)",
                    "VA_DISK_SEQUENTIAL": "DiskSequential"}[mode]


def test_storage_rules(storage):
    lines, _ = storage
    rows = {ln.split()[0] + (" " + ln.split()[1] if ln.startswith("trim") else ""): ln.split()
            for ln in lines if ln}
    # trim writes every tile of the 4 x 3 grid out and frees them
    for m in ("VA_DISK_LRU", "VA_DISK_LRU_MIN"):
        assert rows["trim " + m][2:] == ["0", "12"]
    # tiles are 45 x 38 (0,0) and 50 x 38 ((1,0), (2,0)) cells with halos
    n00, n10, n20 = 45 * 38, 50 * 38, 50 * 38
    # RAM LRU of two: the tile that left comes back zero-filled
    assert rows["lru"][1:] == ["2", float(F(0)).hex().replace("0x0.0p+0", "0x0p+0"),
                               rows["lru"][3]] and hexf(rows["lru"][3]) == F(3 * n20)
    # disk LRU of two: the tile that left is on disk and comes back intact
    r = rows["disklru"]
    assert r[1:3] == ["2", "1"]
    assert (hexf(r[3]), hexf(r[4]), hexf(r[5])) == (F(n00), F(2 * n10), F(3 * n20))
    # three tiles held at once in a store of two; room made down to two
    r = rows["held"]
    assert r[1:3] == ["3", "2"] and [hexf(x) for x in r[3:]] == [F(4), F(5), F(6)]
    assert rows["removed"][1] == "1"
    # one tile at a time: a second get before the release is refused
    r = rows["seq"]
    assert r[1:3] == ["1", "2"] and hexf(r[3]) == F(2 * 50 * 43) and hexf(r[4]) == F(9)
