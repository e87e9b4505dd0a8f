"""hmap::VirtualArray of the C++ drop-in (include/hmap_virtual_array.hpp)
against the restatement in oracle/virtual_array.py, bit for bit, and pins
of the restatement: tiles partition the grid, halos only towards
neighbours, bounding boxes span their cells, same-shape round trips are the
identity, std::lerp as the C++ library computes it, and overlap smoothing
keeps agreeing overlaps unchanged.  No GPU call."""
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
    # the single-array mode: the whole array at once, then back to tiles
    ii = np.arange(nx)[None, :]
    jj = np.arange(ny)[:, None]
    want = src + (ii * ny + jj).astype(np.float32)
    assert np.array_equal(load("single_%d" % k, (ny, nx)), want)
