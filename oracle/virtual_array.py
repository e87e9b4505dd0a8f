"""Restatement of hmap::VirtualArray's tile geometry and whole-array
operations (TEST INFRASTRUCTURE ONLY), in float32 scalar arithmetic
operation by operation, for checking the drop-in's C++ VirtualArray.

Reference: HighMap/src/virtual_array/virtual_array.cpp,
virtual_array_processing.cpp, HighMap/include/highmap/virtual_array/
virtual_array.inl, tile_region.hpp; the resampling of
HighMap/src/array/methods.cpp:203-245,364-383 and
HighMap/src/interpolate/interpolate_array.cpp:15-185 with
HighMap/include/highmap/interpolate/interpolate2d.hpp:94-113; std::lerp as
libstdc++ defines it (<cmath>, C++20).  Arrays are numpy (ny, nx) float32
planes, so element (i, j) of hmap::Array is a[j, i].
"""
import numpy as np

F = np.float32


def ceil_div(a, b):
    return (a + b - 1) // b


class Region:
    def __init__(self, tx, ty, bbox, shape, halo):
        self.key = (tx, ty)
        self.bbox = bbox      # (x1, x2, y1, y2) float32
        self.shape = shape    # (nx, ny), halo included
        self.halo = halo      # (left, right, bottom, top)


class VA:
    """Tiles held as a dict of float32 (ny, nx) planes, halo included."""

    def __init__(self, shape, bbox, tile_shape, halo):
        self.shape = tuple(shape)
        self.bbox = tuple(F(b) for b in bbox)
        self.tile_shape = tuple(tile_shape)
        self.halo = int(halo)
        self.tiles = {}

    def max_tiles(self):
        return (ceil_div(self.shape[0], self.tile_shape[0]),
                ceil_div(self.shape[1], self.tile_shape[1]))

    def region(self, tx, ty):
        """virtual_array.cpp:429-468."""
        sx, sy = self.shape
        tsx, tsy = self.tile_shape
        w = min(tsx, sx - tx * tsx)
        h = min(tsy, sy - ty * tsy)
        txm, tym = ceil_div(sx, tsx) - 1, ceil_div(sy, tsy) - 1
        hl = 0 if tx == 0 else self.halo
        hr = 0 if tx == txm else self.halo
        hb = 0 if ty == 0 else self.halo
        ht = 0 if ty == tym else self.halo
        b0, b1, b2, b3 = self.bbox
        dx, dy = F(b1 - b0), F(b3 - b2)
        x1 = F(b0 + F(F(dx * F(tx * tsx - hl)) / F(sx)))
        x2 = F(b0 + F(F(dx * F(tx * tsx + w + hr)) / F(sx)))
        y1 = F(b2 + F(F(dy * F(ty * tsy - hb)) / F(sy)))
        y2 = F(b2 + F(F(dy * F(ty * tsy + h + ht)) / F(sy)))
        return Region(tx, ty, (x1, x2, y1, y2), (w + hl + hr, h + hb + ht), (hl, hr, hb, ht))

    def regions(self):
        ntx, nty = self.max_tiles()
        return [self.region(tx, ty) for ty in range(nty) for tx in range(ntx)]

    def origin(self, r):
        """tile_region_global_indices: global (i, j) of the tile's cell 0."""
        return (r.key[0] * self.tile_shape[0] - r.halo[0], r.key[1] * self.tile_shape[1] - r.halo[2])

    def tile(self, r):
        if r.key not in self.tiles:
            self.tiles[r.key] = np.zeros((r.shape[1], r.shape[0]), np.float32)
        return self.tiles[r.key]

    def _index(self, r, nx_a, ny_a, interior):
        """Destination/source cell indices of a tile's cells (the float32
        products of virtual_array.cpp:104-130 and 470-501, vectorised)."""
        rx = F(F(nx_a - 1) / F(self.shape[0] - 1))
        ry = F(F(ny_a - 1) / F(self.shape[1] - 1))
        i0, j0 = self.origin(r)
        hl, hr, hb, ht = r.halo if interior else (0, 0, 0, 0)
        ii = np.arange(hl, r.shape[0] - hr)
        jj = np.arange(hb, r.shape[1] - ht)
        xs = (i0 + ii).astype(np.float32) + F(0.5)
        ys = (j0 + jj).astype(np.float32) + F(0.5)
        ig = np.clip(np.floor(xs * rx).astype(np.int64), 0, nx_a - 1)
        jg = np.clip(np.floor(ys * ry).astype(np.int64), 0, ny_a - 1)
        return ii, jj, ig, jg

    def from_array_bicubic(self, a):
        """virtual_array.cpp:176-213: the bicubic sample at each cell centre
        scaled to the array, x = (i + 0.5) (n_a - 1) / (n - 1), corner
        floor(x) clamped into the array, offset x - floor(x) kept."""
        ny_a, nx_a = a.shape
        rx = F(F(nx_a - 1) / F(self.shape[0] - 1))
        ry = F(F(ny_a - 1) / F(self.shape[1] - 1))
        for r in self.regions():
            i0, j0 = self.origin(r)
            xs = ((i0 + np.arange(r.shape[0])).astype(np.float32) + F(0.5)) * rx
            ys = ((j0 + np.arange(r.shape[1])).astype(np.float32) + F(0.5)) * ry
            ig, jg = np.floor(xs).astype(np.int64), np.floor(ys).astype(np.int64)
            u, v = (xs - ig.astype(np.float32)), (ys - jg.astype(np.float32))
            self.tile(r)[...] = bicubic_at(a, np.clip(ig, 0, nx_a - 1)[None, :],
                                           np.clip(jg, 0, ny_a - 1)[:, None],
                                           u[None, :], v[:, None])

    def single_array(self, func, stride=1, k_cutoff=None):
        """virtual_array.inl:225-375 for one output array: VA_SINGLE_ARRAY
        (k_cutoff None) or VA_SINGLE_ARRAY_DOWNSCALED.  func takes and
        returns a (ny, nx) float32 array of the working shape."""
        nx, ny = self.shape
        wrk = (nx, ny)
        if k_cutoff is None and stride > 1:
            wrk = (max(2, int(F(nx) / F(stride))), max(2, int(F(ny) / F(stride))))
        if k_cutoff is not None and F(k_cutoff) != F(1):
            wrk = (max(2, int(F(nx) * F(k_cutoff))), max(2, int(F(ny) * F(k_cutoff))))
        full = self.to_array()
        out = np.asarray(func(resample_bilinear(full, wrk)), np.float32)
        if k_cutoff is not None:
            low = resample_bicubic(resample_bilinear(full, wrk), (nx, ny))
            self.from_array(resample_bicubic(out, (nx, ny)) + (full - low))
        elif wrk == (nx, ny):
            self.from_array(out)
        else:
            self.from_array_bicubic(out)

    def from_array(self, a):
        """virtual_array.cpp:104-130 (nearest source cell of each centre,
        halo cells included)."""
        ny_a, nx_a = a.shape
        for r in self.regions():
            ii, jj, ig, jg = self._index(r, nx_a, ny_a, False)
            self.tile(r)[...] = a[jg[:, None], ig[None, :]]

    def to_array(self, shape=None):
        """virtual_array.cpp:470-501 (interior cells only, tile order; a
        later tile's cell overwrites an earlier one's at a shared
        destination)."""
        nx_a, ny_a = shape or self.shape
        out = np.zeros((ny_a, nx_a), np.float32)
        for r in self.regions():
            ii, jj, ig, jg = self._index(r, nx_a, ny_a, True)
            t = self.tile(r)
            # within a tile the cells are written row by row, and the index
            # maps are non-decreasing: the last write to a destination comes
            # from the last i (j) mapping to its column (row)
            ki, kj = _last(ig), _last(jg)
            out[jg[kj][:, None], ig[ki][None, :]] = t[jj[kj][:, None], ii[ki][None, :]]
        return out

    def local(self, gi, gj):
        """tile_region_from_global_index and local_indices."""
        r = self.region(gi // self.tile_shape[0], gj // self.tile_shape[1])
        i0, j0 = self.origin(r)
        return r, gi - i0, gj - j0

    def get(self, gi, gj):
        r, li, lj = self.local(gi, gj)
        return self.tile(r)[lj, li]

    def set(self, gi, gj, v):
        r, li, lj = self.local(gi, gj)
        self.tile(r)[lj, li] = F(v)

    def _unit(self, x, y):
        b0, b1, b2, b3 = self.bbox
        xr = min(max(F(F(F(x) - b0) / F(b1 - b0)), F(0)), F(1))
        yr = min(max(F(F(F(y) - b2) / F(b3 - b2)), F(0)), F(1))
        return F(xr * F(self.shape[0] - 1)), F(yr * F(self.shape[1] - 1))

    def get_nearest(self, x, y):
        """virtual_array.cpp:267-279."""
        xi, yi = self._unit(x, y)
        return self.get(int(xi), int(yi))

    def get_bilinear(self, x, y):
        """virtual_array.cpp:234-265, bilinear_interp (interpolate2d.hpp:94-105)."""
        xi, yi = self._unit(x, y)
        gi, gj = int(xi), int(yi)
        u, v = F(xi - F(gi)), F(yi - F(gj))
        gi1 = gi - 1 if gi == self.shape[0] - 1 else gi + 1
        gj1 = gj - 1 if gj == self.shape[1] - 1 else gj + 1
        f00, f10 = self.get(gi, gj), self.get(gi1, gj)
        f01, f11 = self.get(gi, gj1), self.get(gi1, gj1)
        a10, a01 = F(f10 - f00), F(f01 - f00)
        a11 = F(F(F(f11 - f10) - f01) + f00)
        return F(F(F(f00 + F(a10 * u)) + F(a01 * v)) + F(F(a11 * u) * v))

    def min(self):
        return min(F(self.tile(r).min()) for r in self.regions())

    def max(self):
        return max(F(self.tile(r).max()) for r in self.regions())

    def mean(self):
        return F(self.sum() / F(self.shape[0] * self.shape[1]))

    def remap(self, vmin, vmax):
        """virtual_array_processing.cpp:209-223 and range.cpp:363-377."""
        lo, hi = self.min(), self.max()
        vmin, vmax = F(vmin), F(vmax)
        for r in self.regions():
            t = self.tile(r)
            if lo == hi:
                t[...] = vmin
                continue
            scale = F(F(vmax - vmin) / F(hi - lo))
            offset = F(vmin - F(lo * scale))
            t[...] = (t * scale).astype(np.float32) + offset

    def range_percentile(self, p_low, p_high, bins=1024):
        """virtual_array_processing.cpp:96-164."""
        lo_r, hi_r = self.min(), self.max()
        if lo_r == hi_r:
            return lo_r, hi_r
        hist = np.zeros(bins, np.int64)
        span = F(hi_r - lo_r)
        for r in self.regions():
            hl, hr, hb, ht = r.halo
            v = self.tile(r)[hb:r.shape[1] - ht, hl:r.shape[0] - hr].ravel()
            idx = (((v - lo_r) / span).astype(np.float32) * F(bins)).astype(np.float32)
            idx = np.minimum(idx.astype(np.int64), bins - 1)
            hist += np.bincount(idx, minlength=bins)
        total = int(hist.sum())
        lo_t = int(F(F(p_low) * F(total)))
        hi_t = int(F(F(p_high) * F(total)))
        acc, lo, hi = 0, lo_r, hi_r
        for b in range(bins):
            acc += int(hist[b])
            at = F(lo_r + F(span * F(F(b) / F(bins))))
            if acc >= lo_t and lo == lo_r:
                lo = at
            if acc >= hi_t:
                hi = at
                break
        return lo, hi

    def unique_values(self):
        """The sorted distinct values of all tiles, halo cells included."""
        return np.unique(np.concatenate([self.tile(r).ravel() for r in self.regions()]))

    def sum(self):
        """virtual_array_processing.cpp:225-251: float sums of interiors in
        row-major order (numpy's float32 cumsum adds sequentially), then over
        tiles in tile order."""
        total = F(0)
        for r in self.regions():
            hl, hr, hb, ht = r.halo
            inner = self.tile(r)[hb:r.shape[1] - ht, hl:r.shape[0] - hr].ravel()
            s = np.cumsum(inner, dtype=np.float32)[-1] if inner.size else F(0)
            total = F(total + s)
        return total

    def smooth_overlap_buffers(self):
        """virtual_array.cpp:337-401: pairs in order; within a pair every
        cell is independent, so a pair is one vector operation."""
        ntx, nty = self.max_tiles()
        hw = self.halo
        ramp = smoothstep5(np.arange(hw, dtype=np.float32) / F(hw - 1))
        for ty in range(nty):
            for tx in range(ntx - 1):
                t0 = self.tile(self.region(tx, ty))
                t1 = self.tile(self.region(tx + 1, ty))
                pb = t0.shape[1] - 2 * hw
                t1[:, :hw] = lerp(t0[:, pb:pb + hw], t1[:, :hw], ramp[None, :])
                t0[:, pb:pb + hw] = t1[:, :hw]
        for ty in range(nty - 1):
            for tx in range(ntx):
                t0 = self.tile(self.region(tx, ty))
                t1 = self.tile(self.region(tx, ty + 1))
                qb = t0.shape[0] - 2 * hw
                t1[:hw, :] = lerp(t0[qb:qb + hw, :], t1[:hw, :], ramp[:, None])
                t0[qb:qb + hw, :] = t1[:hw, :]


def catmull_rom(p0, p1, p2, p3, x):
    """interpolate2d.hpp:107-113, elementwise: the differences p1 - p2 and
    p2 - p0 in float32, every other operation in float64 in the source's
    order, the result rounded to float32."""
    p = [np.asarray(q, np.float32) for q in (p0, p1, p2, p3)]
    d12 = (p[1] - p[2]).astype(np.float64)
    d20 = (p[2] - p[0]).astype(np.float64)
    P0, P1, P2, P3 = (q.astype(np.float64) for q in p)
    x = np.asarray(x, np.float32).astype(np.float64)
    c3 = 3.0 * d12 + P3 - P0
    c2 = 2.0 * P0 - 5.0 * P1 + 4.0 * P2 - P3 + x * c3
    c1 = d20 + x * c2
    return (P1 + 0.5 * x * c1).astype(np.float32)


def bilinear(f00, f10, f01, f11, u, v):
    """interpolate2d.hpp:94-105 in float32, left to right."""
    a10, a01 = f10 - f00, f01 - f00
    a11 = ((f11 - f10) - f01) + f00
    return (((f00 + a10 * u) + a01 * v) + (a11 * u) * v).astype(np.float32)


def bicubic_at(a, i, j, u, v):
    """Array::get_value_bicubic_at (methods.cpp:203-223) for broadcastable
    index and offset arrays: Catmull-Rom along j in the columns i-1..i+2,
    then along i; reads clamped into the array."""
    ny, nx = a.shape
    col = []
    for m in range(-1, 3):
        ic = np.clip(i + m, 0, nx - 1)
        p = [a[np.clip(j + n, 0, ny - 1), ic] for n in range(-1, 3)]
        col.append(catmull_rom(*p, v))
    return catmull_rom(*col, u)


def sample_axis(nt, ns):
    """Source coordinates of the nt target points of one axis
    (interpolate_array.cpp with endpoint false, pixel-centred, unit boxes):
    s = 0.5 / nt, x_i = s + i ((1 - s) / nt), coordinate x_i / (1 / ns) - 0.5,
    all float32."""
    s = F(F(0.5) * F(F(1) / F(nt)))
    step = F(F(F(1) - s) / F(nt))
    x = s + np.arange(nt).astype(np.float32) * step if nt > 1 else np.array([s], np.float32)
    return (x.astype(np.float32) / F(F(1) / F(ns))).astype(np.float32) - F(0.5)


def resample_bilinear(a, shape):
    """Array::resample_to_shape (interpolate_array.cpp:122-185): the lower
    corner truncated and clamped before the offsets are taken."""
    ny, nx = a.shape
    cx, cy = sample_axis(shape[0], nx), sample_axis(shape[1], ny)
    i0 = np.clip(cx.astype(np.int64), 0, nx - 1)
    j0 = np.clip(cy.astype(np.int64), 0, ny - 1)
    u = (cx - i0.astype(np.float32))[None, :]
    v = (cy - j0.astype(np.float32))[:, None]
    i1, j1 = np.minimum(i0 + 1, nx - 1), np.minimum(j0 + 1, ny - 1)
    g = lambda jj, ii: a[jj[:, None], ii[None, :]]
    return bilinear(g(j0, i0), g(j0, i1), g(j1, i0), g(j1, i1), u, v)


def resample_bicubic(a, shape):
    """Array::resample_to_shape_bicubic (interpolate_array.cpp:31-103): the
    corner truncated toward zero, not clamped, for the offsets."""
    ny, nx = a.shape
    cx, cy = sample_axis(shape[0], nx), sample_axis(shape[1], ny)
    i0, j0 = cx.astype(np.int64), cy.astype(np.int64)
    u = (cx - i0.astype(np.float32))[None, :]
    v = (cy - j0.astype(np.float32))[:, None]
    return bicubic_at(a, i0[None, :], j0[:, None], u, v)


def _last(idx):
    """Mask of the last occurrence of each value of a non-decreasing array."""
    keep = np.ones(len(idx), bool)
    keep[:-1] = idx[1:] != idx[:-1]
    return keep


def smoothstep5(x):
    """HighMap/src/math/core.cpp:97-100, float32 (scalars or arrays)."""
    x = np.asarray(x, np.float32)
    return (((x * x) * x) * ((x * ((x * F(6)) - F(15))) + F(10))).astype(np.float32)


def lerp(a, b, t):
    """std::lerp(float, float, float) as libstdc++ (C++20) computes it,
    elementwise on float32 scalars or arrays."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    t = np.asarray(t, np.float32)
    a, b, t = np.broadcast_arrays(a, b, t)
    straddle = ((a <= 0) & (b >= 0)) | ((a >= 0) & (b <= 0))
    r1 = (t * b) + ((F(1) - t) * a)
    x = a + t * (b - a)
    r2 = np.where((t > 1) == (b > a), np.where(b < x, x, b), np.where(b > x, x, b))
    out = np.where(straddle, r1, np.where(t == 1, b, r2))
    return out.astype(np.float32)
