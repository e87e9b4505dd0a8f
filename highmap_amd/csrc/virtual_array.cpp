// virtual_array.cpp -- hmap::VirtualArray of the drop-in (host side).
//
// Tile geometry, the tile loops and the whole-array operations follow
// HighMap/src/virtual_array/virtual_array.cpp,
// virtual_array_processing.cpp and virtual_array.inl; the arithmetic of
// each formula is kept (float operations in the same order) so that tile
// regions, bounding boxes and blended overlaps come out the same.
#include "hmap_virtual_array.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <future>
#include <limits>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace hmap
{

int ceil_div(int a, int b)
{
  return (a + b - 1) / b;
}

// ------------------------------------------------------------- TileRegion

TileRegion::TileRegion(const TileKey    &key,
                       const glm::vec4  &bbox,
                       const glm::ivec2 &shape,
                       const glm::ivec4 &halo)
    : key(key), bbox(bbox), shape(shape), halo(halo)
{
}

glm::vec2 TileRegion::cell_center(int i, int j) const
{
  // tile_region.cpp:21-34
  const float dx = (bbox.y - bbox.x) / float(shape.x);
  const float dy = (bbox.w - bbox.z) / float(shape.y);
  return glm::vec2(bbox.x + (i + 0.5f) * dx, bbox.z + (j + 0.5f) * dy);
}

glm::vec2 TileRegion::cell_corner(int i, int j) const
{
  // tile_region.cpp:36-41, glm::mix(a, b, t) = a * (1 - t) + b * t
  const float tx = float(i) / shape.x, ty = float(j) / shape.y;
  return glm::vec2(bbox.x * (1.f - tx) + bbox.y * tx, bbox.z * (1.f - ty) + bbox.w * ty);
}

std::string TileRegion::key_as_string() const
{
  return std::to_string(key.tx) + "_" + std::to_string(key.ty);
}

// ---------------------------------------------------------------- storage

RamTileStorage::RamTileStorage(const RamTileStorage &other)
{
  std::lock_guard<std::mutex> lk(const_cast<std::mutex &>(other.mutex));
  tiles = other.tiles;
}

Array &RamTileStorage::get_tile(const TileRegion &region)
{
  // ram_tile_storage.cpp:23-32: created zero-filled at the region's shape
  // (halo included) on first access.  Map nodes do not move, so the
  // reference stays valid while other threads insert other tiles.
  std::lock_guard<std::mutex> lk(mutex);
  auto                        it = tiles.find(region.key);
  if (it != tiles.end()) return it->second;
  return tiles.emplace(region.key, Array(region.shape)).first->second;
}

void RamTileStorage::release_tile(const TileRegion &) {}

size_t RamTileStorage::max_live_tiles() const
{
  return std::numeric_limits<size_t>::max();
}

std::unique_ptr<TileStorage> RamTileStorage::clone() const
{
  return std::make_unique<RamTileStorage>(*this);
}

std::unique_ptr<TileStorage> make_storage(glm::ivec2, glm::ivec2, StorageMode mode)
{
  if (mode != StorageMode::VA_RAM)
    throw std::invalid_argument("hmap::VirtualArray: only the RAM tile storage is provided");
  return std::make_unique<RamTileStorage>();
}

// ----------------------------------------------------------- VirtualArray

VirtualArray::VirtualArray(glm::ivec2                   shape,
                           glm::vec4                    bbox,
                           glm::ivec2                   tile_shape,
                           int                          halo,
                           std::unique_ptr<TileStorage> storage)
    : shape(shape), bbox(bbox), tile_shape(tile_shape), halo(halo), storage(std::move(storage))
{
}

VirtualArray::VirtualArray(glm::ivec2  shape,
                           glm::vec4   bbox,
                           glm::ivec2  tile_shape,
                           int         halo,
                           StorageMode storage_mode)
    : shape(shape), bbox(bbox), tile_shape(tile_shape), halo(halo),
      storage(make_storage(shape, tile_shape, storage_mode))
{
}

VirtualArray::VirtualArray(glm::ivec2 shape, glm::ivec2 tile_shape, int halo,
                           StorageMode storage_mode)
    : VirtualArray(shape, glm::vec4(0.f, 1.f, 0.f, 1.f), tile_shape, halo, storage_mode)
{
}

// clone and copy_from start from an empty RAM tile store of the copied
// geometry, whatever store the source uses (as HighMap's do,
// virtual_array.cpp:76-94); the data, if asked for, is copied tile by tile.
static std::unique_ptr<TileStorage> empty_ram_store(glm::ivec2 shape, glm::ivec2 tile_shape)
{
  return make_storage(shape, tile_shape, StorageMode::VA_RAM);
}

std::unique_ptr<VirtualArray> VirtualArray::clone(const ComputeMode &cm, bool deep_copy)
{
  auto copy = std::make_unique<VirtualArray>(shape, bbox, tile_shape, halo,
                                             empty_ram_store(shape, tile_shape));
  if (deep_copy) copy_data(*this, *copy, cm);
  return copy;
}

void VirtualArray::copy_from(VirtualArray &src, const ComputeMode &cm, bool copy_src_data)
{
  if (this == &src) return;
  storage = empty_ram_store(src.shape, src.tile_shape);
  std::tie(shape, bbox, tile_shape, halo) =
      std::make_tuple(src.shape, src.bbox, src.tile_shape, src.halo);
  if (copy_src_data) copy_data(src, *this, cm);
}

glm::ivec2 VirtualArray::get_max_tiles() const
{
  return glm::ivec2(ceil_div(shape.x, tile_shape.x), ceil_div(shape.y, tile_shape.y));
}

int VirtualArray::get_ntiles() const
{
  const glm::ivec2 m = get_max_tiles();
  return m.x * m.y;
}

TileRegion VirtualArray::tile_region_from_tile_coords(int tile_x, int tile_y) const
{
  // virtual_array.cpp:429-468: interior w x h, a halo on every side that
  // has a neighbour tile, bbox from the cell range in float
  const int w = std::min(tile_shape.x, shape.x - tile_x * tile_shape.x);
  const int h = std::min(tile_shape.y, shape.y - tile_y * tile_shape.y);
  const int tx_max = ceil_div(shape.x, tile_shape.x) - 1;
  const int ty_max = ceil_div(shape.y, tile_shape.y) - 1;
  const glm::ivec4 h4(tile_x == 0 ? 0 : halo, tile_x == tx_max ? 0 : halo,
                      tile_y == 0 ? 0 : halo, tile_y == ty_max ? 0 : halo);
  const float sx = bbox.y - bbox.x, sy = bbox.w - bbox.z;
  const float x1 = bbox.x + sx * (tile_x * tile_shape.x - h4.x) / float(shape.x);
  const float x2 = bbox.x + sx * (tile_x * tile_shape.x + w + h4.y) / float(shape.x);
  const float y1 = bbox.z + sy * (tile_y * tile_shape.y - h4.z) / float(shape.y);
  const float y2 = bbox.z + sy * (tile_y * tile_shape.y + h + h4.w) / float(shape.y);
  return TileRegion(TileKey{tile_x, tile_y}, glm::vec4(x1, x2, y1, y2),
                    glm::ivec2(w + h4.x + h4.y, h + h4.z + h4.w), h4);
}

TileRegion VirtualArray::tile_region_from_global_index(int global_i, int global_j) const
{
  return tile_region_from_tile_coords(global_i / tile_shape.x, global_j / tile_shape.y);
}

glm::vec2 VirtualArray::tile_region_global_position(const TileRegion &region) const
{
  return glm::vec2((region.bbox.x - bbox.x) / (bbox.y - bbox.x),
                   (region.bbox.z - bbox.z) / (bbox.w - bbox.z));
}

glm::ivec2 VirtualArray::tile_region_global_indices(const TileRegion &region) const
{
  return glm::ivec2(region.key.tx * tile_shape.x - region.halo.x,
                    region.key.ty * tile_shape.y - region.halo.z);
}

glm::ivec2 VirtualArray::local_indices(const TileRegion &region, int global_i,
                                       int global_j) const
{
  const int tx = global_i / tile_shape.x, ty = global_j / tile_shape.y;
  return glm::ivec2(global_i - (tx * tile_shape.x - region.halo.x),
                    global_j - (ty * tile_shape.y - region.halo.z));
}

float VirtualArray::get(int global_i, int global_j) const
{
  const TileRegion region = tile_region_from_global_index(global_i, global_j);
  const Array     &tile = storage->get_tile(region);
  const glm::ivec2 l = local_indices(region, global_i, global_j);
  const float      v = tile(l.x, l.y);
  storage->release_tile(region);
  return v;
}

void VirtualArray::set(int global_i, int global_j, float v)
{
  const TileRegion region = tile_region_from_global_index(global_i, global_j);
  Array           &tile = storage->get_tile(region);
  const glm::ivec2 l = local_indices(region, global_i, global_j);
  tile(l.x, l.y) = v;
  storage->release_tile(region);
}

float VirtualArray::get_bilinear(float x, float y) const
{
  // virtual_array.cpp:234-265 with bilinear_interp (interpolate2d.hpp:94-105)
  float xr = std::clamp((x - bbox.x) / (bbox.y - bbox.x), 0.f, 1.f);
  float yr = std::clamp((y - bbox.z) / (bbox.w - bbox.z), 0.f, 1.f);
  const float xi = xr * (shape.x - 1.f), yi = yr * (shape.y - 1.f);
  const int   gi = int(xi), gj = int(yi);
  const float u = xi - gi, v = yi - gj;
  const int   gi1 = gi == shape.x - 1 ? gi - 1 : gi + 1;
  const int   gj1 = gj == shape.y - 1 ? gj - 1 : gj + 1;
  const float f00 = get(gi, gj), f10 = get(gi1, gj), f01 = get(gi, gj1), f11 = get(gi1, gj1);
  const float a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
  return f00 + a10 * u + a01 * v + a11 * u * v;
}

float VirtualArray::get_nearest(float x, float y) const
{
  const float xr = std::clamp((x - bbox.x) / (bbox.y - bbox.x), 0.f, 1.f);
  const float yr = std::clamp((y - bbox.z) / (bbox.w - bbox.z), 0.f, 1.f);
  return get(int(xr * (shape.x - 1.f)), int(yr * (shape.y - 1.f)));
}

void VirtualArray::fill(float value, const ComputeMode &cm)
{
  for_each_tile(
      *this, [value](Array &tile, const TileRegion &)
      { std::fill(tile.vector.begin(), tile.vector.end(), value); },
      cm);
}

void VirtualArray::from_array(const Array &array, const ComputeMode &cm)
{
  // virtual_array.cpp:104-130: nearest source cell of each destination
  // cell centre, halo cells included
  auto lambda = [&array, this](Array &tile, const TileRegion &region)
  {
    const float      rx = float(array.shape.x - 1) / float(shape.x - 1);
    const float      ry = float(array.shape.y - 1) / float(shape.y - 1);
    const glm::ivec2 ij0 = tile_region_global_indices(region);
    for (int j = 0; j < region.shape.y; ++j)
      for (int i = 0; i < region.shape.x; ++i)
      {
        const float xd = float(ij0.x + i) + 0.5f, yd = float(ij0.y + j) + 0.5f;
        const int   ig = std::clamp(int(std::floor(xd * rx)), 0, array.shape.x - 1);
        const int   jg = std::clamp(int(std::floor(yd * ry)), 0, array.shape.y - 1);
        tile(i, j) = array(ig, jg);
      }
  };
  for_each_tile(*this, lambda, cm);
}

Array VirtualArray::to_array(const glm::ivec2 &array_shape, const ComputeMode &cm) const
{
  // virtual_array.cpp:470-501: the interior cells of every tile, each to
  // the destination cell holding its centre
  Array array(array_shape);
  auto  lambda = [&array, this](const Array &tile, const TileRegion &region)
  {
    const float      rx = float(array.shape.x - 1) / float(shape.x - 1);
    const float      ry = float(array.shape.y - 1) / float(shape.y - 1);
    const glm::ivec2 ij0 = tile_region_global_indices(region);
    const glm::ivec4 &b = region.halo;
    for (int j = b.z; j < region.shape.y - b.w; ++j)
      for (int i = b.x; i < region.shape.x - b.y; ++i)
      {
        const float xs = float(ij0.x + i) + 0.5f, ys = float(ij0.y + j) + 0.5f;
        const int   ig = std::clamp(int(std::floor(xs * rx)), 0, array.shape.x - 1);
        const int   jg = std::clamp(int(std::floor(ys * ry)), 0, array.shape.y - 1);
        array(ig, jg) = tile(i, j);
      }
  };
  for_each_tile(*this, lambda, cm);
  return array;
}

Array VirtualArray::to_array(const ComputeMode &cm) const
{
  return to_array(shape, cm);
}

Array VirtualArray::to_array_dbg() const
{
  Array array(shape);
  for (int j = 0; j < shape.y; ++j)
    for (int i = 0; i < shape.x; ++i) array(i, j) = get(i, j);
  return array;
}

void VirtualArray::trim_storage() const
{
  storage->trim();
}

// virtual_array_processing.cpp: per-tile partials in tile order.  min, max
// and range read whole tiles (halo included), sum the interior cells.

float VirtualArray::max(const ComputeMode &cm) const
{
  const glm::ivec2 nt = get_max_tiles();
  if (nt.x * nt.y == 0) return 0.f;
  std::vector<float> part(nt.x * nt.y);
  for_each_tile(
      *this, [&part, nt](const Array &tile, const TileRegion &region)
      { part[region.key.ty * nt.x + region.key.tx] = tile.max(); },
      cm);
  return *std::max_element(part.begin(), part.end());
}

float VirtualArray::min(const ComputeMode &cm) const
{
  const glm::ivec2 nt = get_max_tiles();
  if (nt.x * nt.y == 0) return 0.f;
  std::vector<float> part(nt.x * nt.y);
  for_each_tile(
      *this, [&part, nt](const Array &tile, const TileRegion &region)
      { part[region.key.ty * nt.x + region.key.tx] = tile.min(); },
      cm);
  return *std::min_element(part.begin(), part.end());
}

glm::vec2 VirtualArray::range(const ComputeMode &cm) const
{
  const glm::ivec2 nt = get_max_tiles();
  if (nt.x * nt.y == 0) return glm::vec2(0.f, 0.f);
  std::vector<float> lo(nt.x * nt.y), hi(nt.x * nt.y);
  for_each_tile(
      *this,
      [&lo, &hi, nt](const Array &tile, const TileRegion &region)
      {
        const int k = region.key.ty * nt.x + region.key.tx;
        hi[k] = tile.max();
        lo[k] = tile.min();
      },
      cm);
  return glm::vec2(*std::min_element(lo.begin(), lo.end()),
                   *std::max_element(hi.begin(), hi.end()));
}

glm::vec2 VirtualArray::range_percentile(float p_low, float p_high, const ComputeMode &cm,
                                         size_t bins) const
{
  // virtual_array_processing.cpp:96-164: a histogram of the interior cells
  // over the whole-tile range, then the first bins reaching p_low and
  // p_high of the count
  const glm::ivec2 nt = get_max_tiles();
  if (nt.x * nt.y == 0) return glm::vec2(0.f, 0.f);
  const glm::vec2 rg = range(cm);
  if (rg.x == rg.y) return rg;
  std::vector<std::vector<size_t>> part(nt.x * nt.y, std::vector<size_t>(bins, 0));
  for_each_tile(
      *this,
      [&](const Array &tile, const TileRegion &region)
      {
        auto             &hist = part[region.key.ty * nt.x + region.key.tx];
        const glm::ivec4 &h = region.halo;
        for (int j = h.z; j < region.shape.y - h.w; ++j)
          for (int i = h.x; i < region.shape.x - h.y; ++i)
          {
            const float v = tile(i, j);
            hist[std::min<size_t>(bins - 1, size_t((v - rg.x) / (rg.y - rg.x) * bins))]++;
          }
      },
      cm);
  std::vector<size_t> hist(bins, 0);
  for (const auto &ph : part)
    for (size_t b = 0; b < bins; ++b) hist[b] += ph[b];
  size_t total = 0;
  for (size_t v : hist) total += v;
  const size_t lo_target = size_t(p_low * total), hi_target = size_t(p_high * total);
  size_t       acc = 0;
  float        lo = rg.x, hi = rg.y;
  for (size_t b = 0; b < bins; ++b)
  {
    acc += hist[b];
    if (acc >= lo_target && lo == rg.x) lo = rg.x + (rg.y - rg.x) * (float(b) / bins);
    if (acc >= hi_target)
    {
      hi = rg.x + (rg.y - rg.x) * (float(b) / bins);
      break;
    }
  }
  return glm::vec2(lo, hi);
}

std::vector<float> VirtualArray::unique_values(const ComputeMode &cm) const
{
  // virtual_array_processing.cpp:253-281 and vector_unique_values
  // (vector_utils.cpp:280-287): the sorted distinct values of every tile,
  // halo cells included
  const glm::ivec2 nt = get_max_tiles();
  if (nt.x * nt.y == 0) return {0.f};
  auto uniq = [](std::vector<float> &v)
  {
    v.erase(std::unique(v.begin(), v.end()), v.end());
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  std::vector<std::vector<float>> part(nt.x * nt.y);
  for_each_tile(
      *this,
      [&](const Array &tile, const TileRegion &region)
      {
        std::vector<float> v = tile.vector;
        uniq(v);
        part[region.key.ty * nt.x + region.key.tx] = std::move(v);
      },
      cm);
  std::vector<float> all;
  for (const auto &v : part) all.insert(all.end(), v.begin(), v.end());
  uniq(all);
  return all;
}

float VirtualArray::sum(const ComputeMode &cm) const
{
  const glm::ivec2 nt = get_max_tiles();
  if (nt.x * nt.y == 0) return 0.f;
  std::vector<float> part(nt.x * nt.y);
  for_each_tile(
      *this,
      [&part, nt](const Array &tile, const TileRegion &region)
      {
        const glm::ivec4 &h = region.halo;
        float             s = 0.f;
        for (int j = h.z; j < region.shape.y - h.w; ++j)
          for (int i = h.x; i < region.shape.x - h.y; ++i) s += tile(i, j);
        part[region.key.ty * nt.x + region.key.tx] = s;
      },
      cm);
  float s = 0.f;
  for (float v : part) s += v;
  return s;
}

float VirtualArray::mean(const ComputeMode &cm) const
{
  return sum(cm) / (float)(shape.x * shape.y);
}

void VirtualArray::inverse(const ComputeMode &cm)
{
  const float hmax = max(cm);
  for_each_tile(
      *this,
      [hmax](Array &tile, const TileRegion &)
      {
        for (float &x : tile.vector) x = x * -1.f;
        for (float &x : tile.vector) x = x + hmax;
      },
      cm);
}

void VirtualArray::remap(float vmin, float vmax, const ComputeMode &cm)
{
  const float lo = min(cm), hi = max(cm);
  remap(vmin, vmax, lo, hi, cm);
}

void VirtualArray::remap(float vmin, float vmax, float from_min, float from_max,
                         const ComputeMode &cm)
{
  for_each_tile(
      *this, [=](Array &tile, const TileRegion &)
      { hmap::remap(tile, vmin, vmax, from_min, from_max); },
      cm);
}

static float smoothstep5(float x)
{
  // HighMap/src/math/core.cpp:97-100
  return x * x * x * (x * (x * 6.f - 15.f) + 10.f);
}

// The overlap band between two neighbouring tiles along one axis.  Along
// that axis the first `halo` cells of `hi` (the right or upper tile) are the
// grid cells [n - 2 halo, n - halo) of `lo` (n its extent): band position k
// of both becomes lerp(lo's value, hi's value, w_k), w_k = smoothstep5(k /
// (halo - 1)), so the seam fades from lo's result into hi's.  The weights
// are formed once per band; each cell pair is independent.
static void blend_overlap_band(Array &lo, Array &hi, int halo, bool along_x)
{
  std::vector<float> w((size_t)std::max(halo, 0));
  for (int k = 0; k < halo; k++) w[(size_t)k] = smoothstep5(float(k) / float(halo - 1));
  const int first = (along_x ? lo.shape.x : lo.shape.y) - 2 * halo; // band start in lo
  const int across = along_x ? lo.shape.y : lo.shape.x;
  for (int k = 0; k < halo; k++)
    for (int c = 0; c < across; c++)
    {
      float      &a = along_x ? lo(first + k, c) : lo(c, first + k);
      float      &b = along_x ? hi(k, c) : hi(c, k);
      const float v = std::lerp(a, b, w[(size_t)k]);
      b = v;
      a = v;
    }
}

void VirtualArray::smooth_overlap_buffers()
{
  // semantics of HighMap's VirtualArray::smooth_overlap_buffers
  // (virtual_array.cpp:337-401): every pair of neighbours along x, then
  // every pair along y, blended as blend_overlap_band describes
  const size_t live = storage->max_live_tiles();
  if (live < 2)
  {
    std::fprintf(stderr,
                 "hmap::VirtualArray: overlap smoothing needs two tiles resident at once, "
                 "the tile storage keeps %zu; the overlaps are left as they are\n",
                 live);
    return;
  }
  const int tiles_x = ceil_div(shape.x, tile_shape.x), tiles_y = ceil_div(shape.y, tile_shape.y);
  auto blend = [&](int ax, int ay, int bx, int by, bool along_x)
  {
    const TileRegion ra = tile_region_from_tile_coords(ax, ay);
    const TileRegion rb = tile_region_from_tile_coords(bx, by);
    Array           &a = storage->get_tile(ra);
    Array           &b = storage->get_tile(rb);
    blend_overlap_band(a, b, halo, along_x);
    storage->release_tile(ra);
    storage->release_tile(rb);
  };
  for (int ty = 0; ty < tiles_y; ++ty)
    for (int tx = 0; tx + 1 < tiles_x; ++tx) blend(tx, ty, tx + 1, ty, true);
  for (int ty = 0; ty + 1 < tiles_y; ++ty)
    for (int tx = 0; tx < tiles_x; ++tx) blend(tx, ty, tx, ty + 1, false);
}

void copy_data(VirtualArray &src, VirtualArray &dst, const ComputeMode &cm)
{
  for_each_tile(
      std::vector<VirtualArray *>{&src, &dst},
      [](std::vector<Array *> &a, const TileRegion &) { *a[1] = *a[0]; }, cm);
}

void remap(Array &array, float vmin, float vmax, float from_min, float from_max)
{
  // range.cpp:363-377
  if (from_min == from_max)
  {
    std::fill(array.vector.begin(), array.vector.end(), vmin);
    return;
  }
  const float scale = (vmax - vmin) / (from_max - from_min);
  const float offset = vmin - from_min * scale;
  for (float &x : array.vector) x = x * scale + offset;
}

// ------------------------------------------------------------- tile loops

void for_each_tile(const TileAccess &access, const TileFunc &func, const ComputeMode &cm)
{
  // virtual_array.inl:19-109
  if (access.outputs.empty())
  {
    std::fprintf(stderr, "for_each_tile: no output VirtualArray\n");
    return;
  }
  const VirtualArray *ref = access.ref_va();
  if (!ref) return;
  auto same = [ref](const VirtualArray *va)
  { return !va || (va->shape == ref->shape && va->tile_shape == ref->tile_shape); };
  for (auto *va : access.inputs)
    if (!same(va)) throw std::invalid_argument("for_each_tile: incompatible input VirtualArray");
  for (auto *va : access.outputs)
    if (!same(va))
      throw std::invalid_argument("for_each_tile: incompatible output VirtualArray");

  auto dispatch = [&](const TileRegion &region)
  {
    std::vector<const Array *> in;
    std::vector<Array *>       out;
    for (auto *va : access.inputs) in.push_back(va ? &va->storage->get_tile(region) : nullptr);
    for (auto *va : access.outputs) out.push_back(va ? &va->storage->get_tile(region) : nullptr);
    func(in, out, region);
    for (auto *va : access.inputs)
      if (va) va->storage->release_tile(region);
    for (auto *va : access.outputs)
      if (va) va->storage->release_tile(region);
  };

  const int ntx = ceil_div(ref->shape.x, ref->tile_shape.x);
  const int nty = ceil_div(ref->shape.y, ref->tile_shape.y);
  switch (cm.mode)
  {
  case ForEachMode::VA_SEQUENTIAL:
    for (int ty = 0; ty < nty; ++ty)
      for (int tx = 0; tx < ntx; ++tx) dispatch(ref->tile_region_from_tile_coords(tx, ty));
    break;
  case ForEachMode::VA_DISTRIBUTED:
  {
    // worker t takes tiles t, t + nthreads, ... in row-major tile order
    const int ntasks = ntx * nty;
    int       nthreads = (int)std::thread::hardware_concurrency();
    nthreads = std::max(1, std::min(nthreads, ntasks));
    std::vector<std::future<void>> fut;
    for (int t = 0; t < nthreads; ++t)
      fut.emplace_back(std::async(std::launch::async,
                                  [&, t]()
                                  {
                                    for (int k = t; k < ntasks; k += nthreads)
                                      dispatch(ref->tile_region_from_tile_coords(k % ntx, k / ntx));
                                  }));
    for (auto &f : fut) f.get();
    break;
  }
  case ForEachMode::VA_SINGLE_ARRAY:
  {
    if (cm.stride > 1)
      throw std::invalid_argument("for_each_tile: strided VA_SINGLE_ARRAY is not provided");
    const ComputeMode  local{ForEachMode::VA_DISTRIBUTED, false, 1, 1.f};
    std::vector<Array> ain, aout;
    for (auto *va : access.inputs) ain.push_back(va ? va->to_array(local) : Array());
    for (auto *va : access.outputs) aout.push_back(va ? va->to_array(local) : Array());
    std::vector<const Array *> in;
    std::vector<Array *>       out;
    for (size_t k = 0; k < access.inputs.size(); ++k)
      in.push_back(access.inputs[k] ? &ain[k] : nullptr);
    for (size_t k = 0; k < access.outputs.size(); ++k)
      out.push_back(access.outputs[k] ? &aout[k] : nullptr);
    const TileRegion whole(TileKey{0, 0}, ref->bbox, ref->shape, glm::ivec4(0, 0, 0, 0));
    func(in, out, whole);
    for (size_t k = 0; k < access.outputs.size(); ++k)
      if (access.outputs[k]) access.outputs[k]->from_array(aout[k], local);
    break;
  }
  default:
    throw std::invalid_argument("for_each_tile: VA_SINGLE_ARRAY_DOWNSCALED is not provided");
  }
  if (cm.trim_storage)
  {
    for (auto *va : access.inputs)
      if (va) va->trim_storage();
    for (auto *va : access.outputs)
      if (va) va->trim_storage();
  }
}

} // namespace hmap
