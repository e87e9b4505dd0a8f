// virtual_array.cpp -- hmap::VirtualArray of the drop-in (host side).
//
// Tile geometry, the tile loops and the whole-array operations follow
// HighMap/src/virtual_array/virtual_array.cpp,
// virtual_array_processing.cpp and virtual_array.inl; the arithmetic of
// each formula is kept (float operations in the same order) so that tile
// regions, bounding boxes and blended overlaps come out the same.
#include "hmap_virtual_array.hpp"

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <future>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <system_error>
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

// ------------------------------------------------------------ LRU storage

LruTileStorage::LruTileStorage(size_t max_tiles) : max_tiles(std::max<size_t>(1, max_tiles)) {}

Array &LruTileStorage::get_tile(const TileRegion &region)
{
  std::lock_guard<std::mutex> lk(mutex);
  auto                        it = tiles.find(region.key);
  if (it == tiles.end())
  {
    // make room first: least recently used tiles that nobody holds leave
    // until the new one fits (or only held tiles are left)
    while (tiles.size() >= max_tiles)
    {
      auto victim = tiles.end();
      for (auto e = tiles.begin(); e != tiles.end(); ++e)
        if (e->second.in_use == 0 &&
            (victim == tiles.end() || e->second.last_use < victim->second.last_use))
          victim = e;
      if (victim == tiles.end()) break;
      on_evict(victim->first, victim->second.tile);
      tiles.erase(victim);
    }
    Entry e;
    e.tile = Array(region.shape);
    on_create(region, e.tile);
    it = tiles.emplace(region.key, std::move(e)).first;
  }
  it->second.in_use++;
  it->second.last_use = ++clock;
  return it->second.tile;
}

void LruTileStorage::release_tile(const TileRegion &region)
{
  std::lock_guard<std::mutex> lk(mutex);
  auto                        it = tiles.find(region.key);
  if (it != tiles.end() && it->second.in_use > 0) it->second.in_use--;
}

size_t LruTileStorage::max_live_tiles() const
{
  return max_tiles;
}

size_t LruTileStorage::resident_tiles() const
{
  std::lock_guard<std::mutex> lk(mutex);
  return tiles.size();
}

std::unique_ptr<TileStorage> LruTileStorage::clone() const
{
  return std::make_unique<LruTileStorage>(max_tiles);
}

std::string LruTileStorage::info_string() const
{
  return "LRU (capacity=" + std::to_string(max_tiles) + ")";
}

void LruTileStorage::on_evict(const TileKey &, const Array &) {}

bool LruTileStorage::on_create(const TileRegion &, Array &)
{
  return false;
}

// ------------------------------------------------------------ disk storages

// A new directory under the system's temporary directory, named from the
// prefix, the process id and a per-process counter (retried past names
// that exist).
static std::filesystem::path fresh_temp_dir(const char *prefix)
{
  static std::atomic<unsigned> counter{0};
  const auto                   base = std::filesystem::temp_directory_path();
  for (;;)
  {
    const auto dir = base / (std::string(prefix) + "_" + std::to_string((long)::getpid()) + "_" +
                             std::to_string(counter.fetch_add(1)));
    if (std::filesystem::create_directory(dir)) return dir;
  }
}

static std::filesystem::path tile_file(const std::filesystem::path &root, const TileKey &key)
{
  return root / ("tile_" + std::to_string(key.tx) + "_" + std::to_string(key.ty) + ".f32");
}

static void write_tile(const std::filesystem::path &path, const Array &tile)
{
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("hmap::VirtualArray: cannot write " + path.string());
  const size_t n = std::fwrite(tile.vector.data(), sizeof(float), tile.vector.size(), f);
  const bool   ok = std::fclose(f) == 0 && n == tile.vector.size();
  if (!ok) throw std::runtime_error("hmap::VirtualArray: short write to " + path.string());
}

// Reads the tile if its file exists (it must hold exactly the tile's cells).
static bool read_tile(const std::filesystem::path &path, Array &tile)
{
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const size_t n = std::fread(tile.vector.data(), sizeof(float), tile.vector.size(), f);
  const bool   whole = n == tile.vector.size() && std::fgetc(f) == EOF;
  std::fclose(f);
  if (!whole) throw std::runtime_error("hmap::VirtualArray: " + path.string() +
                                       " does not hold the tile's cells");
  return true;
}

DiskLruTileStorage::DiskLruTileStorage(size_t max_tiles)
    : LruTileStorage(max_tiles), root(fresh_temp_dir("hmap_va_lru"))
{
}

DiskLruTileStorage::~DiskLruTileStorage()
{
  std::error_code ec;
  std::filesystem::remove_all(root, ec);
}

std::unique_ptr<TileStorage> DiskLruTileStorage::clone() const
{
  return std::make_unique<DiskLruTileStorage>(max_tiles);
}

std::string DiskLruTileStorage::info_string() const
{
  return "DiskLRU (capacity=" + std::to_string(max_tiles) + ")";
}

void DiskLruTileStorage::on_evict(const TileKey &key, const Array &tile)
{
  write_tile(tile_file(root, key), tile);
}

bool DiskLruTileStorage::on_create(const TileRegion &region, Array &tile)
{
  return read_tile(tile_file(root, region.key), tile);
}

void DiskLruTileStorage::trim()
{
  std::lock_guard<std::mutex> lk(mutex);
  for (auto it = tiles.begin(); it != tiles.end();)
    if (it->second.in_use == 0)
    {
      on_evict(it->first, it->second.tile);
      it = tiles.erase(it);
    }
    else
      ++it;
}

DiskSequentialTileStorage::DiskSequentialTileStorage() : root(fresh_temp_dir("hmap_va_seq")) {}

DiskSequentialTileStorage::~DiskSequentialTileStorage()
{
  std::error_code ec;
  std::filesystem::remove_all(root, ec);
}

std::unique_ptr<TileStorage> DiskSequentialTileStorage::clone() const
{
  // an empty store with a directory of its own
  return std::make_unique<DiskSequentialTileStorage>();
}

Array &DiskSequentialTileStorage::get_tile(const TileRegion &region)
{
  std::lock_guard<std::mutex> lk(mutex);
  if (live)
    throw std::logic_error("hmap::VirtualArray: the disk-sequential tile storage holds one tile; "
                           "tile " + std::to_string(live_key.tx) + "," +
                           std::to_string(live_key.ty) + " is not released yet");
  Array tile(region.shape);
  read_tile(tile_file(root, region.key), tile);
  live_key = region.key;
  live = std::move(tile);
  return *live;
}

void DiskSequentialTileStorage::write_live()
{
  write_tile(tile_file(root, live_key), *live);
  live.reset();
}

void DiskSequentialTileStorage::release_tile(const TileRegion &region)
{
  std::lock_guard<std::mutex> lk(mutex);
  if (!live)
  {
    std::fprintf(stderr, "hmap::VirtualArray: disk-sequential release of tile %d,%d, none held\n",
                 region.key.tx, region.key.ty);
    return;
  }
  if (!(region.key == live_key))
    std::fprintf(stderr,
                 "hmap::VirtualArray: disk-sequential release of tile %d,%d while %d,%d is held; "
                 "the held tile is written back\n",
                 region.key.tx, region.key.ty, live_key.tx, live_key.ty);
  write_live();
}

void DiskSequentialTileStorage::trim()
{
  std::lock_guard<std::mutex> lk(mutex);
  if (live) write_live();
}

std::unique_ptr<TileStorage> make_storage(glm::ivec2 shape, glm::ivec2 tile_shape,
                                          StorageMode mode)
{
  // tile_storage.cpp:12-38
  switch (mode)
  {
  case StorageMode::VA_RAM: return std::make_unique<RamTileStorage>();
  case StorageMode::VA_DISK_LRU:
    return std::make_unique<DiskLruTileStorage>(
        (size_t)ceil_div(shape.x, tile_shape.x) * (size_t)ceil_div(shape.y, tile_shape.y));
  case StorageMode::VA_DISK_LRU_MIN: return std::make_unique<DiskLruTileStorage>(2);
  case StorageMode::VA_DISK_SEQUENTIAL: return std::make_unique<DiskSequentialTileStorage>();
  }
  throw std::invalid_argument("hmap::VirtualArray: unknown storage mode " +
                              std::to_string((int)mode));
}

std::string to_string(StorageMode m)
{
  switch (m)
  {
  case StorageMode::VA_RAM: return "VA_RAM";
  case StorageMode::VA_DISK_LRU: return "VA_DISK_LRU";
  case StorageMode::VA_DISK_LRU_MIN: return "VA_DISK_LRU_MIN";
  case StorageMode::VA_DISK_SEQUENTIAL: return "VA_DISK_SEQUENTIAL";
  }
  return "UNKNOWN";
}

std::string to_string(ForEachMode m)
{
  switch (m)
  {
  case ForEachMode::VA_SEQUENTIAL: return "VA_SEQUENTIAL";
  case ForEachMode::VA_DISTRIBUTED: return "VA_DISTRIBUTED";
  case ForEachMode::VA_SINGLE_ARRAY: return "VA_SINGLE_ARRAY";
  case ForEachMode::VA_SINGLE_ARRAY_DOWNSCALED: return "VA_SINGLE_ARRAY_DOWNSCALED";
  }
  return "UNKNOWN";
}

const std::map<std::string, int> for_each_mode_as_string = {
    {"Distributed", ForEachMode::VA_DISTRIBUTED},
    {"Sequential", ForEachMode::VA_SEQUENTIAL},
    {"Single array", ForEachMode::VA_SINGLE_ARRAY},
    {"Single array downscaled", ForEachMode::VA_SINGLE_ARRAY_DOWNSCALED}};

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

// Cell centre (global index + 0.5) scaled to the array: source position
// x = (i + 0.5) (n_a - 1) / (n - 1), split into the floor and the offset
// (virtual_array.cpp:136-226).
template <class Sample>
static void from_array_sampled(VirtualArray &va, const Array &array, const ComputeMode &cm,
                               Sample &&sample)
{
  auto lambda = [&](Array &tile, const TileRegion &region)
  {
    const float      rx = float(array.shape.x - 1) / float(va.shape.x - 1);
    const float      ry = float(array.shape.y - 1) / float(va.shape.y - 1);
    const glm::ivec2 ij0 = va.tile_region_global_indices(region);
    for (int j = 0; j < region.shape.y; ++j)
      for (int i = 0; i < region.shape.x; ++i)
      {
        const float xs = (float(ij0.x + i) + 0.5f) * rx, ys = (float(ij0.y + j) + 0.5f) * ry;
        const int   ig = int(std::floor(xs)), jg = int(std::floor(ys));
        tile(i, j) = sample(ig, jg, xs - float(ig), ys - float(jg));
      }
  };
  for_each_tile(va, lambda, cm);
}

void VirtualArray::from_array_bilinear(const Array &array, const ComputeMode &cm)
{
  // the corner clamped so that the 2x2 cells exist; offsets kept
  from_array_sampled(*this, array, cm,
                     [&array](int ig, int jg, float u, float v)
                     {
                       return array.get_value_bilinear_at(std::clamp(ig, 0, array.shape.x - 2),
                                                          std::clamp(jg, 0, array.shape.y - 2),
                                                          u, v);
                     });
}

void VirtualArray::from_array_bicubic(const Array &array, const ComputeMode &cm)
{
  // the corner clamped into the array; offsets kept
  from_array_sampled(*this, array, cm,
                     [&array](int ig, int jg, float u, float v)
                     {
                       return array.get_value_bicubic_at(std::clamp(ig, 0, array.shape.x - 1),
                                                         std::clamp(jg, 0, array.shape.y - 1),
                                                         u, v);
                     });
}

std::string VirtualArray::info_string(const ComputeMode &cm, int indent) const
{
  // the fields of virtual_array.cpp:285-309
  const std::string  pad((size_t)std::max(indent, 0), ' ');
  const glm::ivec2   nt = get_max_tiles();
  std::ostringstream o;
  o << pad << "VirtualArray @" << static_cast<const void *>(this) << "\n"
    << pad << "  shape        : " << shape.x << " x " << shape.y << "\n"
    << pad << "  tile_size    : " << tile_shape.x << " x " << tile_shape.y << "\n"
    << pad << "  tile_grid    : " << nt.x << " x " << nt.y << " (" << nt.x * nt.y << " tiles)\n"
    << pad << "  halo         : " << halo << "\n"
    << pad << "  storage      : " << (storage ? storage->info_string() : "<none>") << "\n"
    << pad << "  value_range  : [" << min(cm) << ", " << max(cm) << "]\n";
  return o.str();
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

// The single-array modes (virtual_array.inl:225-375): gather every array,
// resample it bilinearly to the working shape, run `func` once on the whole
// arrays, and put the outputs back as the mode says.
static void single_array_loop(const TileAccess &access, const TileFunc &func,
                              const ComputeMode &cm)
{
  const VirtualArray *ref = access.ref_va();
  const bool          down = cm.mode == ForEachMode::VA_SINGLE_ARRAY_DOWNSCALED;
  const ComputeMode   local{ForEachMode::VA_DISTRIBUTED, false, 1, 1.f};
  glm::ivec2          wrk = ref->shape;
  if (!down && cm.stride > 1)
    wrk = glm::ivec2(std::max(2, int(ref->shape.x / float(cm.stride))),
                     std::max(2, int(ref->shape.y / float(cm.stride))));
  if (down && cm.k_cutoff != 1.f)
    wrk = glm::ivec2(std::max(2, int(ref->shape.x * cm.k_cutoff)),
                     std::max(2, int(ref->shape.y * cm.k_cutoff)));

  const size_t       ni = access.inputs.size(), no = access.outputs.size();
  std::vector<Array> ain(ni), aout(no);
  std::vector<const Array *> in(ni, nullptr);
  std::vector<Array *>       out(no, nullptr);
  for (size_t k = 0; k < ni; ++k)
    if (access.inputs[k])
    {
      ain[k] = access.inputs[k]->to_array(local).resample_to_shape(wrk);
      in[k] = &ain[k];
    }
  for (size_t k = 0; k < no; ++k)
    if (access.outputs[k])
    {
      aout[k] = access.outputs[k]->to_array(local).resample_to_shape(wrk);
      out[k] = &aout[k];
    }

  func(in, out, TileRegion(TileKey{0, 0}, ref->bbox, wrk, glm::ivec4(0, 0, 0, 0)));

  for (size_t k = 0; k < no; ++k)
  {
    VirtualArray *va = access.outputs[k];
    if (!va) continue;
    if (down)
    {
      // coarse result at full size, plus the full array's detail above the
      // working grid's band (its own down-then-up round trip removed)
      const Array full = va->to_array(local);
      const Array low = full.resample_to_shape(wrk).resample_to_shape_bicubic(ref->shape);
      va->from_array(aout[k].resample_to_shape_bicubic(ref->shape) + (full - low), local);
    }
    else if (wrk == ref->shape)
      va->from_array(aout[k], local);
    else
      va->from_array_bicubic(aout[k], local);
  }
}

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
    // every tile taken is given back, also when a storage or func throws
    std::vector<const VirtualArray *> held;
    auto                              give_back = [&]()
    {
      for (const VirtualArray *va : held) va->storage->release_tile(region);
      held.clear();
    };
    try
    {
      std::vector<const Array *> in;
      std::vector<Array *>       out;
      for (auto *va : access.inputs)
      {
        in.push_back(va ? &va->storage->get_tile(region) : nullptr);
        if (va) held.push_back(va);
      }
      for (auto *va : access.outputs)
      {
        out.push_back(va ? &va->storage->get_tile(region) : nullptr);
        if (va) held.push_back(va);
      }
      func(in, out, region);
    }
    catch (...)
    {
      give_back();
      throw;
    }
    give_back();
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
  case ForEachMode::VA_SINGLE_ARRAY_DOWNSCALED:
    single_array_loop(access, func, cm);
    break;
  default:
    throw std::invalid_argument("for_each_tile: unknown mode " + std::to_string((int)cm.mode));
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
