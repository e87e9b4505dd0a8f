// hmap_virtual_array.hpp -- hmap::VirtualArray for the drop-in: a grid held
// as tiles with overlapping halos, processed tile by tile, so that grids
// larger than one device (or one call) run through the hmap::gpu functions
// a tile at a time.
//
// Same names, geometry and semantics as the reference:
//   TileKey, TileRegion        HighMap/include/highmap/virtual_array/tile_region.hpp:21-53
//   TileStorage, StorageMode   HighMap/include/highmap/virtual_array/tile_storage.hpp
//   ComputeMode, ForEachMode,
//   VirtualArray               HighMap/include/highmap/virtual_array/virtual_array.hpp:27-165
//   for_each_tile, for_each_cell HighMap/include/highmap/virtual_array/virtual_array.inl
//
// Provided: every tile storage (RAM, RAM LRU, disk-backed LRU with the
// capacity of the whole grid or of two tiles, disk-sequential) and every
// compute mode (sequential, distributed, single array at full size or
// strided, single array downscaled).
//
// Tiles of one for_each_tile call in VA_DISTRIBUTED mode run on separate
// host threads; the hmap::gpu functions give each calling thread its own
// device context (stream and buffer pool), so those tiles' uploads, kernels
// and downloads overlap on the GPU.  The LRU storages never evict a tile
// that is between get_tile and release_tile, so they are safe in that mode
// too (the store holds more tiles than its capacity while every resident
// tile is in use).
#pragma once

#include <cstdint>
#include <filesystem>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <unordered_map>
#include <vector>

#include "hmap_gpu.hpp"

namespace hmap
{

int ceil_div(int a, int b);

struct TileKey
{
  int tx;
  int ty;

  bool operator==(const TileKey &other) const { return tx == other.tx && ty == other.ty; }
};

struct TileKeyHash
{
  size_t operator()(const TileKey &k) const
  {
    return (static_cast<size_t>(static_cast<uint32_t>(k.tx)) << 32) ^
           static_cast<size_t>(static_cast<uint32_t>(k.ty));
  }
};

struct TileRegion
{
  TileKey    key;   // tile coordinates
  glm::vec4  bbox;  // (x1, x2, y1, y2) of the tile, halo included
  glm::ivec2 shape; // cells of the tile, halo included
  glm::ivec4 halo;  // halo cells on the (left, right, bottom, top) side

  TileRegion(const TileKey    &key,
             const glm::vec4  &bbox,
             const glm::ivec2 &shape,
             const glm::ivec4 &halo = glm::ivec4(0, 0, 0, 0));

  glm::vec2   cell_center(int i, int j) const;
  glm::vec2   cell_corner(int i, int j) const;
  std::string key_as_string() const;
};

enum StorageMode : int
{
  VA_RAM,
  VA_DISK_LRU,       // disk-backed LRU, capacity the number of tiles
  VA_DISK_LRU_MIN,   // disk-backed LRU, capacity two tiles
  VA_DISK_SEQUENTIAL // one tile in memory; no overlap smoothing
};

std::string to_string(StorageMode m);

class TileStorage
{
public:
  virtual ~TileStorage() = default;
  virtual Array                       &get_tile(const TileRegion &region) = 0;
  virtual void                         release_tile(const TileRegion &region) = 0;
  virtual size_t                       max_live_tiles() const = 0;
  virtual std::unique_ptr<TileStorage> clone() const = 0;
  virtual std::string                  info_string() const = 0;
  // drop what memory can be dropped, keeping the data
  virtual void trim() {}
};

// Every tile in host memory, created (zero-filled, halo included) on first
// access.  Safe for concurrent access to distinct tiles.
// (HighMap/src/virtual_array/ram_tile_storage.cpp)
class RamTileStorage : public TileStorage
{
public:
  RamTileStorage() = default;
  RamTileStorage(const RamTileStorage &other);

  Array                       &get_tile(const TileRegion &region) override;
  void                         release_tile(const TileRegion &region) override;
  size_t                       max_live_tiles() const override;
  std::unique_ptr<TileStorage> clone() const override;
  std::string                  info_string() const override { return "RAM"; }

private:
  std::unordered_map<TileKey, Array, TileKeyHash> tiles;
  std::mutex                                      mutex;
};

// At most `max_tiles` tiles in memory (more only while every resident tile
// is in use); the least recently used tile not in use leaves when another
// is needed.  This base class drops the tile that leaves: a later access
// finds it zero-filled again (HighMap/src/virtual_array/
// lru_tile_storage.cpp).  Thread-safe.
class LruTileStorage : public TileStorage
{
public:
  explicit LruTileStorage(size_t max_tiles);

  Array                       &get_tile(const TileRegion &region) override;
  void                         release_tile(const TileRegion &region) override;
  size_t                       max_live_tiles() const override;
  std::unique_ptr<TileStorage> clone() const override;
  std::string                  info_string() const override;

  size_t resident_tiles() const; // tiles in memory now

protected:
  struct Entry
  {
    Array         tile;
    int           in_use = 0;   // get_tile calls not yet released
    std::uint64_t last_use = 0; // access clock
  };

  // With the lock held: a tile leaves memory / a new tile is created (return
  // false to leave it zero-filled).
  virtual void on_evict(const TileKey &key, const Array &tile);
  virtual bool on_create(const TileRegion &region, Array &tile);

  size_t                                          max_tiles;
  std::unordered_map<TileKey, Entry, TileKeyHash> tiles;
  std::uint64_t                                   clock = 0;
  mutable std::mutex                              mutex;
};

// The LRU store with the tiles that leave memory written to a private
// directory under the system's temporary directory (one raw float32 file
// per tile) and read back on the next access; trim() writes every resident
// tile out and frees it.  The directory is removed with the store.
// (HighMap/src/virtual_array/disk_lru_storage.cpp)
class DiskLruTileStorage : public LruTileStorage
{
public:
  explicit DiskLruTileStorage(size_t max_tiles);
  ~DiskLruTileStorage() override;

  std::unique_ptr<TileStorage> clone() const override;
  std::string                  info_string() const override;
  void                         trim() override;

  const std::filesystem::path &directory() const { return root; }

protected:
  void on_evict(const TileKey &key, const Array &tile) override;
  bool on_create(const TileRegion &region, Array &tile) override;

private:
  std::filesystem::path root;
};

// One tile in memory at a time: get_tile reads it from disk (zero-filled
// the first time), release_tile writes it back and frees it.  A second
// get_tile before the release throws std::logic_error, so only sequential
// loops (and single-cell get/set) may use it; overlap smoothing, which needs
// two tiles at once, is skipped with a message.
// (HighMap/src/virtual_array/disk_sequential_storage.cpp)
class DiskSequentialTileStorage : public TileStorage
{
public:
  DiskSequentialTileStorage();
  ~DiskSequentialTileStorage() override;

  Array                       &get_tile(const TileRegion &region) override;
  void                         release_tile(const TileRegion &region) override;
  size_t                       max_live_tiles() const override { return 1; }
  std::unique_ptr<TileStorage> clone() const override;
  std::string                  info_string() const override { return "DiskSequential"; }
  void                         trim() override;

  const std::filesystem::path &directory() const { return root; }

private:
  std::filesystem::path root;
  std::mutex            mutex;
  TileKey               live_key{0, 0};
  std::optional<Array>  live;

  void write_live();
};

std::unique_ptr<TileStorage> make_storage(glm::ivec2 shape, glm::ivec2 tile_shape,
                                          StorageMode mode);

enum ForEachMode : int
{
  VA_SEQUENTIAL,             // tile by tile, one thread
  VA_DISTRIBUTED,            // tile by tile, hardware_concurrency threads
  VA_SINGLE_ARRAY,           // the whole array at once (every `stride`-th cell)
  VA_SINGLE_ARRAY_DOWNSCALED // the whole array at once, downscaled by k_cutoff
};

std::string to_string(ForEachMode m);

// names of the modes as a user interface lists them
// (HighMap/include/highmap/virtual_array/virtual_array.hpp:50-54)
extern const std::map<std::string, int> for_each_mode_as_string;

struct ComputeMode
{
  ForEachMode mode;
  bool        trim_storage = false;
  int         stride = 1;     // VA_SINGLE_ARRAY: shape / stride cells (at least 2)
  float       k_cutoff = 1.f; // VA_SINGLE_ARRAY_DOWNSCALED: shape * k_cutoff cells
};

struct VirtualArray
{
  VirtualArray() = default;
  VirtualArray(glm::ivec2                   shape,
               glm::vec4                    bbox,
               glm::ivec2                   tile_shape,
               int                          halo,
               std::unique_ptr<TileStorage> storage);
  VirtualArray(glm::ivec2  shape,
               glm::vec4   bbox,
               glm::ivec2  tile_shape,
               int         halo,
               StorageMode storage_mode = StorageMode::VA_RAM);
  VirtualArray(glm::ivec2 shape, glm::ivec2 tile_shape, int halo,
               StorageMode storage_mode = StorageMode::VA_RAM);

  std::unique_ptr<VirtualArray> clone(const ComputeMode &cm, bool deep_copy = false);
  void copy_from(VirtualArray &src, const ComputeMode &cm, bool copy_src_data = true);

  float      get(int global_i, int global_j) const;
  float      get_bilinear(float x, float y) const;
  float      get_nearest(float x, float y) const;
  void       set(int global_i, int global_j, float v);
  glm::ivec2 get_max_tiles() const;
  int        get_ntiles() const;

  std::string info_string(const ComputeMode &cm, int indent = 0) const;

  void  fill(float value, const ComputeMode &cm);
  // Every cell (halo included) from the array scaled to this grid: the
  // nearest cell, or the bilinear / bicubic sample at the cell centre.
  void  from_array(const Array &array, const ComputeMode &cm);
  void  from_array_bilinear(const Array &array, const ComputeMode &cm);
  void  from_array_bicubic(const Array &array, const ComputeMode &cm);
  Array to_array(const glm::ivec2 &array_shape, const ComputeMode &cm) const;
  Array to_array(const ComputeMode &cm) const;
  Array to_array_dbg() const;

  float     max(const ComputeMode &cm) const;
  float     mean(const ComputeMode &cm) const;
  float     min(const ComputeMode &cm) const;
  float     sum(const ComputeMode &cm) const;
  glm::vec2 range(const ComputeMode &cm) const;
  glm::vec2 range_percentile(float p_low, float p_high, const ComputeMode &cm,
                             size_t bins = 1024) const;
  std::vector<float> unique_values(const ComputeMode &cm) const;

  void inverse(const ComputeMode &cm);
  void remap(float vmin, float vmax, const ComputeMode &cm);
  void remap(float vmin, float vmax, float from_min, float from_max, const ComputeMode &cm);

  void smooth_overlap_buffers();

  glm::vec2  tile_region_global_position(const TileRegion &region) const;
  glm::ivec2 tile_region_global_indices(const TileRegion &region) const;
  TileRegion tile_region_from_global_index(int global_i, int global_j) const;
  TileRegion tile_region_from_tile_coords(int tile_x, int tile_y) const;
  glm::ivec2 local_indices(const TileRegion &region, int global_i, int global_j) const;

  void trim_storage() const;

  glm::ivec2                   shape;
  glm::vec4                    bbox; // (x1, x2, y1, y2)
  glm::ivec2                   tile_shape;
  int                          halo = 0;
  std::unique_ptr<TileStorage> storage;
};

void copy_data(VirtualArray &src, VirtualArray &dst, const ComputeMode &cm);

// ------------------------------------------------------------ tile loops

struct TileAccess
{
  std::vector<const VirtualArray *> inputs;
  std::vector<VirtualArray *>       outputs;

  const VirtualArray *ref_va() const
  {
    if (!outputs.empty() && outputs.front()) return outputs.front();
    return inputs.empty() ? nullptr : inputs.front();
  }
};

using TileFunc = std::function<void(const std::vector<const Array *> &,
                                    std::vector<Array *> &,
                                    const TileRegion &)>;

// The general loop (HighMap/include/highmap/virtual_array/virtual_array.inl).
// Tile modes: `func` receives the input and output tiles of one region
// (halo included).  Single-array modes: `func` receives whole arrays of a
// working shape and a region covering the grid's bbox without halo:
//   VA_SINGLE_ARRAY             working shape max(2, shape / stride) per
//       axis (the shape itself at stride 1); every array is gathered and
//       resampled bilinearly to it (at stride 1 too, as the reference
//       does, which shifts cells by up to half a cell); the outputs come
//       back cell for cell, or bicubically when strided;
//   VA_SINGLE_ARRAY_DOWNSCALED  working shape max(2, shape * k_cutoff) per
//       axis (the shape at k_cutoff 1); the outputs come back as the
//       working result upsampled bicubically plus the detail the working
//       grid cannot hold: full - bicubic_up(bilinear_down(full)).
// Throws std::invalid_argument for arrays of other shapes or tile shapes.
void for_each_tile(const TileAccess &access, const TileFunc &func, const ComputeMode &cm);

template <typename Func>
void for_each_tile(VirtualArray &va, Func &&func, const ComputeMode &cm)
{
  TileAccess acc;
  acc.outputs = {&va};
  for_each_tile(
      acc,
      [&](const std::vector<const Array *> &, std::vector<Array *> &out,
          const TileRegion &region) { func(*out[0], region); },
      cm);
}

template <typename Func>
void for_each_tile(const VirtualArray &va, Func &&func, const ComputeMode &cm)
{
  TileAccess acc;
  acc.outputs = {const_cast<VirtualArray *>(&va)};
  for_each_tile(
      acc,
      [&](const std::vector<const Array *> &, std::vector<Array *> &out,
          const TileRegion &region) { func(*out[0], region); },
      cm);
}

template <typename Func>
void for_each_tile(const std::vector<VirtualArray *> &outputs, Func &&func,
                   const ComputeMode &cm)
{
  TileAccess acc;
  acc.outputs = outputs;
  for_each_tile(
      acc,
      [&](const std::vector<const Array *> &, std::vector<Array *> &out,
          const TileRegion &region) { func(out, region); },
      cm);
}

template <typename Func>
void for_each_tile(const std::vector<const VirtualArray *> &inputs,
                   const std::vector<VirtualArray *>       &outputs,
                   Func                                   &&func,
                   const ComputeMode                       &cm)
{
  TileAccess acc;
  acc.inputs = inputs;
  acc.outputs = outputs;
  for_each_tile(
      acc,
      [&](const std::vector<const Array *> &in, std::vector<Array *> &out,
          const TileRegion &region) { func(in, out, region); },
      cm);
}

template <typename Func>
void for_each_cell(VirtualArray &va, Func &&func, const ComputeMode &cm)
{
  for_each_tile(
      va,
      [&](Array &tile, const TileRegion &region)
      {
        for (int j = 0; j < region.shape.y; ++j)
          for (int i = 0; i < region.shape.x; ++i) func(tile(i, j), i, j, region);
      },
      cm);
}

// hmap::remap with an explicit source range (HighMap/src/range/range.cpp:
// 363-377).
void remap(Array &array, float vmin, float vmax, float from_min, float from_max);

} // namespace hmap
