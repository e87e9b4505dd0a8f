// va_driver.cpp -- exercises the drop-in's hmap::VirtualArray and writes the
// results as raw float32 (and text) into the directory given, for the tests
// to compare with the restatement in oracle/virtual_array.py.
//
//   va_driver geom DIR   tile geometry, from/to_array, reductions, overlap
//                        smoothing, std::lerp samples (no GPU call)
//   va_driver modes DIR  Array resampling, from_array_bilinear/_bicubic and
//                        the single-array modes (full, strided, downscaled)
//   va_driver storage DIR  one sequence of operations under every tile
//                        storage, and the LRU / disk / sequential rules
//   va_driver gpu DIR    hmap::gpu functions run tile by tile, as
//                        examples/ex_virtual_array does (noise over each
//                        region's bbox), in sequential and distributed modes,
//                        with RAM and disk-backed (two tiles) storage, and
//                        thermal in the downscaled single-array mode
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

#include "hmap_virtual_array.hpp"

static void dump(const std::string &dir, const std::string &name, const hmap::Array &a)
{
  const std::string path = dir + "/" + name + ".f32";
  FILE             *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(a.vector.data(), sizeof(float), a.vector.size(), f);
  std::fclose(f);
}

// a(i, j) = ((7 i + 13 j) mod 101) / 101, exact in float and in numpy
static hmap::Array pattern(glm::ivec2 s)
{
  hmap::Array a(s);
  for (int j = 0; j < s.y; j++)
    for (int i = 0; i < s.x; i++) a(i, j) = (float)((7 * i + 13 * j) % 101) / 101.f;
  return a;
}

struct Geo
{
  glm::ivec2 shape;
  glm::vec4  bbox;
  glm::ivec2 tile;
  int        halo;
};

static const Geo GEOS[] = {
    {{1024, 1024}, {0.f, 1.f, 0.f, 1.f}, {220, 312}, 32}, // ex_virtual_array
    {{1024, 1024}, {0.f, 1.f, 0.f, 1.f}, {8, 8}, 7},      // its small-tile case
    {{130, 97}, {-1.5f, 2.25f, 0.1f, 0.9f}, {40, 33}, 5},
    {{64, 48}, {0.f, 1.f, 0.f, 1.f}, {64, 48}, 3},        // one tile
    {{50, 70}, {0.f, 3.f, -1.f, 1.f}, {17, 70}, 4},       // one row of tiles
};

static int run_geom(const std::string &dir)
{
  const hmap::ComputeMode seq{hmap::VA_SEQUENTIAL};
  const hmap::ComputeMode dist{hmap::VA_DISTRIBUTED};
  FILE                   *txt = std::fopen((dir + "/geom.txt").c_str(), "w");
  int                     k = 0;
  for (const Geo &g : GEOS)
  {
    hmap::VirtualArray va(g.shape, g.bbox, g.tile, g.halo);
    std::fprintf(txt, "va %d %d %d\n", k, va.get_max_tiles().x, va.get_max_tiles().y);
    for (int ty = 0; ty < va.get_max_tiles().y; ty++)
      for (int tx = 0; tx < va.get_max_tiles().x; tx++)
      {
        const hmap::TileRegion r = va.tile_region_from_tile_coords(tx, ty);
        const glm::vec2        c = r.cell_center(1, 2), e = r.cell_corner(3, 1);
        std::fprintf(txt, "r %d %d %a %a %a %a %d %d %d %d %d %d %a %a %a %a\n", tx, ty,
                     r.bbox.x, r.bbox.y, r.bbox.z, r.bbox.w, r.shape.x, r.shape.y, r.halo.x,
                     r.halo.y, r.halo.z, r.halo.w, c.x, c.y, e.x, e.y);
      }
    const hmap::Array src = pattern(g.shape);
    va.from_array(src, dist);
    dump(dir, "same_" + std::to_string(k), va.to_array(seq));
    // several source cells share a destination cell here: sequential, so
    // that the last one in tile order is kept
    dump(dir, "half_" + std::to_string(k),
         va.to_array(glm::ivec2(g.shape.x / 2, g.shape.y / 2), seq));
    hmap::VirtualArray vb(g.shape, g.bbox, g.tile, g.halo);
    vb.from_array(pattern(glm::ivec2(g.shape.x / 3 + 2, g.shape.y / 2 + 3)), seq);
    dump(dir, "from_small_" + std::to_string(k), vb.to_array(seq));
    std::fprintf(txt, "red %d %a %a %a %a %a\n", k, va.sum(seq), va.min(dist), va.max(seq),
                 va.mean(dist), vb.sum(dist));
    const glm::vec2 pr = va.range_percentile(0.05f, 0.95f, dist);
    const glm::vec2 pr2 = va.range_percentile(0.2f, 0.7f, seq, 64);
    const std::vector<float> uv = va.unique_values(dist);
    std::fprintf(txt, "pct %d %a %a %a %a %zu %a %a\n", k, pr.x, pr.y, pr2.x, pr2.y, uv.size(),
                 uv.front(), uv.back());
    // get / set and interpolation
    va.set(g.shape.x - 1, 0, 5.f);
    std::fprintf(txt, "get %d %a %a %a %a\n", k, va.get(g.shape.x / 2, g.shape.y / 3),
                 va.get(g.shape.x - 1, 0), va.get_nearest(g.bbox.x + 0.3f * (g.bbox.y - g.bbox.x),
                                                           g.bbox.w),
                 va.get_bilinear(g.bbox.x + 0.37f * (g.bbox.y - g.bbox.x),
                                 g.bbox.z + 0.61f * (g.bbox.w - g.bbox.z)));
    // overlap smoothing on a field whose tiles disagree in their halos
    hmap::VirtualArray vs(g.shape, g.bbox, g.tile, g.halo);
    hmap::for_each_tile(
        vs,
        [](hmap::Array &t, const hmap::TileRegion &r)
        {
          for (int j = 0; j < r.shape.y; j++)
            for (int i = 0; i < r.shape.x; i++)
              t(i, j) = (float)(r.key.tx * 3 + r.key.ty * 5) + (float)((i + 2 * j) % 7) / 7.f;
        },
        dist);
    vs.smooth_overlap_buffers();
    {
      hmap::Array all; // every tile's cells in tile order, halo included
      for (int ty = 0; ty < vs.get_max_tiles().y; ty++)
        for (int tx = 0; tx < vs.get_max_tiles().x; tx++)
        {
          const hmap::Array &t = vs.storage->get_tile(vs.tile_region_from_tile_coords(tx, ty));
          all.vector.insert(all.vector.end(), t.vector.begin(), t.vector.end());
        }
      all.shape = glm::ivec2((int)all.vector.size(), 1);
      dump(dir, "smooth_tiles_" + std::to_string(k), all);
    }
    vs.remap(0.f, 1.f, seq);
    dump(dir, "smooth_remap_" + std::to_string(k), vs.to_array(dist));
    k++;
  }
  // std::lerp samples
  const float ab[][3] = {{0.3f, 0.7f, 0.25f}, {-1.f, 2.f, 0.4f}, {2.f, -1.f, 0.4f},
                         {1.5f, 1.5f, 0.9f}, {0.1f, 0.2f, 1.f},  {0.9f, 0.2f, 0.999f},
                         {-0.f, 0.f, 0.5f},  {3.f, 5.f, 1.3f},   {5.f, 3.f, -0.2f}};
  for (const auto &t : ab) std::fprintf(txt, "lerp %a %a %a %a\n", t[0], t[1], t[2], std::lerp(t[0], t[1], t[2]));
  std::fclose(txt);
  std::printf("va_driver geom ok\n");
  return 0;
}

static int run_modes(const std::string &dir)
{
  const hmap::ComputeMode seq{hmap::VA_SEQUENTIAL};
  int                     k = 0;
  for (const Geo &g : GEOS)
  {
    const std::string sk = "_" + std::to_string(k++);
    if (k == 2) continue; // the second is the first's grid in 8 x 8 tiles
    const hmap::Array src = pattern(g.shape);
    // Array resampling: down, up and to the same shape
    const glm::ivec2 down(g.shape.x / 3 + 1, g.shape.y / 2 + 5);
    const glm::ivec2 up(2 * g.shape.x - 3, g.shape.y + 7);
    dump(dir, "rs_lin_down" + sk, src.resample_to_shape(down));
    dump(dir, "rs_lin_up" + sk, src.resample_to_shape_bilinear(up));
    dump(dir, "rs_lin_same" + sk, src.resample_to_shape(g.shape));
    dump(dir, "rs_cub_down" + sk, src.resample_to_shape_bicubic(down));
    dump(dir, "rs_cub_up" + sk, src.resample_to_shape_bicubic(up));
    // from_array with bilinear and bicubic samples
    const hmap::Array  small = pattern(glm::ivec2(g.shape.x / 3 + 2, g.shape.y / 2 + 3));
    hmap::VirtualArray va(g.shape, g.bbox, g.tile, g.halo);
    va.from_array_bicubic(small, seq);
    dump(dir, "from_cub" + sk, va.to_array(seq));
    va.from_array_bilinear(small, seq);
    dump(dir, "from_lin" + sk, va.to_array(seq));
    // the single-array modes, each from the pattern
    auto single = [&](const hmap::ComputeMode &cm, const std::string &name, auto &&f)
    {
      hmap::VirtualArray vc(g.shape, g.bbox, g.tile, g.halo);
      vc.from_array(src, seq);
      hmap::for_each_tile(vc, f, cm);
      dump(dir, name + sk, vc.to_array(seq));
    };
    auto add_index = [](hmap::Array &a, const hmap::TileRegion &r)
    {
      for (int j = 0; j < r.shape.y; j++)
        for (int i = 0; i < r.shape.x; i++) a(i, j) = a(i, j) + (float)(i * r.shape.y + j);
    };
    auto halve = [](hmap::Array &a, const hmap::TileRegion &)
    {
      for (float &x : a.vector) x = x * 0.5f;
    };
    single(hmap::ComputeMode{hmap::VA_SINGLE_ARRAY}, "single", add_index);
    single(hmap::ComputeMode{hmap::VA_SINGLE_ARRAY, false, 3}, "single_s3", add_index);
    single(hmap::ComputeMode{hmap::VA_SINGLE_ARRAY_DOWNSCALED, false, 1, 0.5f}, "down_05", halve);
    single(hmap::ComputeMode{hmap::VA_SINGLE_ARRAY_DOWNSCALED, false, 1, 1.f}, "down_1", halve);
  }
  std::printf("va_driver modes ok\n");
  return 0;
}

static size_t files_in(const std::filesystem::path &d)
{
  size_t n = 0;
  for (const auto &e : std::filesystem::directory_iterator(d))
    if (e.is_regular_file()) n++;
  return n;
}

static int run_storage(const std::string &dir)
{
  const hmap::ComputeMode seq{hmap::VA_SEQUENTIAL};
  const hmap::ComputeMode dist{hmap::VA_DISTRIBUTED};
  const Geo              &g = GEOS[2]; // 130 x 97 in 4 x 3 tiles of 40 x 33, halo 5
  const hmap::Array       src = pattern(g.shape);
  FILE                   *txt = std::fopen((dir + "/storage.txt").c_str(), "w");
  const hmap::StorageMode modes[] = {hmap::VA_RAM, hmap::VA_DISK_LRU, hmap::VA_DISK_LRU_MIN,
                                     hmap::VA_DISK_SEQUENTIAL};
  for (hmap::StorageMode m : modes)
  {
    const std::string       name = hmap::to_string(m);
    const bool              one = m == hmap::VA_DISK_SEQUENTIAL;
    const hmap::ComputeMode cm = one ? seq : dist; // one tile at a time: sequential loops only
    hmap::VirtualArray      va(g.shape, g.bbox, g.tile, g.halo, m);
    hmap::VirtualArray      vb(g.shape, g.bbox, g.tile, g.halo, m);
    va.from_array(src, cm);
    hmap::for_each_tile(
        va,
        [](hmap::Array &t, const hmap::TileRegion &r)
        {
          for (int j = 0; j < r.shape.y; j++)
            for (int i = 0; i < r.shape.x; i++)
              t(i, j) = t(i, j) + (float)(r.key.tx * 3 + r.key.ty * 5) + (float)((i + 2 * j) % 7) / 7.f;
        },
        cm);
    va.smooth_overlap_buffers(); // skipped, with a message, by the sequential store
    va.remap(0.f, 1.f, seq);
    va.set(7, 9, 2.5f);
    hmap::for_each_tile(
        std::vector<const hmap::VirtualArray *>{&va}, std::vector<hmap::VirtualArray *>{&vb},
        [](const std::vector<const hmap::Array *> &in, std::vector<hmap::Array *> &out,
           const hmap::TileRegion &)
        {
          for (size_t c = 0; c < in[0]->vector.size(); c++)
            out[0]->vector[c] = 2.f * in[0]->vector[c] + 1.f;
        },
        hmap::ComputeMode{cm.mode, true}); // trim both stores after the loop
    if (auto *d = dynamic_cast<hmap::DiskLruTileStorage *>(vb.storage.get()))
    {
      // trimmed: nothing resident, every tile on disk
      std::fprintf(txt, "trim %s %zu %zu\n", name.c_str(), d->resident_tiles(),
                   files_in(d->directory()));
    }
    dump(dir, "st_a_" + name, va.to_array(seq));
    dump(dir, "st_b_" + name, vb.to_array(cm));
    const float got = va.get(7, 9), sum = vb.sum(seq);
    std::fprintf(txt, "mode %s %zu %a %a |%s\n", name.c_str(), va.storage->max_live_tiles(), got,
                 sum, va.storage->info_string().c_str());
  }
  // the rules of the stores, on 40 x 33 tiles
  hmap::VirtualArray geo(g.shape, g.bbox, g.tile, g.halo);
  auto               reg = [&](int tx, int ty) { return geo.tile_region_from_tile_coords(tx, ty); };
  auto               fill_release = [&](hmap::TileStorage &st, int tx, int ty, float v)
  {
    hmap::Array &t = st.get_tile(reg(tx, ty));
    std::fill(t.vector.begin(), t.vector.end(), v);
    st.release_tile(reg(tx, ty));
  };
  auto read_release = [&](hmap::TileStorage &st, int tx, int ty)
  {
    const float s = st.get_tile(reg(tx, ty)).sum();
    st.release_tile(reg(tx, ty));
    return s;
  };
  {
    // RAM LRU of two: the first tile leaves when the third comes, and is
    // zero again when asked for later
    hmap::LruTileStorage lru(2);
    fill_release(lru, 0, 0, 1.f);
    fill_release(lru, 1, 0, 2.f);
    fill_release(lru, 2, 0, 3.f);
    const size_t res = lru.resident_tiles();
    const float  s0 = read_release(lru, 0, 0), s2 = read_release(lru, 2, 0);
    std::fprintf(txt, "lru %zu %a %a\n", res, s0, s2);
  }
  std::filesystem::path gone;
  {
    // disk LRU of two: the same tile comes back from disk
    hmap::DiskLruTileStorage dl(2);
    gone = dl.directory();
    fill_release(dl, 0, 0, 1.f);
    fill_release(dl, 1, 0, 2.f);
    fill_release(dl, 2, 0, 3.f);
    const size_t res = dl.resident_tiles(), on_disk = files_in(dl.directory());
    const float  s0 = read_release(dl, 0, 0), s1 = read_release(dl, 1, 0);
    const float  s2 = read_release(dl, 2, 0);
    std::fprintf(txt, "disklru %zu %zu %a %a %a\n", res, on_disk, s0, s1, s2);
    // tiles in use are never evicted: three held at once in a store of two
    hmap::Array &a = dl.get_tile(reg(0, 1));
    hmap::Array &b = dl.get_tile(reg(1, 1));
    hmap::Array &c = dl.get_tile(reg(2, 1));
    a.vector[0] = 4.f;
    b.vector[0] = 5.f;
    c.vector[0] = 6.f;
    const size_t held = dl.resident_tiles();
    dl.release_tile(reg(0, 1));
    dl.release_tile(reg(1, 1));
    dl.release_tile(reg(2, 1));
    fill_release(dl, 3, 1, 7.f); // room is made down to the capacity
    const size_t after = dl.resident_tiles();
    const float  v0 = dl.get_tile(reg(0, 1)).vector[0];
    const float  v1 = dl.get_tile(reg(1, 1)).vector[0];
    const float  v2 = dl.get_tile(reg(2, 1)).vector[0];
    std::fprintf(txt, "held %zu %zu %a %a %a\n", held, after, v0, v1, v2);
  }
  std::fprintf(txt, "removed %d\n", (int)!std::filesystem::exists(gone));
  {
    // one tile at a time in the sequential store
    hmap::DiskSequentialTileStorage ds;
    fill_release(ds, 1, 1, 2.f);
    hmap::Array &t = ds.get_tile(reg(0, 0));
    int          refused = 0;
    try
    {
      ds.get_tile(reg(1, 1));
    }
    catch (const std::logic_error &)
    {
      refused = 1;
    }
    t.vector[0] = 9.f;
    ds.release_tile(reg(0, 0));
    const size_t on_disk = files_in(ds.directory());
    const float  s11 = read_release(ds, 1, 1), s00 = read_release(ds, 0, 0);
    std::fprintf(txt, "seq %d %zu %a %a\n", refused, on_disk, s11, s00);
  }
  std::fclose(txt);
  std::printf("va_driver storage ok\n");
  return 0;
}

static int run_gpu(const std::string &dir)
{
  if (!hmap::gpu::init_opencl())
  {
    std::fprintf(stderr, "init_opencl failed\n");
    return 1;
  }
  const glm::ivec2 shape{1024, 1024};
  const glm::vec4  bbox{0.f, 1.f, 0.f, 1.f};
  for (int m = 0; m < 3; m++)
  {
    // sequential, distributed, and distributed over the disk-backed store
    // that keeps two tiles in memory
    const hmap::ComputeMode cm{m == 0 ? hmap::VA_SEQUENTIAL : hmap::VA_DISTRIBUTED};
    const std::string       sfx = m == 0 ? "_seq" : m == 1 ? "_dist" : "_disk";
    hmap::VirtualArray      va(shape, bbox, glm::ivec2(220, 312), 32,
                               m == 2 ? hmap::VA_DISK_LRU_MIN : hmap::VA_RAM);
    // ex_virtual_array's generate_noise, on the GPU
    hmap::for_each_tile(
        va,
        [](hmap::Array &tile, const hmap::TileRegion &region)
        {
          tile = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, region.shape, {2.f, 2.f}, 0, 8,
                                      0.7f, 0.5f, 2.f, nullptr, nullptr, nullptr, nullptr,
                                      region.bbox);
        },
        cm);
    dump(dir, "noise" + sfx, va.to_array(cm));
    hmap::for_each_tile(
        va, [](hmap::Array &tile, const hmap::TileRegion &)
        { hmap::gpu::thermal(tile, 2.f / 1024.f, 20); },
        cm);
    dump(dir, "thermal" + sfx, va.to_array(cm));
    hmap::for_each_tile(
        va, [](hmap::Array &tile, const hmap::TileRegion &)
        { hmap::gpu::hydraulic_vpipes(tile, 1e-3f, true, 0.01f, 20); },
        cm);
    dump(dir, "vpipes" + sfx, va.to_array(cm));
    va.smooth_overlap_buffers();
    va.remap(0.f, 1.f, cm);
    dump(dir, "final" + sfx, va.to_array(cm));
  }
  {
    // thermal on the grid downscaled by 2, its result upsampled with the
    // full grid's finer detail added back
    hmap::VirtualArray va(shape, bbox, glm::ivec2(220, 312), 32);
    hmap::for_each_tile(
        va,
        [](hmap::Array &tile, const hmap::TileRegion &region)
        {
          tile = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, region.shape, {2.f, 2.f}, 0, 8,
                                      0.7f, 0.5f, 2.f, nullptr, nullptr, nullptr, nullptr,
                                      region.bbox);
        },
        hmap::ComputeMode{hmap::VA_DISTRIBUTED});
    hmap::for_each_tile(
        va, [](hmap::Array &a, const hmap::TileRegion &) { hmap::gpu::thermal(a, 2.f / 512.f, 20); },
        hmap::ComputeMode{hmap::VA_SINGLE_ARRAY_DOWNSCALED, false, 1, 0.5f});
    dump(dir, "thermal_downscaled", va.to_array(hmap::ComputeMode{hmap::VA_SEQUENTIAL}));
  }
  std::printf("va_driver gpu ok\n");
  return 0;
}

int main(int argc, char **argv)
{
  if (argc < 3)
  {
    std::fprintf(stderr, "usage: %s geom|modes|storage|gpu out_dir\n", argv[0]);
    return 2;
  }
  if (!std::strcmp(argv[1], "geom")) return run_geom(argv[2]);
  if (!std::strcmp(argv[1], "modes")) return run_modes(argv[2]);
  if (!std::strcmp(argv[1], "storage")) return run_storage(argv[2]);
  if (!std::strcmp(argv[1], "gpu")) return run_gpu(argv[2]);
  return 2;
}
