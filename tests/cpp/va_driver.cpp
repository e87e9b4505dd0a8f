// va_driver.cpp -- exercises the drop-in's hmap::VirtualArray and writes the
// results as raw float32 (and text) into the directory given, for the tests
// to compare with the restatement in oracle/virtual_array.py.
//
//   va_driver geom DIR   tile geometry, from/to_array, reductions, overlap
//                        smoothing, std::lerp samples (no GPU call)
//   va_driver gpu DIR    hmap::gpu functions run tile by tile, as
//                        examples/ex_virtual_array does (noise over each
//                        region's bbox), in sequential and distributed modes
#include <cmath>
#include <cstdio>
#include <cstring>
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
  const hmap::ComputeMode single{hmap::VA_SINGLE_ARRAY};
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
    // the single-array mode sees the whole array at once
    hmap::VirtualArray vc(g.shape, g.bbox, g.tile, g.halo);
    vc.from_array(src, seq);
    hmap::for_each_tile(
        vc,
        [](hmap::Array &a, const hmap::TileRegion &r)
        {
          for (int j = 0; j < r.shape.y; j++)
            for (int i = 0; i < r.shape.x; i++) a(i, j) = a(i, j) + (float)(i * r.shape.y + j);
        },
        single);
    dump(dir, "single_" + std::to_string(k), vc.to_array(seq));
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

static int run_gpu(const std::string &dir)
{
  if (!hmap::gpu::init_opencl())
  {
    std::fprintf(stderr, "init_opencl failed\n");
    return 1;
  }
  const glm::ivec2 shape{1024, 1024};
  const glm::vec4  bbox{0.f, 1.f, 0.f, 1.f};
  for (int m = 0; m < 2; m++)
  {
    const hmap::ComputeMode cm{m == 0 ? hmap::VA_SEQUENTIAL : hmap::VA_DISTRIBUTED};
    const std::string       sfx = m == 0 ? "_seq" : "_dist";
    hmap::VirtualArray      va(shape, bbox, glm::ivec2(220, 312), 32);
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
  std::printf("va_driver gpu ok\n");
  return 0;
}

int main(int argc, char **argv)
{
  if (argc < 3)
  {
    std::fprintf(stderr, "usage: %s geom|gpu out_dir\n", argv[0]);
    return 2;
  }
  if (!std::strcmp(argv[1], "geom")) return run_geom(argv[2]);
  if (!std::strcmp(argv[1], "gpu")) return run_gpu(argv[2]);
  return 2;
}
