// va_bench.cpp -- times hmap::gpu::hydraulic_vpipes run tile by tile over a
// hmap::VirtualArray (the out-of-core pattern of HighMap's VirtualArray:
// host-resident tiles with halos, each sent through the GPU by the drop-in
// call), in VA_SEQUENTIAL and VA_DISTRIBUTED modes.  Prints one JSON line.
//
//   va_bench N TILE HALO ITERATIONS
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>

#include "hmap_virtual_array.hpp"

int main(int argc, char **argv)
{
  if (argc < 5)
  {
    std::fprintf(stderr, "usage: %s N TILE HALO ITERATIONS\n", argv[0]);
    return 2;
  }
  const int n = std::atoi(argv[1]), tile = std::atoi(argv[2]), halo = std::atoi(argv[3]);
  const int iters = std::atoi(argv[4]);
  if (!hmap::gpu::init_opencl())
  {
    std::fprintf(stderr, "no GPU\n");
    return 1;
  }
  const glm::ivec2 shape{n, n};
  const glm::vec4  bbox{0.f, 1.f, 0.f, 1.f};
  const auto       noise = [](hmap::Array &t, const hmap::TileRegion &r)
  {
    t = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, r.shape, {4.f, 4.f}, 2, 8, 0.7f, 0.5f, 2.f,
                             nullptr, nullptr, nullptr, nullptr, r.bbox);
    hmap::remap(t);
  };
  const auto erode = [iters](hmap::Array &t, const hmap::TileRegion &)
  { hmap::gpu::hydraulic_vpipes(t, 1e-3f, true, 0.01f, iters, 0.5f, 0.1f, 0.05f, 0.05f); };

  double       secs[2] = {0.0, 0.0};
  hmap::Array  res[2];
  long long    tile_cells = 0;
  const hmap::ComputeMode modes[2] = {{hmap::VA_SEQUENTIAL}, {hmap::VA_DISTRIBUTED}};
  for (int m = 0; m < 2; m++)
  {
    hmap::VirtualArray va(shape, bbox, glm::ivec2(tile, tile), halo);
    hmap::for_each_tile(va, noise, modes[1]);
    // warm-up on one tile's worth (buffer pools, code objects)
    {
      hmap::Array w = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, {tile, tile}, {4.f, 4.f}, 2);
      hmap::gpu::hydraulic_vpipes(w, 1e-3f, true, 0.01f, 2);
    }
    const auto t0 = std::chrono::steady_clock::now();
    hmap::for_each_tile(va, erode, modes[m]);
    secs[m] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res[m] = va.to_array(modes[1]);
    tile_cells = 0;
    for (int ty = 0; ty < va.get_max_tiles().y; ty++)
      for (int tx = 0; tx < va.get_max_tiles().x; tx++)
      {
        const hmap::TileRegion r = va.tile_region_from_tile_coords(tx, ty);
        tile_cells += (long long)r.shape.x * r.shape.y;
      }
    std::fprintf(stderr, "va_bench: mode %d done in %.3f s\n", m, secs[m]);
  }
  const bool   same = res[0].vector == res[1].vector;
  const double cells = (double)n * n;
  std::printf("{\"n\": %d, \"tile\": %d, \"halo\": %d, \"iterations\": %d, \"tiles\": %d, "
              "\"tile_cells\": %lld, \"seq_s\": %.4f, \"dist_s\": %.4f, "
              "\"seq_gcell_iter_s\": %.3f, \"dist_gcell_iter_s\": %.3f, "
              "\"dist_tile_gcell_iter_s\": %.3f, \"modes_equal\": %s, \"host_threads\": %u}\n",
              n, tile, halo, iters, (int)((n + tile - 1) / tile) * ((n + tile - 1) / tile),
              tile_cells, secs[0], secs[1], cells * iters / secs[0] / 1e9,
              cells * iters / secs[1] / 1e9, (double)tile_cells * iters / secs[1] / 1e9,
              same ? "true" : "false", std::thread::hardware_concurrency());
  return same ? 0 : 1;
}
