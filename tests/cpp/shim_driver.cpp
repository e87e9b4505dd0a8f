// shim_driver.cpp -- calls the hmap::gpu drop-in exactly as HighMap's own
// examples do (examples/ex_thermal/ex_thermal.cpp:7-16,
// examples/ex_hydraulic_vpipes_gpu/ex_hydraulic_vpipes_gpu.cpp:7-22,
// examples/ex_flow_simulation/ex_flow_simulation.cpp:5-26,
// examples/ex_thermal_auto_bedrock/ex_thermal_auto_bedrock.cpp:5-15,
// examples/ex_sediment_deposition/ex_sediment_deposition.cpp:5-18,
// examples/ex_snow_simulation/ex_snow_simulation.cpp:5-38) and
// writes every result as raw float32 into the directory given, for the GPU
// test to compare with the oracle.
#include <cstdio>
#include <string>

#include "hmap_gpu.hpp"

static void dump(const std::string &dir, const char *name, const hmap::Array &a)
{
  std::string path = dir + "/" + name + ".f32";
  FILE       *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(a.vector.data(), sizeof(float), a.vector.size(), f);
  std::fclose(f);
}

int main(int argc, char **argv)
{
  if (argc < 2)
  {
    std::fprintf(stderr, "usage: %s out_dir\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  if (!hmap::gpu::init_opencl())
  {
    std::fprintf(stderr, "init_opencl failed\n");
    return 1;
  }

  // thermal, as ex_thermal: noise_fbm(PERLIN, {256,256}, {4,4}, seed 1)
  glm::ivec2  shape = {256, 256};
  hmap::Array z = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, shape, {4.f, 4.f}, 1);
  dump(dir, "noise256", z);
  hmap::remap(z);
  hmap::Array z0 = z;
  hmap::Array dep;
  hmap::gpu::thermal(z, 2.f / 256.f, 50, nullptr, &dep);
  dump(dir, "thermal256", z);
  dump(dir, "deposition256", dep);

  // masked overload with a talus plane
  hmap::Array zm = z0, mask(shape, 0.f), talus(shape, 1.f / 256.f);
  for (int j = 0; j < shape.y; j++)
    for (int i = 0; i < shape.x; i++) mask(i, j) = (float)i / (float)(shape.x - 1);
  hmap::gpu::thermal(zm, &mask, talus, 7);
  dump(dir, "thermal_masked256", zm);

  // vpipes, as ex_hydraulic_vpipes_gpu with the 1024^2 workload's
  // parameters on a 192 x 128 map (seed 2)
  glm::ivec2  s2 = {192, 128};
  hmap::Array h = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, s2, {4.f, 4.f}, 2, 8, 0.7f);
  hmap::remap(h);
  dump(dir, "vp_in", h);
  hmap::Array d, sed, u, v;
  hmap::gpu::hydraulic_vpipes(h, 1e-3f, true, 0.01f, 40, 0.5f, 0.1f, 0.05f, 0.05f, 1.f, 10.f,
                              true, 0.01f, nullptr, &d, &sed, &u, &v);
  dump(dir, "vp_z", h);
  dump(dir, "vp_d", d);
  dump(dir, "vp_s", sed);
  dump(dir, "vp_u", u);
  dump(dir, "vp_v", v);

  // reference defaults on the same input
  hmap::Array h2(s2);
  {
    FILE *f = std::fopen((dir + "/vp_in.f32").c_str(), "rb");
    if (!f || std::fread(h2.vector.data(), sizeof(float), h2.size(), f) != h2.size())
      return 1;
    std::fclose(f);
  }
  hmap::gpu::hydraulic_vpipes(h2);
  dump(dir, "vp_default_z", h2);

  // flow_simulation, as ex_flow_simulation (its CPU noise and bulkify are
  // replaced by the GPU noise of seed 2, remapped)
  hmap::Array zf = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, shape, {4.f, 4.f}, 2, 8, 0.7f);
  hmap::remap(zf);
  hmap::Array water_depth = hmap::gpu::flow_simulation(zf, 1e-2f, hmap::Array(shape, 1.f), 400);
  dump(dir, "flow256", water_depth);

  // thermal_auto_bedrock and sediment_deposition, as their examples (GPU
  // noise in place of the CPU noise)
  hmap::Array za = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, shape, {2.f, 2.f}, 1);
  dump(dir, "auto_in", za);
  hmap::gpu::thermal_auto_bedrock(za, 0.5f / shape.x, 500);
  dump(dir, "auto256", za);
  hmap::Array zs = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, shape, {4.f, 4.f}, 1);
  hmap::Array talus_s(shape, 2.f / (float)shape.x), dep_s;
  hmap::gpu::sediment_deposition(zs, talus_s, &dep_s, 0.01f, 10, 50);
  dump(dir, "sed256", zs);
  dump(dir, "sed_dep256", dep_s);

  // snow_simulation, as ex_snow_simulation (GPU noise in place of its CPU
  // noise and bulkify; a ramp melting map in place of snow_melting_map)
  hmap::Array zw = hmap::gpu::noise_fbm(hmap::NoiseType::PERLIN, shape, {4.f, 4.f}, 2);
  hmap::remap(zw);
  dump(dir, "snow_in", zw);
  const hmap::Array talus_w(shape, 2.5f / shape.x);
  hmap::Array snow0 = hmap::gpu::snow_simulation(zw, 1e-1f, hmap::Array(shape, 1.f),
                                                 hmap::Array(shape, 0.f), talus_w, 500);
  dump(dir, "snow0", snow0);
  hmap::Array melt(shape);
  for (int j = 0; j < shape.y; j++)
    for (int i = 0; i < shape.x; i++) melt(i, j) = (float)j / (float)(shape.y - 1);
  hmap::Array snow1 = hmap::gpu::snow_simulation(zw, 1e-1f, hmap::Array(shape, 1.f), melt,
                                                 talus_w, 500);
  dump(dir, "snow1", snow1);
  std::printf("shim_driver ok\n");
  return 0;
}
