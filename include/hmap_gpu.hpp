// hmap_gpu.hpp -- C++ drop-in for HighMap's grid-erosion GPU functions.
//
// Declares, with the reference's names, signatures and default arguments:
//   hmap::gpu::init_opencl        HighMap/include/highmap/opencl/gpu_opencl.hpp:30
//   hmap::gpu::noise              HighMap/include/highmap/primitives/coherent_noise.hpp:804-812
//   hmap::gpu::noise_fbm          HighMap/include/highmap/primitives/coherent_noise.hpp:821-834
//   hmap::gpu::thermal (3 overloads)  HighMap/include/highmap/erosion.hpp:1784-1801
//   hmap::gpu::thermal_auto_bedrock (3) HighMap/include/highmap/erosion.hpp:1824-1838
//   hmap::gpu::sediment_deposition (2)  HighMap/include/highmap/erosion.hpp:1366-1379
//   hmap::gpu::hydraulic_vpipes   HighMap/include/highmap/erosion.hpp:1206-1223
//   hmap::gpu::flow_simulation    HighMap/include/highmap/hydrology/hydrology.hpp:875-882
//   hmap::gpu::snow_simulation    HighMap/include/highmap/hydrology/hydrology.hpp:942-956
// plus the minimal hmap::Array (HighMap/include/highmap/array.hpp:31-44,
// 236-239: float storage, shape {nx, ny}, index j*nx + i) and hmap::NoiseType
// (HighMap/include/highmap/functions.hpp:66-81) those signatures use.
//
// The functions run on libhmg (include/hmg.h) on the first HIP device.  The
// reference returns void and reports nothing; here a failure throws
// std::runtime_error carrying hmg_last_error(), so no result is silently
// left unchanged.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#if __has_include(<glm/glm.hpp>)
#include <glm/glm.hpp>
#else
// The few glm types the signatures use, when glm itself is not installed.
namespace glm
{
struct ivec2
{
  int x = 0, y = 0;
  ivec2() = default;
  ivec2(int a, int b) : x(a), y(b) {}
};
struct vec2
{
  float x = 0.f, y = 0.f;
  vec2() = default;
  vec2(float a, float b) : x(a), y(b) {}
};
struct vec4
{
  float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
  vec4() = default;
  vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
};
struct ivec4
{
  int x = 0, y = 0, z = 0, w = 0;
  ivec4() = default;
  ivec4(int a, int b, int c, int d) : x(a), y(b), z(c), w(d) {}
};
inline bool operator==(const ivec2 &a, const ivec2 &b) { return a.x == b.x && a.y == b.y; }
inline bool operator!=(const ivec2 &a, const ivec2 &b) { return !(a == b); }
} // namespace glm
#endif

namespace hmap
{

enum NoiseType : int
{
  PARBERRY,
  PERLIN,
  PERLIN_BILLOW,
  PERLIN_HALF,
  SIMPLEX2,
  SIMPLEX2S,
  VALUE,
  VALUE_CUBIC,
  VALUE_DELAUNAY,
  VALUE_LINEAR,
  WORLEY,
  WORLEY_DOUBLE,
  WORLEY_VALUE,
};

// 2-D float grid with hmap::Array's layout: shape {ni, nj}, index j*ni + i.
class Array
{
public:
  glm::ivec2         shape = glm::ivec2(0, 0);
  std::vector<float> vector;

  Array() = default;
  explicit Array(glm::ivec2 s) : shape(s), vector((size_t)s.x * (size_t)s.y, 0.f) {}
  Array(glm::ivec2 s, float value) : shape(s), vector((size_t)s.x * (size_t)s.y, value) {}

  float &operator()(int i, int j) { return vector[(size_t)j * shape.x + i]; }
  const float &operator()(int i, int j) const { return vector[(size_t)j * shape.x + i]; }
  size_t       size() const { return vector.size(); }
  float        sum() const;
  float        min() const;
  float        max() const;

  // Interchange formats (HighMap/src/array/io.cpp:28-66,111-131).
  // Raw float32 in vector order (j*nx + i); from_file reads size() floats
  // into the current shape.
  void to_file(const std::string &fname) const;
  void from_file(const std::string &fname);
  // .npy: written as float32 of shape (nx, ny) in Fortran order (so element
  // (i, j) is at j*nx + i, as in memory); read back in either order, the
  // shape taken from the file.  Throws std::runtime_error on I/O errors or
  // a dtype other than little-endian float32.
  void to_numpy(const std::string &fname) const;
  void from_numpy(const std::string &fname);

  // Sampling between cells, (i, j) the lower corner, (u, v) the offsets
  // (HighMap/src/array/methods.cpp:203-245): bilinear over the 2x2 cells
  // (a corner on the last row or column steps one cell back with offset 1),
  // bicubic Catmull-Rom over the 4x4 cells around, indices clamped.
  float get_value_bilinear_at(int i, int j, float u, float v) const;
  float get_value_bicubic_at(int i, int j, float u, float v) const;

  // Resampling to another shape (HighMap/src/array/methods.cpp:364-383 over
  // hmap::interpolate_array_bilinear / _bicubic with their defaults,
  // HighMap/src/interpolate/interpolate_array.cpp:15-185): target point i
  // of n_t sits at 0.5/n_t + i (1 - 0.5/n_t) / n_t of the unit square (the
  // reference's grid: pixel-centred start, end point excluded), i.e. at
  // source index x n_s - 0.5.  resample_to_shape is the bilinear form.  The
  // same shape is not an identity: the last cell samples half a cell in.
  Array resample_to_shape(glm::ivec2 new_shape) const;
  Array resample_to_shape_bilinear(glm::ivec2 new_shape) const;
  Array resample_to_shape_bicubic(glm::ivec2 new_shape) const;

  // Cell by cell; the shapes must agree (std::invalid_argument otherwise).
  Array operator+(const Array &other) const;
  Array operator-(const Array &other) const;
};

// hmap::remap (HighMap/src/range/range.cpp:342-361).
void remap(Array &array, float vmin = 0.f, float vmax = 1.f);

namespace gpu
{

bool init_opencl();

Array noise(NoiseType     noise_type,
            glm::ivec2    shape,
            glm::vec2     kw,
            std::uint32_t seed,
            const Array  *p_noise_x = nullptr,
            const Array  *p_noise_y = nullptr,
            const Array  *p_stretching = nullptr,
            glm::vec4     bbox = {0.f, 1.f, 0.f, 1.f},
            glm::ivec2    period = {0, 0});

Array noise_fbm(NoiseType     noise_type,
                glm::ivec2    shape,
                glm::vec2     kw,
                std::uint32_t seed,
                int           octaves = 8,
                float         weight = 0.7f,
                float         persistence = 0.5f,
                float         lacunarity = 2.f,
                const Array  *p_ctrl_param = nullptr,
                const Array  *p_noise_x = nullptr,
                const Array  *p_noise_y = nullptr,
                const Array  *p_stretching = nullptr,
                glm::vec4     bbox = {0.f, 1.f, 0.f, 1.f},
                glm::ivec2    period = {0, 0});

void thermal(Array       &z,
             const Array &talus,
             int          iterations = 10,
             Array       *p_bedrock = nullptr,
             Array       *p_deposition_map = nullptr);

void thermal(Array       &z,
             const Array *p_mask,
             const Array &talus,
             int          iterations = 10,
             Array       *p_bedrock = nullptr,
             Array       *p_deposition_map = nullptr);

void thermal(Array &z,
             float  talus,
             int    iterations = 10,
             Array *p_bedrock = nullptr,
             Array *p_deposition_map = nullptr);

void thermal_auto_bedrock(Array       &z,
                          const Array &talus,
                          int          iterations = 10,
                          Array       *p_deposition_map = nullptr);

void thermal_auto_bedrock(Array       &z,
                          const Array *p_mask,
                          const Array &talus,
                          int          iterations = 10,
                          Array       *p_deposition_map = nullptr);

void thermal_auto_bedrock(Array &z,
                          float  talus,
                          int    iterations = 10,
                          Array *p_deposition_map = nullptr);

void sediment_deposition(Array       &z,
                         Array       *p_mask,
                         const Array &talus,
                         Array       *p_deposition_map = nullptr,
                         float        max_deposition = 0.01f,
                         int          iterations = 5,
                         int          thermal_subiterations = 10);

void sediment_deposition(Array       &z,
                         const Array &talus,
                         Array       *p_deposition_map = nullptr,
                         float        max_deposition = 0.01f,
                         int          iterations = 5,
                         int          thermal_subiterations = 10);

void hydraulic_vpipes(Array &z,
                      float  water_height = 1e-2f,
                      bool   maintain_water_volume = true,
                      float  evap_rate = 0.1f,
                      int    iterations = 50,
                      float  dt = 0.5f,
                      float  k_capacity = 0.5f,
                      float  k_erode = 0.001f,
                      float  k_depose = 0.01f,
                      float  k_discharge_exp = 1.f,
                      float  downcutting_max_depth_ratio = 10.f,
                      bool   flux_diffusion = true,
                      float  flux_diffusion_strength = 0.01f,
                      Array *p_rain_map = nullptr,
                      Array *p_water_depth = nullptr,
                      Array *p_sediment = nullptr,
                      Array *p_vel_u = nullptr,
                      Array *p_vel_v = nullptr);

// The parameter is named water_depth in the reference's declaration and
// water_height in its definition (flow_simulation.cpp:26-33).
Array flow_simulation(const Array &z,
                      float        water_depth,
                      const Array &depth_map,
                      int          iterations,
                      float        dt = 0.5f,
                      bool         flux_diffusion = true,
                      float        flux_diffusion_strength = 0.01f,
                      float        dry_out_ratio = 0.f);

Array snow_simulation(const Array &z,
                      float        snow_depth,
                      const Array &fall_map,
                      const Array &melting_map,
                      const Array &talus,
                      int          iterations,
                      float        dt = 0.5f,
                      float        fall_iterations_ratio = 1.f,
                      float        k_snow = 0.5f,
                      float        k_visc = 0.1f,
                      float        k_melt_factor = 0.8f,
                      float        k_depth_ratio = 1.f,
                      float        k_depth_slope_ratio = 1.f,
                      bool         post_filter = true,
                      float        thermal_talus_ratio = 0.2f);

} // namespace gpu
} // namespace hmap
