// hmap_gpu.cpp -- the hmap::gpu drop-in functions over libhmg's C ABI.
//
// Each function keeps the reference's host-side contract: host arrays in,
// results written back into the same arrays, optional outputs assigned when
// their pointer is not null.  The device work is libhmg's; nothing here
// computes the hot path on the CPU.
#include "hmap_gpu.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>

#include "hmg.h"

namespace hmap
{

float Array::sum() const
{
  // Array::sum (HighMap/src/array/methods.cpp:428-431)
  return std::accumulate(vector.begin(), vector.end(), 0.f);
}

float Array::min() const
{
  return vector.empty() ? 0.f : *std::min_element(vector.begin(), vector.end());
}

float Array::max() const
{
  return vector.empty() ? 0.f : *std::max_element(vector.begin(), vector.end());
}

// ------------------------------------------------------------ file formats

void Array::to_file(const std::string &fname) const
{
  // Array::to_file (HighMap/src/array/io.cpp:111-121): the floats as stored
  FILE *f = std::fopen(fname.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + fname + " for writing");
  const size_t n = std::fwrite(vector.data(), sizeof(float), vector.size(), f);
  std::fclose(f);
  if (n != vector.size()) throw std::runtime_error("short write to " + fname);
}

void Array::from_file(const std::string &fname)
{
  // Array::from_file (io.cpp:28-37): size() floats into the current shape
  FILE *f = std::fopen(fname.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + fname);
  const size_t n = std::fread(vector.data(), sizeof(float), vector.size(), f);
  std::fclose(f);
  if (n != vector.size()) throw std::runtime_error("short read from " + fname);
}

namespace
{

// The .npy header dictionary for a float32 (nx, ny) array; padded with
// spaces and a newline so that the data starts on a 64-byte boundary.
std::string npy_header(int nx, int ny, bool fortran)
{
  std::string d = "{'descr': '<f4', 'fortran_order': ";
  d += fortran ? "True" : "False";
  d += ", 'shape': (" + std::to_string(nx) + ", " + std::to_string(ny) + "), }";
  const size_t pre = 10; // magic (6) + version (2) + header length (2)
  size_t       total = pre + d.size() + 1;
  total = (total + 63) / 64 * 64;
  d.append(total - pre - d.size() - 1, ' ');
  d += '\n';
  return d;
}

// Value text following `key` in the header dictionary.
std::string npy_field(const std::string &h, const std::string &key)
{
  const size_t k = h.find("'" + key + "'");
  if (k == std::string::npos) throw std::runtime_error(".npy header without " + key);
  size_t v = h.find(':', k);
  if (v == std::string::npos) throw std::runtime_error("bad .npy header");
  v++;
  while (v < h.size() && h[v] == ' ') v++;
  size_t e = v;
  if (h[v] == '(')
    e = h.find(')', v) + 1;
  else
    while (e < h.size() && h[e] != ',' && h[e] != '}') e++;
  return h.substr(v, e - v);
}

} // namespace

void Array::to_numpy(const std::string &fname) const
{
  // Array::to_numpy (io.cpp:123-131): shape (nx, ny), Fortran order, so the
  // data is the vector as stored
  const std::string h = npy_header(shape.x, shape.y, true);
  FILE             *f = std::fopen(fname.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + fname + " for writing");
  const unsigned char pre[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t      hl = (uint16_t)h.size();
  const unsigned char hlb[2] = {(unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
  bool ok = std::fwrite(pre, 1, 8, f) == 8 && std::fwrite(hlb, 1, 2, f) == 2 &&
            std::fwrite(h.data(), 1, h.size(), f) == h.size() &&
            std::fwrite(vector.data(), sizeof(float), vector.size(), f) == vector.size();
  std::fclose(f);
  if (!ok) throw std::runtime_error("short write to " + fname);
}

void Array::from_numpy(const std::string &fname)
{
  // Array::from_numpy (io.cpp:39-66): shape (nx, ny) from the file; Fortran
  // order is the memory order, C order is transposed on the way in
  FILE *f = std::fopen(fname.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + fname);
  unsigned char pre[8];
  std::string   h;
  bool          ok = std::fread(pre, 1, 8, f) == 8 && std::memcmp(pre + 1, "NUMPY", 5) == 0 &&
            pre[0] == 0x93;
  uint32_t hl = 0;
  if (ok && pre[6] == 1)
  {
    unsigned char b[2];
    ok = std::fread(b, 1, 2, f) == 2;
    hl = (uint32_t)b[0] | ((uint32_t)b[1] << 8);
  }
  else if (ok && (pre[6] == 2 || pre[6] == 3))
  {
    unsigned char b[4];
    ok = std::fread(b, 1, 4, f) == 4;
    hl = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
  }
  else
    ok = false;
  if (ok)
  {
    h.resize(hl);
    ok = std::fread(&h[0], 1, hl, f) == hl;
  }
  if (!ok)
  {
    std::fclose(f);
    throw std::runtime_error(fname + " is not a .npy file");
  }
  const std::string descr = npy_field(h, "descr");
  if (descr != "'<f4'")
  {
    std::fclose(f);
    throw std::runtime_error(fname + ": dtype " + descr + " is not little-endian float32");
  }
  const bool        fortran = npy_field(h, "fortran_order") == "True";
  const std::string sh = npy_field(h, "shape");
  int               a = 0, b = 0;
  if (std::sscanf(sh.c_str(), "(%d, %d)", &a, &b) != 2 || a < 0 || b < 0)
  {
    std::fclose(f);
    throw std::runtime_error(fname + ": shape " + sh + " is not 2-D");
  }
  std::vector<float> d((size_t)a * (size_t)b);
  ok = std::fread(d.data(), sizeof(float), d.size(), f) == d.size();
  std::fclose(f);
  if (!ok) throw std::runtime_error("short read from " + fname);
  shape = glm::ivec2(a, b);
  vector.assign(d.size(), 0.f);
  if (fortran)
    vector = std::move(d);
  else
    for (int j = 0; j < b; j++)
      for (int i = 0; i < a; i++) (*this)(i, j) = d[(size_t)i * b + j];
}

void remap(Array &array, float vmin, float vmax)
{
  if (array.vector.empty()) return;
  const float mn = array.min(), mx = array.max();
  if (mn == mx)
  {
    std::fill(array.vector.begin(), array.vector.end(), vmin);
    return;
  }
  const float scale = (vmax - vmin) / (mx - mn);
  const float offset = vmin - mn * scale;
  for (float &x : array.vector) x = x * scale + offset;
}

// ------------------------------------------------------------- resampling

// Catmull-Rom through p[0..3] at x in [0, 1) (between p[1] and p[2]) with
// the reference's types (HighMap/include/highmap/interpolate/
// interpolate2d.hpp:107-113): the two differences of samples taken in
// float, every other operation in double, the result rounded to float.
static float catmull_rom(const float p[4], float x)
{
  const double xd = x;
  const double c3 = 3.0 * (double)(p[1] - p[2]) + (double)p[3] - (double)p[0];
  const double c2 = 2.0 * p[0] - 5.0 * p[1] + 4.0 * p[2] - (double)p[3] + xd * c3;
  const double c1 = (double)(p[2] - p[0]) + xd * c2;
  return (float)(p[1] + 0.5 * xd * c1);
}

// f00 + (f10 - f00) u + (f01 - f00) v + (f11 - f10 - f01 + f00) u v, in
// float, left to right (interpolate2d.hpp:94-105)
static float bilinear(float f00, float f10, float f01, float f11, float u, float v)
{
  const float a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
  return f00 + a10 * u + a01 * v + a11 * u * v;
}

float Array::get_value_bilinear_at(int i, int j, float u, float v) const
{
  if (i == shape.x - 1) i = shape.x - 2, u = 1.f;
  if (j == shape.y - 1) j = shape.y - 2, v = 1.f;
  const Array &a = *this;
  return bilinear(a(i, j), a(i + 1, j), a(i, j + 1), a(i + 1, j + 1), u, v);
}

float Array::get_value_bicubic_at(int i, int j, float u, float v) const
{
  // four Catmull-Rom curves along j (one per column i - 1 .. i + 2), then
  // one along i through their values
  float col[4];
  for (int m = 0; m < 4; m++)
  {
    const int ic = std::clamp(i + m - 1, 0, shape.x - 1);
    float     p[4];
    for (int n = 0; n < 4; n++) p[n] = (*this)(ic, std::clamp(j + n - 1, 0, shape.y - 1));
    col[m] = catmull_rom(p, v);
  }
  return catmull_rom(col, u);
}

// Source coordinates (cell centres at integers) of the n_t target points of
// one axis: x_i = s + i ((1 - s) / n_t), s = 0.5 / n_t, in float with i
// converted first; coordinate x_i / (1 / n_s) - 0.5.
static std::vector<float> sample_axis(int n_t, int n_s)
{
  std::vector<float> c((size_t)std::max(n_t, 0));
  const float        s = 0.5f * (1.f / (float)n_t);
  const float        step = (1.f - s) / (float)n_t;
  const float        cell = 1.f / (float)n_s;
  for (int i = 0; i < n_t; i++) c[(size_t)i] = (n_t == 1 ? s : s + (float)i * step) / cell - 0.5f;
  return c;
}

template <class Sample>
static Array resample(const Array &src, glm::ivec2 new_shape, Sample &&sample)
{
  Array out(new_shape);
  if (out.vector.empty()) return out;
  if (src.vector.empty()) throw std::invalid_argument("hmap::Array: resampling an empty array");
  const std::vector<float> cx = sample_axis(new_shape.x, src.shape.x);
  const std::vector<float> cy = sample_axis(new_shape.y, src.shape.y);
  for (int j = 0; j < new_shape.y; j++)
    for (int i = 0; i < new_shape.x; i++) out(i, j) = sample(cx[(size_t)i], cy[(size_t)j]);
  return out;
}

Array Array::resample_to_shape(glm::ivec2 new_shape) const
{
  return resample_to_shape_bilinear(new_shape);
}

Array Array::resample_to_shape_bilinear(glm::ivec2 new_shape) const
{
  const Array &a = *this;
  return resample(a, new_shape,
                  [&a](float xc, float yc)
                  {
                    // the lower corner clamped into the grid before the
                    // offsets are taken (interpolate_array.cpp:122-185)
                    const int i0 = std::clamp((int)xc, 0, a.shape.x - 1);
                    const int j0 = std::clamp((int)yc, 0, a.shape.y - 1);
                    const int i1 = std::min(i0 + 1, a.shape.x - 1);
                    const int j1 = std::min(j0 + 1, a.shape.y - 1);
                    return bilinear(a(i0, j0), a(i1, j0), a(i0, j1), a(i1, j1), xc - (float)i0,
                                    yc - (float)j0);
                  });
}

Array Array::resample_to_shape_bicubic(glm::ivec2 new_shape) const
{
  const Array &a = *this;
  return resample(a, new_shape,
                  [&a](float xc, float yc)
                  {
                    // offsets from the unclamped corner (truncated toward
                    // zero); the 4x4 reads are clamped
                    // (interpolate_array.cpp:31-103)
                    const int i0 = (int)xc, j0 = (int)yc;
                    return a.get_value_bicubic_at(i0, j0, xc - (float)i0, yc - (float)j0);
                  });
}

static void same_shape(const Array &a, const Array &b)
{
  if (!(a.shape == b.shape)) throw std::invalid_argument("hmap::Array: shapes differ");
}

Array Array::operator+(const Array &other) const
{
  same_shape(*this, other);
  Array out(shape);
  for (size_t k = 0; k < vector.size(); k++) out.vector[k] = vector[k] + other.vector[k];
  return out;
}

Array Array::operator-(const Array &other) const
{
  same_shape(*this, other);
  Array out(shape);
  for (size_t k = 0; k < vector.size(); k++) out.vector[k] = vector[k] - other.vector[k];
  return out;
}

namespace gpu
{

namespace
{
// Calls from different host threads run on different libhmg contexts, each
// with its own HIP stream and buffer pool (at most HMAP_GPU_CONTEXTS of
// them, default 4; a thread keeps the one it first gets), so that tiles
// processed on several threads (VirtualArray's VA_DISTRIBUTED mode) overlap
// their uploads, kernels and downloads on the device instead of queueing on
// one context's lock.  Results do not depend on the context.
std::mutex              g_mu;
std::vector<hmg_ctx *>  g_ctxs;
size_t                  g_next = 0;
thread_local hmg_ctx   *t_ctx = nullptr;

size_t max_contexts()
{
  const char *e = std::getenv("HMAP_GPU_CONTEXTS");
  const long  v = e ? std::strtol(e, nullptr, 10) : 4;
  return (size_t)std::max(1L, std::min(16L, v));
}

hmg_ctx *first_ctx_locked()
{
  if (g_ctxs.empty())
  {
    hmg_ctx *c = nullptr;
    if (hmg_init(0, &c) != HMG_OK)
      throw std::runtime_error(std::string("hmap::gpu: no GPU: ") + hmg_last_error());
    g_ctxs.push_back(c);
  }
  return g_ctxs.front();
}

hmg_ctx *ctx()
{
  if (t_ctx) return t_ctx;
  std::lock_guard<std::mutex> lk(g_mu);
  first_ctx_locked();
  const size_t k = g_next++ % max_contexts();
  if (k >= g_ctxs.size())
  {
    hmg_ctx *c = nullptr;
    if (hmg_init(0, &c) != HMG_OK)
      throw std::runtime_error(std::string("hmap::gpu: context: ") + hmg_last_error());
    g_ctxs.push_back(c);
  }
  t_ctx = g_ctxs[std::min(k, g_ctxs.size() - 1)];
  return t_ctx;
}

void check(int rc, const char *what)
{
  if (rc != HMG_OK)
    throw std::runtime_error(std::string("hmap::gpu::") + what + ": " + hmg_last_error());
}

void check_shape(const Array &a, glm::ivec2 s, const char *what)
{
  if (a.shape.x != s.x || a.shape.y != s.y || a.vector.size() != (size_t)s.x * (size_t)s.y)
    throw std::invalid_argument(std::string("hmap::gpu: shape mismatch for ") + what);
}
} // namespace

bool init_opencl()
{
  std::lock_guard<std::mutex> lk(g_mu);
  try
  {
    first_ctx_locked();
  }
  catch (const std::runtime_error &)
  {
    return false;
  }
  return true;
}

// HighMap/src/primitives/coherent_noise/primitives_gpu.cpp:504-541
Array noise(NoiseType     noise_type,
            glm::ivec2    shape,
            glm::vec2     kw,
            std::uint32_t seed,
            const Array  *p_noise_x,
            const Array  *p_noise_y,
            const Array * /* p_stretching: ignored, as in the reference */,
            glm::vec4     bbox,
            glm::ivec2    period)
{
  Array out(shape);
  if (p_noise_x) check_shape(*p_noise_x, shape, "noise_x");
  if (p_noise_y) check_shape(*p_noise_y, shape, "noise_y");
  const float bb[4] = {bbox.x, bbox.y, bbox.z, bbox.w};
  const int   per[2] = {period.x, period.y};
  check(hmg_noise(ctx(),
                  out.vector.data(),
                  shape.x,
                  shape.y,
                  (int)noise_type,
                  kw.x,
                  kw.y,
                  seed,
                  p_noise_x ? p_noise_x->vector.data() : nullptr,
                  p_noise_y ? p_noise_y->vector.data() : nullptr,
                  bb,
                  per),
        "noise");
  return out;
}

Array noise_fbm(NoiseType     noise_type,
                glm::ivec2    shape,
                glm::vec2     kw,
                std::uint32_t seed,
                int           octaves,
                float         weight,
                float         persistence,
                float         lacunarity,
                const Array  *p_ctrl_param,
                const Array  *p_noise_x,
                const Array  *p_noise_y,
                const Array * /* p_stretching: ignored, as in the reference */,
                glm::vec4     bbox,
                glm::ivec2    period)
{
  Array out(shape);
  if (p_ctrl_param) check_shape(*p_ctrl_param, shape, "ctrl_param");
  if (p_noise_x) check_shape(*p_noise_x, shape, "noise_x");
  if (p_noise_y) check_shape(*p_noise_y, shape, "noise_y");
  const float bb[4] = {bbox.x, bbox.y, bbox.z, bbox.w};
  const int   per[2] = {period.x, period.y};
  check(hmg_noise_fbm(ctx(),
                      out.vector.data(),
                      shape.x,
                      shape.y,
                      (int)noise_type,
                      kw.x,
                      kw.y,
                      seed,
                      octaves,
                      weight,
                      persistence,
                      lacunarity,
                      p_ctrl_param ? p_ctrl_param->vector.data() : nullptr,
                      p_noise_x ? p_noise_x->vector.data() : nullptr,
                      p_noise_y ? p_noise_y->vector.data() : nullptr,
                      bb,
                      per),
        "noise_fbm");
  return out;
}

void thermal(Array       &z,
             const Array &talus,
             int          iterations,
             Array       *p_bedrock,
             Array       *p_deposition_map)
{
  check_shape(talus, z.shape, "talus");
  if (p_bedrock) check_shape(*p_bedrock, z.shape, "bedrock");
  if (p_deposition_map) *p_deposition_map = Array(z.shape);
  check(hmg_thermal(ctx(),
                    z.vector.data(),
                    z.shape.x,
                    z.shape.y,
                    talus.vector.data(),
                    0.f,
                    iterations,
                    p_bedrock ? p_bedrock->vector.data() : nullptr,
                    p_deposition_map ? p_deposition_map->vector.data() : nullptr,
                    nullptr),
        "thermal");
}

void thermal(Array       &z,
             const Array *p_mask,
             const Array &talus,
             int          iterations,
             Array       *p_bedrock,
             Array       *p_deposition_map)
{
  check_shape(talus, z.shape, "talus");
  if (p_mask) check_shape(*p_mask, z.shape, "mask");
  if (p_bedrock) check_shape(*p_bedrock, z.shape, "bedrock");
  if (p_deposition_map) *p_deposition_map = Array(z.shape);
  check(hmg_thermal(ctx(),
                    z.vector.data(),
                    z.shape.x,
                    z.shape.y,
                    talus.vector.data(),
                    0.f,
                    iterations,
                    p_bedrock ? p_bedrock->vector.data() : nullptr,
                    p_deposition_map ? p_deposition_map->vector.data() : nullptr,
                    p_mask ? p_mask->vector.data() : nullptr),
        "thermal");
}

void thermal(Array &z, float talus, int iterations, Array *p_bedrock, Array *p_deposition_map)
{
  // the reference materialises Array(shape, talus) (thermal_gpu.cpp:94);
  // the scalar is passed through instead, with identical results
  if (p_bedrock) check_shape(*p_bedrock, z.shape, "bedrock");
  if (p_deposition_map) *p_deposition_map = Array(z.shape);
  check(hmg_thermal(ctx(),
                    z.vector.data(),
                    z.shape.x,
                    z.shape.y,
                    nullptr,
                    talus,
                    iterations,
                    p_bedrock ? p_bedrock->vector.data() : nullptr,
                    p_deposition_map ? p_deposition_map->vector.data() : nullptr,
                    nullptr),
        "thermal");
}

// HighMap/src/erosion/thermal_gpu.cpp:98-146
void thermal_auto_bedrock(Array &z, const Array &talus, int iterations, Array *p_deposition_map)
{
  thermal_auto_bedrock(z, nullptr, talus, iterations, p_deposition_map);
}

void thermal_auto_bedrock(Array       &z,
                          const Array *p_mask,
                          const Array &talus,
                          int          iterations,
                          Array       *p_deposition_map)
{
  check_shape(talus, z.shape, "talus");
  if (p_mask) check_shape(*p_mask, z.shape, "mask");
  if (p_deposition_map) *p_deposition_map = Array(z.shape);
  check(hmg_thermal_auto_bedrock(ctx(),
                                 z.vector.data(),
                                 z.shape.x,
                                 z.shape.y,
                                 talus.vector.data(),
                                 0.f,
                                 iterations,
                                 p_deposition_map ? p_deposition_map->vector.data() : nullptr,
                                 p_mask ? p_mask->vector.data() : nullptr),
        "thermal_auto_bedrock");
}

void thermal_auto_bedrock(Array &z, float talus, int iterations, Array *p_deposition_map)
{
  if (p_deposition_map) *p_deposition_map = Array(z.shape);
  check(hmg_thermal_auto_bedrock(ctx(),
                                 z.vector.data(),
                                 z.shape.x,
                                 z.shape.y,
                                 nullptr,
                                 talus,
                                 iterations,
                                 p_deposition_map ? p_deposition_map->vector.data() : nullptr,
                                 nullptr),
        "thermal_auto_bedrock");
}

// HighMap/src/erosion/deposition.cpp:19-63
void sediment_deposition(Array       &z,
                         Array       *p_mask,
                         const Array &talus,
                         Array       *p_deposition_map,
                         float        max_deposition,
                         int          iterations,
                         int          thermal_subiterations)
{
  check_shape(talus, z.shape, "talus");
  if (p_mask) check_shape(*p_mask, z.shape, "mask");
  if (p_deposition_map) *p_deposition_map = Array(z.shape);
  check(hmg_sediment_deposition(ctx(),
                                z.vector.data(),
                                z.shape.x,
                                z.shape.y,
                                talus.vector.data(),
                                0.f,
                                p_deposition_map ? p_deposition_map->vector.data() : nullptr,
                                max_deposition,
                                iterations,
                                thermal_subiterations,
                                p_mask ? p_mask->vector.data() : nullptr),
        "sediment_deposition");
}

void sediment_deposition(Array       &z,
                         const Array &talus,
                         Array       *p_deposition_map,
                         float        max_deposition,
                         int          iterations,
                         int          thermal_subiterations)
{
  sediment_deposition(z,
                      nullptr,
                      talus,
                      p_deposition_map,
                      max_deposition,
                      iterations,
                      thermal_subiterations);
}

void hydraulic_vpipes(Array &z,
                      float  water_height,
                      bool   maintain_water_volume,
                      float  evap_rate,
                      int    iterations,
                      float  dt,
                      float  k_capacity,
                      float  k_erode,
                      float  k_depose,
                      float  k_discharge_exp,
                      float  downcutting_max_depth_ratio,
                      bool   flux_diffusion,
                      float  flux_diffusion_strength,
                      Array *p_rain_map,
                      Array *p_water_depth,
                      Array *p_sediment,
                      Array *p_vel_u,
                      Array *p_vel_v)
{
  hmg_vpipes_params p;
  hmg_vpipes_defaults(&p);
  p.water_height = water_height;
  p.maintain_water_volume = maintain_water_volume ? 1 : 0;
  p.evap_rate = evap_rate;
  p.iterations = iterations;
  p.dt = dt;
  p.k_capacity = k_capacity;
  p.k_erode = k_erode;
  p.k_depose = k_depose;
  p.k_discharge_exp = k_discharge_exp;
  p.downcutting_max_depth_ratio = downcutting_max_depth_ratio;
  p.flux_diffusion = flux_diffusion ? 1 : 0;
  p.flux_diffusion_strength = flux_diffusion_strength;
  if (p_rain_map) check_shape(*p_rain_map, z.shape, "rain_map");
  // optional outputs are assigned by value, as hydraulic_vpipes_gpu.cpp:174-177
  Array *outs[4] = {p_water_depth, p_sediment, p_vel_u, p_vel_v};
  for (Array *o : outs)
    if (o) *o = Array(z.shape);
  check(hmg_vpipes(ctx(),
                   z.vector.data(),
                   z.shape.x,
                   z.shape.y,
                   &p,
                   p_rain_map ? p_rain_map->vector.data() : nullptr,
                   p_water_depth ? p_water_depth->vector.data() : nullptr,
                   p_sediment ? p_sediment->vector.data() : nullptr,
                   p_vel_u ? p_vel_u->vector.data() : nullptr,
                   p_vel_v ? p_vel_v->vector.data() : nullptr),
        "hydraulic_vpipes");
}

// HighMap/src/hydrology/flow_simulation.cpp:26-110
Array flow_simulation(const Array &z,
                      float        water_height,
                      const Array &depth_map,
                      int          iterations,
                      float        dt,
                      bool         flux_diffusion,
                      float        flux_diffusion_strength,
                      float        dry_out_ratio)
{
  check_shape(depth_map, z.shape, "depth_map");
  Array d(z.shape);
  check(hmg_flow_simulation(ctx(),
                            z.vector.data(),
                            z.shape.x,
                            z.shape.y,
                            water_height,
                            depth_map.vector.data(),
                            iterations,
                            dt,
                            flux_diffusion ? 1 : 0,
                            flux_diffusion_strength,
                            dry_out_ratio,
                            d.vector.data()),
        "flow_simulation");
  return d;
}

// HighMap/src/hydrology/snow_simulation.cpp:81-154
Array snow_simulation(const Array &z,
                      float        snow_depth,
                      const Array &fall_map,
                      const Array &melting_map,
                      const Array &talus,
                      int          iterations,
                      float        dt,
                      float        fall_iterations_ratio,
                      float        k_snow,
                      float        k_visc,
                      float        k_melt_factor,
                      float        k_depth_ratio,
                      float        k_depth_slope_ratio,
                      bool         post_filter,
                      float        thermal_talus_ratio)
{
  check_shape(fall_map, z.shape, "fall_map");
  check_shape(melting_map, z.shape, "melting_map");
  check_shape(talus, z.shape, "talus");
  Array s(z.shape);
  check(hmg_snow_simulation(ctx(),
                            z.vector.data(),
                            z.shape.x,
                            z.shape.y,
                            snow_depth,
                            fall_map.vector.data(),
                            melting_map.vector.data(),
                            talus.vector.data(),
                            iterations,
                            dt,
                            fall_iterations_ratio,
                            k_snow,
                            k_visc,
                            k_melt_factor,
                            k_depth_ratio,
                            k_depth_slope_ratio,
                            post_filter ? 1 : 0,
                            thermal_talus_ratio,
                            s.vector.data()),
        "snow_simulation");
  return s;
}

} // namespace gpu
} // namespace hmap
