// snow.hip -- hmap::gpu::snow_simulation on the device
// (HighMap/src/hydrology/snow_simulation.cpp:81-154, kernel
// HighMap/src/gpu_opencl/kernels/snow_simulation.cl:6-167).
//
// The reference uploads s, runs one kernel launch, downloads s and fills
// its borders on the host, once per iteration.  Here the whole loop stays
// in HBM: each launch computes the kernel's cell update, applies
// fill_borders by evaluating border cells at their fill source (the inward
// cell, the diagonal one at corners), and adds the next iteration's snow
// fall, so one launch per iteration reads s once and writes it once.  The
// post-filter (mask, gradient talus, masked thermal) runs on the device too.
//
// FP32 contract: one IEEE operation per source operation in the kernel's
// order, correctly rounded division and sqrt, hypot as sqrt(x*x + y*y),
// OpenCL smoothstep/mix by their specification formulas, and the project
// exp (hm_exp).
#include <algorithm>

#include "hmg_internal.hpp"

namespace hmg
{

struct SnowConst
{
  int   nx, ny;
  float dt, snow_depth, k_snow, k_melt, k_visc, k_depth_slope_ratio;
  float sc0; // k_depth_ratio * snow_depth
};

__device__ __forceinline__ float snow_smoothstep(float e0, float e1, float x)
{
  float t = (x - e0) / (e1 - e0);
  t = fminf(fmaxf(t, 0.f), 1.f);
  return t * t * (3.f - 2.f * t);
}

// Drop-weighted share that neighbour n (height hn) sends to the centre
// (height hc): hn's drops to the centre's opposite axis neighbour (ho,
// weight 1) and to the two diagonal cells beside n (ha, hb, weight diag)
// form the denominator (snow_simulation.cl:83-147).
__device__ __forceinline__ float snow_inflow(float hn, float hc, float ho, float ha, float hb,
                                             float repose, const SnowConst &k)
{
  constexpr float diag = 0.70710678f;
  const float     drop = hn - hc;
  const float     w = fmaxf(0.f, drop);
  const float     f = k.k_snow * fmaxf(0.f, drop - repose) * k.dt;
  const float     sum = fmaxf(0.f, hn - ho) + fmaxf(0.f, hn - ha) * diag + fmaxf(0.f, hn - hb) * diag;
  return sum > 1e-6f ? f * w / sum : 0.f;
}

// The kernel's update of cell (i, j), reads clamped to the grid.
__device__ __forceinline__ float snow_cell(const float *__restrict__ z,
                                           const float *__restrict__ s,
                                           const float *__restrict__ talus,
                                           const float *__restrict__ melt,
                                           int                      i,
                                           int                      j,
                                           const SnowConst         &k)
{
  constexpr float diag = 0.70710678f;
  const int       nx = k.nx;
  const int       il = max(i - 1, 0), ir = min(i + 1, nx - 1);
  const size_t    rd = (size_t)max(j - 1, 0) * nx, rc = (size_t)j * nx,
               ru = (size_t)min(j + 1, k.ny - 1) * nx;
  // snow of the 3x3 neighbourhood
  const float sc = s[rc + i];
  const float sl = s[rc + il], sr = s[rc + ir], sd = s[rd + i], su = s[ru + i];
  const float sdl = s[rd + il], sdr = s[rd + ir], sul = s[ru + il], sur = s[ru + ir];

  const float repose =
      talus[rc + i] * (1.f + k.k_depth_slope_ratio * snow_smoothstep(0.f, k.sc0, sc));
  const float hc = z[rc + i] + sc;
  const float hl = z[rc + il] + sl, hr = z[rc + ir] + sr;
  const float hd = z[rd + i] + sd, hu = z[ru + i] + su;
  const float hdl = z[rd + il] + sdl, hdr = z[rd + ir] + sdr;
  const float hul = z[ru + il] + sul, hur = z[ru + ir] + sur;

  // outflow from the cell's own slope above its repose slope
  const float gx = 0.5f * (hr - hl), gy = 0.5f * (hu - hd);
  const float over = fmaxf(0.f, hm_hypot(gx, gy) - repose);
  float       flux = fminf(k.k_snow * over * k.dt, sc);
  flux *= hm_exp(-sc / k.sc0);
  const float wsum = fmaxf(0.f, hc - hl) + fmaxf(0.f, hc - hr) + fmaxf(0.f, hc - hd) +
                     fmaxf(0.f, hc - hu) +
                     diag * (fmaxf(0.f, hc - hdl) + fmaxf(0.f, hc - hdr) + fmaxf(0.f, hc - hul) +
                             fmaxf(0.f, hc - hur));
  const float out = wsum > 1e-6f ? flux : 0.f;

  // inflow from the left, right, lower and upper neighbours, in that order
  float in = 0.f;
  in += snow_inflow(hl, hc, hr, hdl, hul, repose, k);
  in += snow_inflow(hr, hc, hl, hdr, hur, repose, k);
  in += snow_inflow(hd, hc, hu, hdl, hdr, repose, k);
  in += snow_inflow(hu, hc, hd, hul, hur, repose, k);
  float sn = sc - out + in;

  // viscosity towards the weighted 8-neighbour mean (mix), then melting
  const float avg = (sl + sr + sd + su + diag * (sdl + sdr + sul + sur)) / (4.f + 4.f * diag);
  sn = sn + (avg - sn) * k.k_visc;
  sn -= fminf(k.k_melt * melt[rc + i] * k.dt, sn);
  return fmaxf(0.f, sn);
}

// First fall: s = 0 + c * fall (the reference adds to a zero array).
__global__ void __launch_bounds__(256) k_snow_init(float *__restrict__ s,
                                                   const float *__restrict__ fall,
                                                   float  c,
                                                   size_t n)
{
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    s[q] = 0.f + c * fall[q];
}

// One iteration: the kernel's update, fill_borders, and (ADD) the next
// iteration's fall c * fall_map.  Blocks of 64 x 4 cells.
template <bool ADD>
__global__ void __launch_bounds__(256) k_snow_step(const float *__restrict__ z,
                                                   const float *__restrict__ s,
                                                   const float *__restrict__ talus,
                                                   const float *__restrict__ melt,
                                                   const float *__restrict__ fall,
                                                   float *__restrict__ so,
                                                   SnowConst k,
                                                   float     c)
{
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  // fill_borders (boundary.cpp:119-135): columns 0 and nx-1 copy the
  // inward column, then rows 0 and ny-1 the inward row
  const int si = min(max(i, 1), k.nx - 2), sj = min(max(j, 1), k.ny - 2);
  float     v = snow_cell(z, s, talus, melt, si, sj, k);
  const size_t q = (size_t)j * k.nx + i;
  if (ADD) v = v + c * fall[q];
  so[q] = v;
}

// Post-filter inputs: mask = clamp_max_smooth(s / snow_depth, 1, 0.2)
// (range.cpp:98-110, std::max/std::min rules) and zw = z + s.
__global__ void __launch_bounds__(256) k_snow_prep(const float *__restrict__ s,
                                                   const float *__restrict__ z,
                                                   float *__restrict__ mask,
                                                   float *__restrict__ zw,
                                                   float  snow_depth,
                                                   size_t n)
{
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
  {
    const float x = s[q] / snow_depth;
    const float d = 0.2f - fabsf(x - 1.f);
    const float h = (d < 0.f ? 0.f : d) / 0.2f;
    const float lo = 1.f < x ? 1.f : x;
    mask[q] = lo - h * h * h * 0.2f / 6.f;
    zw[q] = z[q] + s[q];
  }
}

// talus = ratio * gradient_norm(a) (gradient.cpp:124-135,175-217:
// central differences inside, one-sided ones without the 0.5 at the first
// and last column/row; math.cpp:142-151 hypot).
__global__ void __launch_bounds__(256) k_snow_talus(const float *__restrict__ a,
                                                    float *__restrict__ t,
                                                    int   nx,
                                                    int   ny,
                                                    float ratio)
{
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const size_t c = (size_t)j * nx + i;
  float        dx, dy;
  if (i == 0) dx = a[c + 1] - a[c];
  else if (i == nx - 1) dx = a[c] - a[c - 1];
  else dx = 0.5f * (a[c + 1] - a[c - 1]);
  if (j == 0) dy = a[c + nx] - a[c];
  else if (j == ny - 1) dy = a[c] - a[c - nx];
  else dy = 0.5f * (a[c + nx] - a[c - nx]);
  t[c] = ratio * sqrtf(dx * dx + dy * dy);
}

// s = clamp_min(lerp(zw, zf, mask) - z, 0) with the post-filter, else
// clamp_min(s, 0) (range.cpp:127-135: x > 0 ? x : 0).
__global__ void __launch_bounds__(256) k_snow_finish(float *__restrict__ s,
                                                     const float *__restrict__ z,
                                                     const float *__restrict__ zw,
                                                     const float *__restrict__ zf,
                                                     const float *__restrict__ mask,
                                                     size_t n)
{
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
  {
    float v = s[q];
    if (mask)
    {
      const float t = mask[q];
      v = (zw[q] * (1.f - t) + zf[q] * t) - z[q];
    }
    s[q] = v > 0.f ? v : 0.f;
  }
}

static int flat_blocks(size_t n)
{
  return (int)std::min<size_t>((n + 255) / 256, 4096);
}

static int check_snow_args(int nx, int ny, int iterations)
{
  if (nx < 3 || ny < 3) return set_error(HMG_ERR_INVALID, "snow_simulation needs nx, ny >= 3");
  if ((size_t)nx * (size_t)ny > ((size_t)1 << 31))
    return set_error(HMG_ERR_INVALID, "grid larger than 2^31 cells");
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  return HMG_OK;
}

// The whole call on device planes; s receives the snow depth.
static int snow_run(hmg_ctx     *ctx,
                    const float *z,
                    int          nx,
                    int          ny,
                    float        snow_depth,
                    const float *fall,
                    const float *melt,
                    const float *talus,
                    int          iterations,
                    float        dt,
                    float        fall_iterations_ratio,
                    float        k_snow,
                    float        k_visc,
                    float        k_melt_factor,
                    float        k_depth_ratio,
                    float        k_depth_slope_ratio,
                    int          post_filter,
                    float        thermal_talus_ratio,
                    float       *s)
{
  hipStream_t  st = ctx->stream;
  const size_t n = (size_t)nx * ny, bytes = n * sizeof(float);
  SnowConst    k;
  k.nx = nx;
  k.ny = ny;
  k.dt = dt;
  k.snow_depth = snow_depth;
  k.k_snow = k_snow;
  k.k_melt = k_melt_factor / (float)iterations; // snow_simulation.cpp:100
  k.k_visc = k_visc;
  k.k_depth_slope_ratio = k_depth_slope_ratio;
  k.sc0 = k_depth_ratio * snow_depth;
  int n_fall = (int)(fall_iterations_ratio * (float)iterations);
  n_fall = std::max(1, n_fall);
  const float c = snow_depth / (float)n_fall;

  DevBuf other;
  HMG_TRY(other.get(ctx->pool, bytes));
  // ping-pong so that the last iteration writes s
  float *cur = (iterations & 1) ? other.as<float>() : s;
  float *nxt = (iterations & 1) ? s : other.as<float>();
  if (iterations > 0)
  {
    k_snow_init<<<flat_blocks(n), 256, 0, st>>>(cur, fall, c, n);
    HMG_TRY(check_launch("snow_init"));
  }
  else // the fall is added inside the iteration loop only
    HMG_HIP(hipMemsetAsync(s, 0, bytes, st));
  const dim3 blk(64, 4), grd((nx + 63) / 64, (ny + 3) / 4);
  for (int it = 0; it < iterations; it++)
  {
    if (it + 1 < iterations && it + 1 < n_fall)
      k_snow_step<true><<<grd, blk, 0, st>>>(z, cur, talus, melt, fall, nxt, k, c);
    else
      k_snow_step<false><<<grd, blk, 0, st>>>(z, cur, talus, melt, fall, nxt, k, c);
    HMG_TRY(check_launch("snow_step"));
    std::swap(cur, nxt);
  }
  HMG_TRY(launch_extrapolate_borders(s, nx, ny, st));
  if (!post_filter)
  {
    k_snow_finish<<<flat_blocks(n), 256, 0, st>>>(s, z, nullptr, nullptr, nullptr, n);
    HMG_TRY(check_launch("snow_finish"));
    // `other` returns to the pool when this function ends
    HMG_HIP(hipStreamSynchronize(st));
    return HMG_OK;
  }
  // post-filter: thermal on z + s with talus from its gradient, blended by
  // the snow mask (snow_simulation.cpp:139-150)
  DevBuf mask, zw, zf, tal;
  HMG_TRY(mask.get(ctx->pool, bytes));
  HMG_TRY(zw.get(ctx->pool, bytes));
  HMG_TRY(zf.get(ctx->pool, bytes));
  HMG_TRY(tal.get(ctx->pool, bytes));
  k_snow_prep<<<flat_blocks(n), 256, 0, st>>>(s, z, mask.as<float>(), zw.as<float>(), snow_depth,
                                              n);
  HMG_TRY(check_launch("snow_prep"));
  k_snow_talus<<<grd, blk, 0, st>>>(zw.as<float>(), tal.as<float>(), nx, ny, thermal_talus_ratio);
  HMG_TRY(check_launch("snow_talus"));
  HMG_HIP(hipMemcpyAsync(zf.ptr, zw.ptr, bytes, hipMemcpyDeviceToDevice, st));
  // the thermal scratch plane: `other` is free again
  HMG_TRY(launch_thermal(ctx, zf.as<float>(), other.as<float>(), nx, ny, tal.as<float>(), 0.f,
                         iterations, nullptr, st, nullptr, true));
  k_snow_finish<<<flat_blocks(n), 256, 0, st>>>(s, z, zw.as<float>(), zf.as<float>(),
                                                mask.as<float>(), n);
  HMG_TRY(check_launch("snow_finish"));
  // the pooled planes return to the pool when this function ends: wait for
  // the kernels that use them
  HMG_HIP(hipStreamSynchronize(st));
  return HMG_OK;
}

} // namespace hmg

using namespace hmg;

extern "C" int hmg_snow_simulation_dev(hmg_ctx     *ctx,
                                       const float *z_dev,
                                       int          nx,
                                       int          ny,
                                       float        snow_depth,
                                       const float *fall_map_dev,
                                       const float *melting_map_dev,
                                       const float *talus_dev,
                                       int          iterations,
                                       float        dt,
                                       float        fall_iterations_ratio,
                                       float        k_snow,
                                       float        k_visc,
                                       float        k_melt_factor,
                                       float        k_depth_ratio,
                                       float        k_depth_slope_ratio,
                                       int          post_filter,
                                       float        thermal_talus_ratio,
                                       float       *s_dev)
{
  clear_error();
  HMG_TRY(check_snow_args(nx, ny, iterations));
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_dev || !fall_map_dev || !melting_map_dev || !talus_dev || !s_dev)
    return set_error(HMG_ERR_INVALID, "null plane");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  HMG_TRY(snow_run(ctx, z_dev, nx, ny, snow_depth, fall_map_dev, melting_map_dev, talus_dev,
                   iterations, dt, fall_iterations_ratio, k_snow, k_visc, k_melt_factor,
                   k_depth_ratio, k_depth_slope_ratio, post_filter, thermal_talus_ratio, s_dev));
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

extern "C" int hmg_snow_simulation(hmg_ctx     *ctx,
                                   const float *z_host,
                                   int          nx,
                                   int          ny,
                                   float        snow_depth,
                                   const float *fall_map_host,
                                   const float *melting_map_host,
                                   const float *talus_host,
                                   int          iterations,
                                   float        dt,
                                   float        fall_iterations_ratio,
                                   float        k_snow,
                                   float        k_visc,
                                   float        k_melt_factor,
                                   float        k_depth_ratio,
                                   float        k_depth_slope_ratio,
                                   int          post_filter,
                                   float        thermal_talus_ratio,
                                   float       *s_host)
{
  clear_error();
  HMG_TRY(check_snow_args(nx, ny, iterations));
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_host || !fall_map_host || !melting_map_host || !talus_host || !s_host)
    return set_error(HMG_ERR_INVALID, "null array");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  hipStream_t  st = ctx->stream;
  const size_t bytes = (size_t)nx * ny * sizeof(float);
  DevBuf       zb, fb, mb, tb, sb;
  HMG_TRY(zb.get(ctx->pool, bytes));
  HMG_TRY(fb.get(ctx->pool, bytes));
  HMG_TRY(mb.get(ctx->pool, bytes));
  HMG_TRY(tb.get(ctx->pool, bytes));
  HMG_TRY(sb.get(ctx->pool, bytes));
  HMG_HIP(hipMemcpyAsync(zb.ptr, z_host, bytes, hipMemcpyHostToDevice, st));
  HMG_HIP(hipMemcpyAsync(fb.ptr, fall_map_host, bytes, hipMemcpyHostToDevice, st));
  HMG_HIP(hipMemcpyAsync(mb.ptr, melting_map_host, bytes, hipMemcpyHostToDevice, st));
  HMG_HIP(hipMemcpyAsync(tb.ptr, talus_host, bytes, hipMemcpyHostToDevice, st));
  HMG_TRY(snow_run(ctx, zb.as<float>(), nx, ny, snow_depth, fb.as<float>(), mb.as<float>(),
                   tb.as<float>(), iterations, dt, fall_iterations_ratio, k_snow, k_visc,
                   k_melt_factor, k_depth_ratio, k_depth_slope_ratio, post_filter,
                   thermal_talus_ratio, sb.as<float>()));
  HMG_HIP(hipMemcpyAsync(s_host, sb.ptr, bytes, hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  return HMG_OK;
}
