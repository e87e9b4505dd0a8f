// noise.hip -- hmap::gpu::noise_fbm, PERLIN branch, for gfx950.
//
// Follows the reference kernel noise_fbm (HighMap/src/gpu_opencl/kernels/
// noise.cl:456-504) with base_perlin_fbm (:194-216), base_perlin (:28-63),
// wrap_lattice/scale_period (:12-26), g_to_xy (_common_index.cl:8-24) and
// the hashes of _common_rand.cl; host wrapper
// HighMap/src/primitives/coherent_noise/primitives_gpu.cpp:543-591.
//
// The kernel is VALU-bound (8 octaves x 4 corners x one hash sin and one
// sincos), not memory-bound: one 4-byte store per cell.  sin/cos are the
// project-written routines of hmg_cell.hpp, so the field matches the FP32
// restatement bit for bit.
#include "hmg_internal.hpp"

namespace hmg
{

__device__ __forceinline__ float wrap_lattice(float c, int period)
{
  if (period > 0) c = c - (float)period * floorf(c / (float)period);
  return c;
}

__device__ __forceinline__ int scale_period(int period, float lacunarity)
{
  if (period > 0) period = (int)((float)period * lacunarity + 0.5f);
  return period;
}

// grad22f at a lattice corner dotted with the offset (dx, dy).
__device__ __forceinline__ float perlin_corner(float cx, float cy, float fseed, float dx, float dy)
{
  float angle = 6.2831853f * hash12f(cx, cy, fseed);
  float s, c;
  hm_sincos(angle, s, c);
  return c * dx + s * dy;
}

__device__ __forceinline__ float fade5(float x)
{
  return x * x * x * (x * (6.f * x - 15.f) + 10.f);
}

__device__ __forceinline__ float lerp_cl(float a, float b, float t)
{
  return (1.f - t) * a + t * b;
}

__device__ __forceinline__ float base_perlin(float px, float py, float fseed, int per_x, int per_y)
{
  float ix = floorf(px), iy = floorf(py);
  float fx = fract_cl(px), fy = fract_cl(py);
  float x0 = wrap_lattice(ix, per_x), x1 = wrap_lattice(ix + 1.f, per_x);
  float y0 = wrap_lattice(iy, per_y), y1 = wrap_lattice(iy + 1.f, per_y);
  float n00 = perlin_corner(x0, y0, fseed, fx, fy);
  float n10 = perlin_corner(x1, y0, fseed, fx - 1.f, fy);
  float n01 = perlin_corner(x0, y1, fseed, fx, fy - 1.f);
  float n11 = perlin_corner(x1, y1, fseed, fx - 1.f, fy - 1.f);
  float ux = fade5(fx), uy = fade5(fy);
  float a = lerp_cl(n00, n10, ux);
  float b = lerp_cl(n01, n11, ux);
  return 1.42857f * lerp_cl(a, b, uy);
}

struct NoiseArgs
{
  int   nx, ny;
  float kx, ky;
  float fseed;
  int   octaves;
  float weight, persistence, lacunarity;
  float b0, b1, b2, b3;
  int   per_x, per_y;
};

__global__ void __launch_bounds__(256) k_noise_fbm_perlin(float *__restrict__ out,
                                                          const float *__restrict__ ctrl,
                                                          const float *__restrict__ nxp,
                                                          const float *__restrict__ nyp,
                                                          NoiseArgs a)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  size_t idx = (size_t)j * a.nx + i;
  float  ct = ctrl ? ctrl[idx] : 1.f;
  float  dx = nxp ? nxp[idx] : 0.f;
  float  dy = nyp ? nyp[idx] : 0.f;
  float  w = (1.f - ct) + ct * a.weight;

  float x = (float)i / (float)a.nx;
  float y = (float)j / (float)a.ny;
  x = a.kx * (x * (a.b1 - a.b0) + a.b0) + a.kx * dx;
  y = a.ky * (y * (a.b3 - a.b2) + a.b2) + a.ky * dy;

  float n = 0.f, nf = 1.f, na = 0.6f;
  int   px = a.per_x, py = a.per_y;
  for (int k = 0; k < a.octaves; k++)
  {
    float v = base_perlin(x * nf, y * nf, a.fseed, px, py);
    n += v * na;
    na *= (1.f - w) + w * fminf(v + 1.f, 2.f) * 0.5f;
    na *= a.persistence;
    nf *= a.lacunarity;
    px = scale_period(px, a.lacunarity);
    py = scale_period(py, a.lacunarity);
  }
  out[idx] = n;
}

int launch_noise_fbm(float       *out,
                     int          nx,
                     int          ny,
                     float        kx,
                     float        ky,
                     uint32_t     seed,
                     int          octaves,
                     float        weight,
                     float        persistence,
                     float        lacunarity,
                     const float *ctrl,
                     const float *noise_x,
                     const float *noise_y,
                     const float *bbox4,
                     const int   *period2,
                     hipStream_t  st)
{
  NoiseArgs a;
  a.nx = nx;
  a.ny = ny;
  a.kx = kx;
  a.ky = ky;
  // fseed = rand(wang_hash(seed)) (noise.cl:483-484, _common_rand.cl:84-88)
  a.fseed = (float)xorshift32(wang_hash(seed)) * (1.f / 4294967296.f);
  a.octaves = octaves;
  a.weight = weight;
  a.persistence = persistence;
  a.lacunarity = lacunarity;
  a.b0 = bbox4 ? bbox4[0] : 0.f;
  a.b1 = bbox4 ? bbox4[1] : 1.f;
  a.b2 = bbox4 ? bbox4[2] : 0.f;
  a.b3 = bbox4 ? bbox4[3] : 1.f;
  a.per_x = period2 ? period2[0] : 0;
  a.per_y = period2 ? period2[1] : 0;
  dim3 blk(64, 4);
  dim3 grd((nx + 63) / 64, (ny + 3) / 4);
  k_noise_fbm_perlin<<<grd, blk, 0, st>>>(out, ctrl, noise_x, noise_y, a);
  return check_launch("noise_fbm");
}

static int check_noise_args(int nx, int ny, int noise_type, int octaves)
{
  if (nx <= 0 || ny <= 0) return set_error(HMG_ERR_INVALID, "shape must be positive");
  if ((size_t)nx * (size_t)ny > ((size_t)1 << 31))
    return set_error(HMG_ERR_INVALID, "grid larger than 2^31 cells");
  if (octaves < 0) return set_error(HMG_ERR_INVALID, "octaves must be >= 0");
  if (noise_type != 1)
    return set_error(HMG_ERR_UNSUPPORTED,
                     "noise type %d not implemented (only PERLIN = 1)", noise_type);
  return HMG_OK;
}

// --- remap on the device: min/max reduction then an affine pass --------

__device__ __forceinline__ float wave_min(float v)
{
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ void __launch_bounds__(256) k_minmax(const float *__restrict__ a, size_t n, float *mm)
{
  float  lo = INFINITY, hi = -INFINITY;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
  {
    float x = a[k];
    lo = fminf(lo, x);
    hi = fmaxf(hi, x);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  __shared__ float slo[4], shi[4];
  int              w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
  {
    slo[w] = lo;
    shi[w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    for (int q = 1; q < 4; q++)
    {
      lo = fminf(lo, slo[q]);
      hi = fmaxf(hi, shi[q]);
    }
    // per-block partials; float min/max do not depend on order
    mm[2 + 2 * blockIdx.x] = lo;
    mm[3 + 2 * blockIdx.x] = hi;
  }
}

__global__ void __launch_bounds__(256) k_remap(float *__restrict__ a,
                                               size_t n,
                                               const float *__restrict__ mm,
                                               int          nblk,
                                               float        vmin,
                                               float        vmax)
{
  __shared__ float s_scale, s_offset, s_flat;
  if (threadIdx.x == 0)
  {
    float lo = mm[2], hi = mm[3];
    for (int b = 1; b < nblk; b++)
    {
      lo = fminf(lo, mm[2 + 2 * b]);
      hi = fmaxf(hi, mm[3 + 2 * b]);
    }
    // range.cpp:342-361: scale = (vmax - vmin)/(max - min),
    // offset = vmin - min*scale; x*scale + offset
    s_flat = lo == hi ? 1.f : 0.f;
    float scale = lo == hi ? 0.f : (vmax - vmin) / (hi - lo);
    s_scale = scale;
    s_offset = vmin - lo * scale;
  }
  __syncthreads();
  float  scale = s_scale, offset = s_offset;
  bool   flat = s_flat != 0.f;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    a[k] = flat ? vmin : a[k] * scale + offset;
}

int launch_minmax(const float *a, size_t n, float *mm, int nblk, hipStream_t st)
{
  k_minmax<<<nblk, 256, 0, st>>>(a, n, mm);
  return check_launch("minmax");
}

} // namespace hmg

using namespace hmg;

extern "C" {

int hmg_noise_fbm_dev(hmg_ctx     *ctx,
                      float       *out_dev,
                      int          nx,
                      int          ny,
                      int          noise_type,
                      float        kx,
                      float        ky,
                      uint32_t     seed,
                      int          octaves,
                      float        weight,
                      float        persistence,
                      float        lacunarity,
                      const float *ctrl_dev,
                      const float *noise_x_dev,
                      const float *noise_y_dev,
                      const float *bbox4,
                      const int   *period2)
{
  clear_error();
  HMG_TRY(check_noise_args(nx, ny, noise_type, octaves));
  if (!out_dev) return set_error(HMG_ERR_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  return launch_noise_fbm(out_dev, nx, ny, kx, ky, seed, octaves, weight, persistence,
                          lacunarity, ctrl_dev, noise_x_dev, noise_y_dev, bbox4, period2,
                          ctx->stream);
}

int hmg_noise_fbm(hmg_ctx     *ctx,
                  float       *out_host,
                  int          nx,
                  int          ny,
                  int          noise_type,
                  float        kx,
                  float        ky,
                  uint32_t     seed,
                  int          octaves,
                  float        weight,
                  float        persistence,
                  float        lacunarity,
                  const float *ctrl_host,
                  const float *noise_x_host,
                  const float *noise_y_host,
                  const float *bbox4,
                  const int   *period2)
{
  clear_error();
  HMG_TRY(check_noise_args(nx, ny, noise_type, octaves));
  if (!out_host) return set_error(HMG_ERR_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  size_t bytes = (size_t)nx * ny * sizeof(float);
  DevBuf out, c, gx, gy;
  HMG_TRY(out.get(ctx->pool, bytes));
  const float *pc = nullptr, *px = nullptr, *py = nullptr;
  if (ctrl_host)
  {
    HMG_TRY(c.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(c.ptr, ctrl_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    pc = c.as<float>();
  }
  if (noise_x_host)
  {
    HMG_TRY(gx.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(gx.ptr, noise_x_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    px = gx.as<float>();
  }
  if (noise_y_host)
  {
    HMG_TRY(gy.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(gy.ptr, noise_y_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    py = gy.as<float>();
  }
  HMG_TRY(launch_noise_fbm(out.as<float>(), nx, ny, kx, ky, seed, octaves, weight,
                           persistence, lacunarity, pc, px, py, bbox4, period2,
                           ctx->stream));
  HMG_HIP(hipMemcpyAsync(out_host, out.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

int hmg_remap_dev(hmg_ctx *ctx, float *a_dev, size_t n, float vmin, float vmax)
{
  clear_error();
  if (!a_dev && n) return set_error(HMG_ERR_INVALID, "null array");
  if (n == 0) return HMG_OK;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  int nblk = (int)((n + 255) / 256);
  if (nblk > 1024) nblk = 1024;
  DevBuf mm;
  HMG_TRY(mm.get(ctx->pool, sizeof(float) * (2 + 2 * (size_t)nblk)));
  k_minmax<<<nblk, 256, 0, ctx->stream>>>(a_dev, n, mm.as<float>());
  HMG_TRY(check_launch("minmax"));
  k_remap<<<nblk, 256, 0, ctx->stream>>>(a_dev, n, mm.as<float>(), nblk, vmin, vmax);
  HMG_TRY(check_launch("remap"));
  // the scratch buffer returns to the pool; keep it alive until the stream
  // has consumed it
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

} // extern "C"
