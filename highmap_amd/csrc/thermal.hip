// thermal.hip -- hmap::gpu::thermal for gfx950.
//
// Reference: kernel `thermal` and `thermal_with_bedrock`
// (HighMap/src/gpu_opencl/kernels/thermal.cl:5-100), border rule
// apply_boundaries (_common_index.cl:106-132), host wrapper
// HighMap/src/erosion/thermal_gpu.cpp:17-96, extrapolate_borders
// (HighMap/src/boundary/boundary.cpp:19-41), apply_with_mask
// (HighMap/include/highmap/operator.hpp:405-418).
//
// The reference updates z in place inside one NDRange, so neighbours read
// either the old or the new value depending on scheduling.  Here every step
// is Jacobi: step t is read, step t+1 is written to the other buffer.
//
// Kernels
//  k_thermal_step    one Jacobi step, one cell per lane, 9-point stencil
//                    served from L1/LDS-backed cache lines.  HBM-bound:
//                    8 B/cell (z read + write), 12 B/cell with a talus plane.
//  k_thermal_blocked T Jacobi steps per launch on an LDS tile with a T-cell
//                    halo (temporal blocking): the tile is read once and
//                    written once per T steps, so HBM traffic falls to
//                    ~8/T B per cell-step plus the halo overhead.
//  k_extrap_cols / k_extrap_rows  extrapolate_borders, columns then rows.
#include "hmg_internal.hpp"

namespace hmg
{

// Neighbour heights in the reference order (di,dj) = (-1,0),(0,1),(0,-1),
// (1,0),(-1,-1),(-1,1),(1,-1),(1,1) from a row-major array with row pitch p.
template <class Get>
__device__ __forceinline__ void th_gather(Get g, float n[8])
{
  n[0] = g(-1, 0);
  n[1] = g(0, 1);
  n[2] = g(0, -1);
  n[3] = g(1, 0);
  n[4] = g(-1, -1);
  n[5] = g(-1, 1);
  n[6] = g(1, -1);
  n[7] = g(1, 1);
}

__global__ void __launch_bounds__(256) k_thermal_step(const float *__restrict__ z,
                                                      float *__restrict__ zn,
                                                      int nx,
                                                      int ny,
                                                      const float *__restrict__ talus,
                                                      float ts,
                                                      const float *__restrict__ bedrock)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nx || j >= ny) return;
  size_t idx = (size_t)j * nx + i;
  float  out;
  if (i == 0)
    out = z[idx + 1];
  else if (i == nx - 1)
    out = z[idx - 1];
  else if (j == 0)
    out = z[idx + nx];
  else if (j == ny - 1)
    out = z[idx - nx];
  else
  {
    float val = z[idx];
    if (bedrock)
    {
      // thermal_with_bedrock (thermal.cl:75-99) ends with
      // z[index] = max(val, z_bedrock), val being the value read at the
      // start, which overwrites the exchange it has just added: the
      // interior result is max(old z, bedrock).
      out = fmaxf(val, bedrock[idx]);
    }
    else
    {
      float n[8];
      th_gather([&](int di, int dj) { return z[(size_t)((long)idx + (long)dj * nx + di)]; }, n);
      out = th_interior(val, n, talus ? talus[idx] : ts);
    }
  }
  zn[idx] = out;
}

// Temporal blocking.  A block owns a TW x TH output tile and loads the tile
// plus a T-cell halo into LDS; after step s (1..T) the LDS cells
// [s, L - s) hold step-s values, so after T steps the tile itself does.  Cells outside the grid are never read: each step only updates
// cells whose 8 neighbours are inside the grid or are border cells, whose
// update is the copy rule.
template <int TW, int TH, int T>
__global__ void __launch_bounds__(256) k_thermal_blocked(const float *__restrict__ z,
                                                         float *__restrict__ zn,
                                                         int   nx,
                                                         int   ny,
                                                         float ts,
                                                         int   steps)
{
  constexpr int LW = TW + 2 * T, LH = TH + 2 * T;
  __shared__ float buf[2][LH][LW];
  const int        i0 = blockIdx.x * TW - T; // global i of buf column 0
  const int        j0 = blockIdx.y * TH - T;

  for (int e = threadIdx.x; e < LW * LH; e += blockDim.x)
  {
    int   lx = e % LW, ly = e / LW;
    int   gi = i0 + lx, gj = j0 + ly;
    float v = 0.f;
    if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) v = z[(size_t)gj * nx + gi];
    buf[0][ly][lx] = v;
  }
  __syncthreads();

  int cur = 0;
  for (int s = 1; s <= steps; s++)
  {
    const int g = s; // cells [s, L - s) are valid after step s
    for (int e = threadIdx.x; e < LW * LH; e += blockDim.x)
    {
      int lx = e % LW, ly = e / LW;
      if (lx < g || lx >= LW - g || ly < g || ly >= LH - g) continue;
      int gi = i0 + lx, gj = j0 + ly;
      if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) continue;
      float out;
      if (gi == 0)
        out = buf[cur][ly][lx + 1];
      else if (gi == nx - 1)
        out = buf[cur][ly][lx - 1];
      else if (gj == 0)
        out = buf[cur][ly + 1][lx];
      else if (gj == ny - 1)
        out = buf[cur][ly - 1][lx];
      else
      {
        float n[8];
        th_gather([&](int di, int dj) { return buf[cur][ly + dj][lx + di]; }, n);
        out = th_interior(buf[cur][ly][lx], n, ts);
      }
      buf[cur ^ 1][ly][lx] = out;
    }
    cur ^= 1;
    __syncthreads();
  }

  for (int e = threadIdx.x; e < TW * TH; e += blockDim.x)
  {
    int lx = T + e % TW, ly = T + e / TW;
    int gi = i0 + lx, gj = j0 + ly;
    if (gi < nx && gj < ny) zn[(size_t)gj * nx + gi] = buf[cur][ly][lx];
  }
}

__global__ void k_extrap_cols(float *__restrict__ z, int nx, int ny)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ny) return;
  float *row = z + (size_t)j * nx;
  row[0] = 2.f * row[1] - row[2];
  row[nx - 1] = 2.f * row[nx - 2] - row[nx - 3];
}

__global__ void k_extrap_rows(float *__restrict__ z, int nx, int ny)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  z[i] = 2.f * z[(size_t)nx + i] - z[(size_t)2 * nx + i];
  z[(size_t)(ny - 1) * nx + i] = 2.f * z[(size_t)(ny - 2) * nx + i] -
                                 z[(size_t)(ny - 3) * nx + i];
}

int launch_extrapolate_borders(float *z, int nx, int ny, hipStream_t st)
{
  k_extrap_cols<<<(ny + 255) / 256, 256, 0, st>>>(z, nx, ny);
  HMG_TRY(check_launch("extrapolate_borders(cols)"));
  k_extrap_rows<<<(nx + 255) / 256, 256, 0, st>>>(z, nx, ny);
  return check_launch("extrapolate_borders(rows)");
}

// Temporal-blocking depth used for uniform talus; 0 selects one step per
// launch.  Exposed for the tuner.
static int g_thermal_block_steps = 8;

// Runs `iterations` Jacobi steps.  z holds the input and receives the
// result; zscratch is a second plane of the same size.
int launch_thermal(hmg_ctx     *ctx,
                   float       *z,
                   float       *zscratch,
                   int          nx,
                   int          ny,
                   const float *talus,
                   float        talus_scalar,
                   int          iterations,
                   const float *bedrock,
                   hipStream_t  st)
{
  (void)ctx;
  float *a = z, *b = zscratch;
  int    done = 0;
  const int T = g_thermal_block_steps;
  if (!talus && !bedrock && T == 8)
  {
    constexpr int TW = 64, TH = 16, TT = 8;
    dim3          grd((nx + TW - 1) / TW, (ny + TH - 1) / TH);
    while (iterations - done >= 2)
    {
      int s = iterations - done < TT ? iterations - done : TT;
      k_thermal_blocked<TW, TH, TT><<<grd, 256, 0, st>>>(a, b, nx, ny, talus_scalar, s);
      HMG_TRY(check_launch("thermal_blocked"));
      float *t = a;
      a = b;
      b = t;
      done += s;
    }
  }
  dim3 blk(64, 4);
  dim3 grd((nx + 63) / 64, (ny + 3) / 4);
  for (; done < iterations; done++)
  {
    k_thermal_step<<<grd, blk, 0, st>>>(a, b, nx, ny, talus, talus_scalar, bedrock);
    HMG_TRY(check_launch("thermal_step"));
    float *t = a;
    a = b;
    b = t;
  }
  if (a != z)
    HMG_HIP(hipMemcpyAsync(z, a, (size_t)nx * ny * sizeof(float), hipMemcpyDeviceToDevice, st));
  return HMG_OK;
}

static int check_thermal_args(int nx, int ny, int iterations)
{
  if (nx < 3 || ny < 3) return set_error(HMG_ERR_INVALID, "thermal needs nx, ny >= 3");
  if ((size_t)nx * (size_t)ny > ((size_t)1 << 31))
    return set_error(HMG_ERR_INVALID, "grid larger than 2^31 cells");
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  return HMG_OK;
}

} // namespace hmg

using namespace hmg;

extern "C" {

int hmg_thermal_set_block_steps(int steps)
{
  if (steps != 0 && steps != 8) return set_error(HMG_ERR_INVALID, "block steps must be 0 or 8");
  g_thermal_block_steps = steps;
  return HMG_OK;
}

int hmg_thermal_dev(hmg_ctx     *ctx,
                    float       *z_dev,
                    int          nx,
                    int          ny,
                    const float *talus_dev,
                    float        talus_scalar,
                    int          iterations,
                    const float *bedrock_dev)
{
  clear_error();
  HMG_TRY(check_thermal_args(nx, ny, iterations));
  if (!z_dev) return set_error(HMG_ERR_INVALID, "null z");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  DevBuf tmp;
  HMG_TRY(tmp.get(ctx->pool, (size_t)nx * ny * sizeof(float)));
  HMG_TRY(launch_thermal(ctx, z_dev, tmp.as<float>(), nx, ny, talus_dev, talus_scalar,
                         iterations, bedrock_dev, ctx->stream));
  HMG_TRY(launch_extrapolate_borders(z_dev, nx, ny, ctx->stream));
  // the scratch plane returns to the pool when this call ends
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

int hmg_thermal(hmg_ctx     *ctx,
                float       *z_host,
                int          nx,
                int          ny,
                const float *talus_host,
                float        talus_scalar,
                int          iterations,
                const float *bedrock_host,
                float       *deposition_host,
                const float *mask_host)
{
  clear_error();
  HMG_TRY(check_thermal_args(nx, ny, iterations));
  if (!z_host) return set_error(HMG_ERR_INVALID, "null z");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  const size_t n = (size_t)nx * ny;
  const size_t bytes = n * sizeof(float);
  DevBuf       zb, tmp, tb, bb;
  HMG_TRY(zb.get(ctx->pool, bytes));
  HMG_TRY(tmp.get(ctx->pool, bytes));
  const float *pt = nullptr, *pb = nullptr;
  HMG_HIP(hipMemcpyAsync(zb.ptr, z_host, bytes, hipMemcpyHostToDevice, ctx->stream));
  if (talus_host)
  {
    HMG_TRY(tb.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(tb.ptr, talus_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    pt = tb.as<float>();
  }
  if (bedrock_host)
  {
    HMG_TRY(bb.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(bb.ptr, bedrock_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    pb = bb.as<float>();
  }
  HMG_TRY(launch_thermal(ctx, zb.as<float>(), tmp.as<float>(), nx, ny, pt, talus_scalar,
                         iterations, pb, ctx->stream));
  HMG_TRY(launch_extrapolate_borders(zb.as<float>(), nx, ny, ctx->stream));

  std::vector<float> z0;
  if (deposition_host || mask_host) z0.assign(z_host, z_host + n);
  HMG_HIP(hipMemcpyAsync(z_host, zb.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HMG_HIP(hipStreamSynchronize(ctx->stream));

  // thermal_gpu.cpp:71: deposition = maximum(z - z_bckp, 0.f), i.e.
  // std::max(dz, 0.f) per cell (range.cpp:253-262)
  if (deposition_host)
    for (size_t k = 0; k < n; k++)
    {
      float dz = z_host[k] - z0[k];
      deposition_host[k] = (dz < 0.f) ? 0.f : dz;
    }
  // apply_with_mask (operator.hpp:405-418) with lerp (math.cpp:183-195):
  // z = z0 * (1 - t) + z * t
  if (mask_host)
    for (size_t k = 0; k < n; k++)
    {
      float t = mask_host[k];
      z_host[k] = z0[k] * (1.f - t) + z_host[k] * t;
    }
  return HMG_OK;
}

} // extern "C"
