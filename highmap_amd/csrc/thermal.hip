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
//  k_thermal_stream  one Jacobi step with a uniform talus: wavefronts stream
//                    strips of 62 columns down the rows, each exchange formed
//                    once for the pair of cells it joins; the last step of a
//                    thermal() call also applies extrapolate_borders.
//                    8 B/cell (z read + write).
//  k_thermal_stream2 two Jacobi steps per launch, the second from the first's
//                    rows kept in registers: a uniform talus (calls of two or
//                    more steps from 2^18 cells; 4 B per cell-step) or, with
//                    PLANE, each cell's own talus (from 4 x 4; 6 B per
//                    cell-step with a plane).
//                    With PLANE, each cell forms its eight exchanges with its
//                    own talus: a talus plane (12 B/cell) or a negative
//                    scalar.
//  k_thermal_step    one Jacobi step, one cell per lane, 9-point stencil, for
//                    thermal_with_bedrock and grids below 4 x 4.
//  k_thermal_blocked T Jacobi steps per launch on an LDS tile with a T-cell
//                    halo (temporal blocking), for grids below 2^18 cells
//                    where launches dominate.
//  k_extrap_cols / k_extrap_rows  extrapolate_borders, columns then rows,
//                    where it is not folded into the last step.
//  k_thermal_auto_stream  one thermal_auto_bedrock Jacobi step, streamed
//                    (each cell's own talus; the bedrock plane updated in
//                    place by the owner of each column), from 4 x 4.
//  k_thermal_auto_step  the same step one cell per lane, below 4 x 4.
//  k_thermal_post    the deposition map and the mask blend after a call.
//  k_sd_pre / k_sd_post / k_sd_finish  sediment_deposition's element-wise
//                    rounds around its thermal sub-iterations.
#include <algorithm>
#include <cfloat>
#include <mutex>

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

// thermal_auto_bedrock (thermal.cl:102-145), one Jacobi step.  bed is the
// per-cell bedrock (read and written by its own cell only), z0 the input of
// the call; `first` marks step 0, where the kernel sets bedrock = -FLT_MAX.
__global__ void __launch_bounds__(256) k_thermal_auto_step(const float *__restrict__ z,
                                                           float *__restrict__ zn,
                                                           int nx,
                                                           int ny,
                                                           const float *__restrict__ talus,
                                                           float ts,
                                                           float *__restrict__ bed,
                                                           const float *__restrict__ z0,
                                                           int first)
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
    const float zb = first ? -FLT_MAX : bed[idx];
    const float val = z[idx];
    out = val;
    if (val >= zb)
    {
      float n[8];
      th_gather([&](int di, int dj) { return z[(size_t)((long)idx + (long)dj * nx + di)]; }, n);
      const float z0v = z0[idx];
      out = fmaxf(fmaxf(th_interior(val, n, talus ? talus[idx] : ts), z0v), zb);
      bed[idx] = fmaxf(z0v, out);
    }
    else if (first)
      bed[idx] = zb;
  }
  zn[idx] = out;
}

// After a thermal call: deposition = max(z - z0, 0) with std::max's rule
// (thermal_gpu.cpp:71, range.cpp:253-262) if dep is given, then
// apply_with_mask's lerp z = z0 * (1 - t) + z * t (operator.hpp:405-418,
// math.cpp:183-195) if mask is given.
__global__ void __launch_bounds__(256) k_thermal_post(float *__restrict__ z,
                                                      const float *__restrict__ z0,
                                                      float *__restrict__ dep,
                                                      const float *__restrict__ mask,
                                                      size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
  {
    const float zk = z[k], z0k = z0[k];
    if (dep)
    {
      const float dz = zk - z0k;
      dep[k] = dz < 0.f ? 0.f : dz;
    }
    if (mask)
    {
      const float t = mask[k];
      z[k] = z0k * (1.f - t) + zk * t;
    }
  }
}

// sediment_deposition (deposition.cpp:19-38) element-wise parts.
__global__ void __launch_bounds__(256) k_sd_pre(float *__restrict__ smap,
                                                const float *__restrict__ z,
                                                float *__restrict__ ztot,
                                                float  step,
                                                size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
  {
    const float sk = smap[k] + step; // smap = smap + deposition_step
    smap[k] = sk;
    ztot[k] = z[k] + sk; // z_tot = z + smap
  }
}

__global__ void __launch_bounds__(256) k_sd_post(float *__restrict__ smap,
                                                 const float *__restrict__ ztot,
                                                 const float *__restrict__ z,
                                                 size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
  {
    const float dz = ztot[k] - z[k]; // smap = maximum(z_tot - z, 0.f)
    smap[k] = dz < 0.f ? 0.f : dz;
  }
}

__global__ void __launch_bounds__(256) k_sd_finish(float *__restrict__ z,
                                                   const float *__restrict__ smap,
                                                   size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    z[k] = z[k] + smap[k]; // z = z + smap
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

// ------------------------------------------------------ streaming step
//
// One Jacobi step with a uniform talus, one launch.  Each wavefront owns a
// strip of 64 columns (62 written, one halo column each side) and a chunk
// of rows, and streams down the rows with the heights in a register ring
// (rows loaded P rows ahead, one load per cell).
//
// The exchange between two cells is antisymmetric: with d = a - b, the
// reference's helper_thermal_exchange gives E(a, b) = -E(b, a) exactly
// (|d| and the threshold are the same from both sides, IEEE negation and
// subtraction are exact).  So each cell forms only the exchanges with its
// right (i+1, j), upper (i, j+1), upper-right and upper-left neighbours,
// and takes the other four, negated, from the lane to its left or right
// (DPP) or from the previous row's registers.  That halves the work of the
// step.  The eight terms are then added in the reference's order
// (-1,0),(0,1),(0,-1),(1,0),(-1,-1),(-1,1),(1,-1),(1,1), a negated term as
// a subtraction.
//
// One exchange is copysign(m, other - self) with
// m = max(0.2 (|other - self| - c T), 0) (divided by 1.414 on diagonals):
// m > 0 exactly when |d| > c T (denormals kept, so a - b > 0 iff a > b), and
// where the reference adds +0 this may add -0, which leaves the running sum
// unchanged (the sum starts as 0 - term and is never -0).  The pair form
// needs one talus >= 0 for every cell (the host checks): a talus plane or a
// negative scalar takes the PLANE form, eight exchanges per cell.
//
// EXTRAP: the step is the last of a thermal() call, and extrapolate_borders
// (boundary.cpp:19-41) is applied in the same launch: border columns of the
// interior rows from the two cells inward (lanes to the right or left), then
// rows 0 and ny - 1 from the two rows inward (kept in registers; row 0 is
// written when row 2 is done).  Needs nx, ny >= 4; a chunk that holds row
// ny - 1 starts at ny - 3 at the latest, and chunks are >= 3 rows, so both
// row pairs are in one wavefront's registers.  Where chunks overlap or the
// last strip is moved left to end at column nx - 1, two wavefronts write the
// same values.

constexpr int TH_OUT = 62; // columns written per wavefront

__device__ __forceinline__ float th_left(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float th_right(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// E(self, other) of an axial neighbour (c = 1).
__device__ __forceinline__ float th_pair_axial(float self, float other, float md)
{
  const float nd = other - self;
  const float m = fmaxf(0.2f * (fabsf(nd) - md), 0.f);
  return copysignf(m, nd);
}

// E(self, other) of a diagonal neighbour (c = 1.414): m / 1.414 through
// div_1414's product and fma correction, which is the correctly rounded
// quotient for m = 0 and every m in [2^-100, inf).  `odd` is set where m is
// outside those (tiny or infinite); the caller then recomputes the row's
// diagonal exchanges with the IEEE division.
template <bool IEEE>
__device__ __forceinline__ float th_pair_diag(float self, float other, float mdd, bool &odd)
{
  constexpr float B = 1.414f;
  constexpr float R = 0x1.6a17e6p-1f; // RN(1 / 1.414f)
  const float     nd = other - self;
  const float     m = fmaxf(0.2f * (fabsf(nd) - mdd), 0.f);
  float           q;
  if (IEEE)
    q = m / B;
  else
  {
    const uint32_t b = __float_as_uint(m);
    odd = odd || (b - 1u < 0x0d7fffffu) || b >= 0x7f800000u; // (0, 2^-100) or inf
    const float q0 = m * R;
    q = fmaf(fmaf(-q0, B, m), R, q0);
  }
  return copysignf(q, nd);
}

// Workgroups (4 wavefront items each) of a streaming launch, rounded up to
// a multiple of 8 so that th_wave_item's XCD-contiguous order applies at
// every size; the extra workgroups hold no item and return at once.
static inline int th_grid(long items) { return (int)(((items + 3) / 4 + 7) & ~7L); }

__device__ __forceinline__ int th_wave_item()
{
  const int G = gridDim.x;
  const int b = (G & 7) ? blockIdx.x : (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  return __builtin_amdgcn_readfirstlane(b * 4 + (threadIdx.x >> 6));
}

// The exchanges of a row (z0) with the row above it (z1): axial and both
// diagonals, the diagonals with the IEEE division for the whole wave when a
// lane holds a quotient outside div_1414's range.
__device__ __forceinline__ void th_row_pairs(
    float z0, float z1, float md, float mdd, float &eU, float &eUR, float &eUL)
{
  bool odd = false;
  eU = th_pair_axial(z0, z1, md);
  eUR = th_pair_diag<false>(z0, th_right(z1), mdd, odd);
  eUL = th_pair_diag<false>(z0, th_left(z1), mdd, odd);
  if (__builtin_expect(__any(odd), 0))
  {
    eUR = th_pair_diag<true>(z0, th_right(z1), mdd, odd);
    eUL = th_pair_diag<true>(z0, th_left(z1), mdd, odd);
  }
}

// One Jacobi step of a row before the border rule: z0, z1 the row and the
// row above it in the step's input; pU/pUR/pUL the exchanges of the row
// below with this one, replaced by this row's with the row above.  The
// eight terms in the reference's order (see above).
__device__ __forceinline__ float th_row_step(
    float z0, float z1, float md, float mdd, float &pU, float &pUR, float &pUL)
{
  float eU, eUR, eUL;
  th_row_pairs(z0, z1, md, mdd, eU, eUR, eUL);
  const float eR = th_pair_axial(z0, th_right(z0), md);
  // (-1, 0) as 0 - term: the empty asm keeps the compiler from folding
  // (0 - a) + b into b - a, which differs from it in the sign of a zero
  float amount = 0.f - th_left(eR); // (-1, 0)
  asm volatile("" : "+v"(amount));
  amount = amount + eU;             // (0, 1)
  amount = amount - pU;             // (0, -1)
  amount = amount + eR;             // (1, 0)
  amount = amount - th_left(pUR);   // (-1, -1)
  amount = amount + eUL;            // (-1, 1)
  amount = amount - th_right(pUL);  // (1, -1)
  amount = amount + eUR;            // (1, 1)
  float step = z0 + amount;
  // formed in every lane, before any EXEC mask of the caller (the lane
  // exchanges above then fold into the additions)
  asm volatile("" : "+v"(step));
  pU = eU;
  pUR = eUR;
  pUL = eUL;
  return step;
}

// apply_boundaries (_common_index.cl:106-132) on row r of a step's output:
// i == 0, i == nx - 1, j == 0, j == ny - 1 in that order copy the inward
// neighbour of the step's input (zm, z0, z1: rows r-1, r, r+1).  Strips
// away from columns 0 and nx - 1 (edge false) skip the column rules.  The
// lane exchange runs in every lane before the selects (a DPP read from a
// lane the EXEC mask disables gives 0).
__device__ __forceinline__ float th_border(
    float out, float zm, float z0, float z1, int r, int ny, bool edge, bool c0, bool cn)
{
  if (edge || r == 0 || r == ny - 1)
  {
    float zl = th_left(z0), zr = th_right(z0);
    asm volatile("" : "+v"(zl), "+v"(zr));
    if (r == 0) out = z1;
    if (r == ny - 1) out = zm;
    out = cn ? zl : out;
    out = c0 ? zr : out;
  }
  return out;
}

// One Jacobi step of a row with each cell's own talus t (a talus plane, or
// a scalar of any sign): all eight exchanges of the cell, each
// copysign(max(0.2 (|d| - c t), 0), other - self) (/ 1.414 on diagonals),
// which equals the reference's exchange for every t (a negative one
// included: every pair is then over the threshold, and other == self gives
// +m as the reference's else branch does) and for NaNs (0); the terms in the
// reference's order, the sum started as 0 + term (never -0).  zm, z0, z1 are
// rows r-1, r, r+1; the lane exchanges fold into the arithmetic.
__device__ __forceinline__ float th_cell_step(float zm, float z0, float z1, float t)
{
  constexpr float B = 1.414f;
  constexpr float R = 0x1.6a17e6p-1f; // RN(1 / 1.414f)
  const float     md = t, mdd = B * t;
  const float     n[8] = {th_left(z0), z1,          zm,           th_right(z0),
                          th_left(zm), th_left(z1), th_right(zm), th_right(z1)};
  float           e[8];
#pragma unroll
  for (int k = 0; k < 4; k++) e[k] = th_pair_axial(z0, n[k], md);
  // diagonals: m / 1.414 by div_1414's product and fma correction when
  // every lane's m is 0 or in [2^-100, inf) (m >= 0: as integers, the
  // minimum of bits - 1 and the maximum of bits bound all four), else by
  // the IEEE division for the whole wave
  float    nd[4], m[4];
  uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
  for (int k = 0; k < 4; k++)
  {
    nd[k] = n[4 + k] - z0;
    m[k] = fmaxf(0.2f * (fabsf(nd[k]) - mdd), 0.f);
    const uint32_t b = __float_as_uint(m[k]);
    lo = min(lo, b - 1u);
    hi = max(hi, b);
  }
  const bool odd = lo < 0x0d7fffffu || hi >= 0x7f800000u;
  if (__builtin_expect(__any(odd), 0))
  {
#pragma unroll
    for (int k = 0; k < 4; k++) e[4 + k] = copysignf(m[k] / B, nd[k]);
  }
  else
  {
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
      const float q0 = m[k] * R;
      e[4 + k] = copysignf(fmaf(fmaf(-q0, B, m[k]), R, q0), nd[k]);
    }
  }
  float amount = 0.f + e[0];
  asm volatile("" : "+v"(amount));
#pragma unroll
  for (int k = 1; k < 8; k++) amount = amount + e[k];
  float step = z0 + amount;
  asm volatile("" : "+v"(step));
  return step;
}

// PLANE: the exchanges per cell with its own talus (th_cell_step): a talus
// plane (talus, 4 more bytes per cell, its rows loaded with the heights) or,
// with talus null, a negative scalar ts.  ZD (uniform talus only): z and zn
// are float2 planes, heights in x and water depths in y; each row load
// brings both, and every cell written gets its depth back unchanged beside
// its new height (16 B per cell).
template <int P, bool EXTRAP, bool PLANE, bool ZD = false>
__global__ void __launch_bounds__(256) k_thermal_stream(const float *__restrict__ z,
                                                        float *__restrict__ zn,
                                                        int   nx,
                                                        int   ny,
                                                        float ts,
                                                        const float *__restrict__ talus,
                                                        int   strips,
                                                        int   chunk_rows,
                                                        int   n_items)
{
  const int item = th_wave_item();
  if (item >= n_items) return;
  const int lane = threadIdx.x & 63;
  const int strip = item % strips, chunk = item / strips;
  // the last strip ends at column nx - 1 (its columns overlap the previous
  // strip's; both write the same values)
  const int      x0 = nx > TH_OUT ? min(strip * TH_OUT - 1, nx - TH_OUT - 1) : -1;
  const int      gx = x0 + lane;
  const uint32_t off = (uint32_t)min(max(gx, 0), nx - 1); // column loaded
  const uint32_t soff = (uint32_t)max(gx, 0);              // column stored
  const bool     writer = lane >= 1 && lane <= TH_OUT && gx < nx;
  const bool     c0 = gx == 0, cn = gx == nx - 1;
  // strips that hold column 0 or nx - 1 take the border rules in every row
  const bool edge_strip = x0 < 0 || x0 + TH_OUT + 1 >= nx;
  int        j0 = chunk * chunk_rows;
  const int  j1 = min(j0 + chunk_rows, ny);
  if (EXTRAP && j1 - j0 < 3) j0 = j1 - 3;
  const float md = ts, mdd = 1.414f * ts;

  // a row's base in scalar registers, the column as a 32-bit lane offset.
  // Ordinary loads and stores: the nontemporal hint made 4096^2 steps take
  // 42 us instead of 31 (the strip edges are read by two wavefronts, and a
  // step's output is the next step's input, held by the 256 MB cache).
  static_assert(!(ZD && PLANE), "the float2 layout takes a uniform talus");
  constexpr int U = P + 3; // rows r-1 .. r+P+1
  float         R[U];
  float         TR[U]; // PLANE: the talus of the rows in R
  float         DR[U]; // ZD: the water depths of the rows in R
  // row r into ring slot q (ZD: the depth beside it)
  auto load_to = [&](int q, int r) {
    const size_t ro = (size_t)min(max(r, 0), ny - 1) * (size_t)nx;
    if constexpr (ZD)
    {
      const float2 v = at_off(reinterpret_cast<const float2 *>(z) + ro, off);
      R[q] = v.x;
      DR[q] = v.y;
    }
    else
      R[q] = at_off(z + ro, off);
  };
  auto load_t = [&](int r) {
    if (!talus) return ts;
    const float *row = talus + (size_t)min(max(r, 0), ny - 1) * (size_t)nx;
    return at_off(row, off);
  };
  // store of a height at (row r, this lane's column) with depth d (ZD)
  auto store = [&](int r, float v, float d) {
    if constexpr (ZD)
      at_off(reinterpret_cast<float2 *>(zn) + (size_t)r * (size_t)nx, soff) = make_float2(v, d);
    else
      at_off(zn + (size_t)r * (size_t)nx, soff) = v;
  };

  // pre-roll: rows j0-1 .. j0+P, and the exchanges of row j0-1 with row j0
#pragma unroll
  for (int q = 0; q <= P + 1; q++) load_to(q, j0 - 1 + q);
  if (PLANE)
  {
    TR[0] = ts; // row j0 - 1: no step of it here
#pragma unroll
    for (int q = 1; q <= P + 1; q++) TR[q] = load_t(j0 - 1 + q);
  }
  float pU = 0.f, pUR = 0.f, pUL = 0.f; // exchanges of the previous row with the row above it
  if (!PLANE) th_row_pairs(R[0], R[1], md, mdd, pU, pUR, pUL);
  float f1 = 0.f, f2 = 0.f; // EXTRAP: final values of rows r-1 and r-2
  float d_row0 = 0.f;        // EXTRAP, ZD: the depth of row 0, written with row 2

  for (int r0 = j0; r0 < j1; r0 += U)
  {
#pragma unroll
    for (int q = 0; q < U; q++)
    {
      const int r = r0 + q;
      if (r >= j1) break;
      const float zm = R[q % U], z0 = R[(q + 1) % U], z1 = R[(q + 2) % U];
      const float dc = ZD ? DR[(q + 1) % U] : 0.f; // depth of row r
      // the row P + 1 below the next one, into the slot of row r - 1
      if (r + P + 1 <= j1)
      {
        load_to((q + P + 2) % U, r + P + 1);
        if (PLANE) TR[(q + P + 2) % U] = load_t(r + P + 1);
      }

      const float step = PLANE ? th_cell_step(zm, z0, z1, TR[(q + 1) % U])
                               : th_row_step(z0, z1, md, mdd, pU, pUR, pUL);

      if (!EXTRAP)
      {
        const float out = th_border(step, zm, z0, z1, r, ny, edge_strip, c0, cn);
        if (writer) store(r, out, dc);
      }
      else
      {
        // extrapolate_borders: columns (interior rows), then rows
        float f = step;
        if (edge_strip)
        {
          float s1r = th_right(step), s1l = th_left(step);
          asm volatile("" : "+v"(s1r), "+v"(s1l));
          float s2r = th_right(s1r), s2l = th_left(s1l);
          asm volatile("" : "+v"(s2r), "+v"(s2l));
          f = cn ? 2.f * s1l - s2l : f;
          f = c0 ? 2.f * s1r - s2r : f;
        }
        if (r == ny - 1) f = 2.f * f1 - f2;
        if (r == 0) d_row0 = dc;
        if (r == 2 && j0 == 0 && writer) store(0, 2.f * f1 - f, d_row0);
        if (r != 0 && writer) store(r, f, dc); // row 0 is written with row 2
        f2 = f1;
        f1 = f;
      }
    }
  }
}

// Launch geometry of k_thermal_stream: strips of 62 columns and chunks sized
// so that the work items just fill the device's wavefront slots (one round).
// ------------------------------------------------- two steps per launch
//
// k_thermal_stream2: two Jacobi steps with a uniform talus in one launch.
// Each wavefront streams its rows once: step 1 of row r from the input rows
// r-1..r+1 (loaded P rows ahead), then step 2 of row r-1 from step 1's
// rows r-2..r, held in registers, so a step costs half a read and half a
// write per cell.  Step 1 is valid in lanes 1..62 and step 2 in lanes
// 2..61: a strip writes 60 columns (the last strip moved to end at column
// nx - 1).  Each step applies the border rule to its own output, so step 2
// reads exactly the field that a separate launch would have written.  Rows
// above row 0 and below row ny - 1 come from clamped loads; they reach only
// values that the border rule replaces.  A chunk of rows j0..j1-1 runs step
// 1 on rows j0-1..j1.

constexpr int TH2_OUT = 60; // columns written per wavefront

// PLANE: each cell's own talus (th_cell_step), from a talus plane (rows
// loaded with the heights, a ring beside theirs) or, with talus null, a
// negative scalar ts.
template <int P, bool PLANE>
__global__ void __launch_bounds__(256) k_thermal_stream2(const float *__restrict__ z,
                                                         float *__restrict__ zn,
                                                         int   nx,
                                                         int   ny,
                                                         float ts,
                                                         const float *__restrict__ talus,
                                                         int   strips,
                                                         int   chunk_rows,
                                                         int   n_items)
{
  const int item = th_wave_item();
  if (item >= n_items) return;
  const int      lane = threadIdx.x & 63;
  const int      strip = item % strips, chunk = item / strips;
  const int      x0 = nx > TH2_OUT ? min(strip * TH2_OUT - 2, nx - TH2_OUT - 2) : -2;
  const int      gx = x0 + lane;
  const uint32_t off = (uint32_t)min(max(gx, 0), nx - 1); // column loaded
  const uint32_t soff = (uint32_t)max(gx, 0);              // column stored
  const bool     writer = lane >= 2 && lane < 2 + TH2_OUT && gx < nx;
  const bool     c0 = gx == 0, cn = gx == nx - 1;
  const bool     edge_strip = x0 < 0 || x0 + TH2_OUT + 2 >= nx;
  const int      j0 = chunk * chunk_rows;
  const int      j1 = min(j0 + chunk_rows, ny);
  const float    md = ts, mdd = 1.414f * ts;
  auto           load = [&](int r) {
    const float *row = z + (size_t)min(max(r, 0), ny - 1) * (size_t)nx;
    return at_off(row, off);
  };
  auto load_t = [&](int r) {
    if (!talus) return ts;
    const float *row = talus + (size_t)min(max(r, 0), ny - 1) * (size_t)nx;
    return at_off(row, off);
  };

  // rings of U slots indexed by compile-time constants (rows rotate by
  // renaming): with step-1 row r, Z holds input rows r-1 .. r+P+1 (TR
  // their talus, PLANE) and S1 step-1 rows r-2 .. r
  constexpr int U = P + 3;
  float         Z[U], S1[U], TR[U];
  float         pU = 0.f, pUR = 0.f, pUL = 0.f; // step 1: exchanges of input row r-1 with row r
  float         qU = 0.f, qUR = 0.f, qUL = 0.f; // step 2: exchanges of step-1 row r-2 with row r-1
  {
    // pre-roll: step 1 of rows j0-1 and j0, step 2's exchanges between them
    const float za = load(j0 - 2), zb = load(j0 - 1);
#pragma unroll
    for (int q = 0; q <= P + 1; q++) Z[q] = load(j0 + q);
    if (PLANE)
    {
#pragma unroll
      for (int q = 0; q <= P + 1; q++) TR[q] = load_t(j0 + q);
      TR[U - 1] = load_t(j0 - 1); // row j0 - 1, its slot in the ring
      float s = th_cell_step(za, zb, Z[0], TR[U - 1]);
      S1[0] = th_border(s, za, zb, Z[0], j0 - 1, ny, edge_strip, c0, cn);
      s = th_cell_step(zb, Z[0], Z[1], TR[0]);
      S1[1] = th_border(s, zb, Z[0], Z[1], j0, ny, edge_strip, c0, cn);
    }
    else
    {
      th_row_pairs(za, zb, md, mdd, pU, pUR, pUL);
      float s = th_row_step(zb, Z[0], md, mdd, pU, pUR, pUL);
      S1[0] = th_border(s, za, zb, Z[0], j0 - 1, ny, edge_strip, c0, cn);
      s = th_row_step(Z[0], Z[1], md, mdd, pU, pUR, pUL);
      S1[1] = th_border(s, zb, Z[0], Z[1], j0, ny, edge_strip, c0, cn);
      th_row_pairs(S1[0], S1[1], md, mdd, qU, qUR, qUL);
    }
  }

  for (int r0 = j0 + 1; r0 <= j1; r0 += U)
  {
#pragma unroll
    for (int q = 0; q < U; q++)
    {
      const int r = r0 + q;
      if (r > j1) break;
      const float zm = Z[q % U], z0 = Z[(q + 1) % U], z1 = Z[(q + 2) % U];
      // step 1 of row r and step 2 of row r-1 take the talus of rows r and
      // r-1 (PLANE), in slots the load below does not replace
      const float tr = TR[(q + 1) % U], trm = TR[q % U];
      // input row r+P+1 into the slot of row r-2 (the last needed is j1+1)
      if (r + P + 1 <= j1 + 1)
      {
        Z[(q + P + 2) % U] = load(r + P + 1);
        if (PLANE) TR[(q + P + 2) % U] = load_t(r + P + 1);
      }
      // step 1 of row r
      float s = PLANE ? th_cell_step(zm, z0, z1, tr) : th_row_step(z0, z1, md, mdd, pU, pUR, pUL);
      s = th_border(s, zm, z0, z1, r, ny, edge_strip, c0, cn);
      S1[(q + 2) % U] = s;
      // step 2 of row r-1
      const float a = S1[q % U], b = S1[(q + 1) % U];
      float       o = PLANE ? th_cell_step(a, b, s, trm) : th_row_step(b, s, md, mdd, qU, qUR, qUL);
      o = th_border(o, a, b, s, r - 1, ny, edge_strip, c0, cn);
      if (writer) at_off(zn + (size_t)(r - 1) * (size_t)nx, soff) = o;
    }
  }
}

template <int P, bool PLANE = false>
static int launch_thermal_stream2(const float *a,
                                  float       *b,
                                  int          nx,
                                  int          ny,
                                  float        ts,
                                  int          num_cus,
                                  hipStream_t  st,
                                  const float *talus = nullptr)
{
  static const int blocks_per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_thermal_stream2<P, PLANE>, 256, 0) !=
            hipSuccess ||
        n < 1)
    {
      (void)hipGetLastError();
      n = 2;
    }
    return n;
  }();
  const long strips = nx > TH2_OUT ? (nx + TH2_OUT - 1) / TH2_OUT : 1;
  const long slots = 2 * 4L * blocks_per_cu * (num_cus > 0 ? num_cus : 256);
  long       ch = (strips * ny + slots - 1) / slots;
  ch = ch < 3 ? 3 : ch;
  const long chunks = (ny + ch - 1) / ch;
  const long items = strips * chunks;
  const int  G = th_grid(items);
  k_thermal_stream2<P, PLANE>
      <<<G, 256, 0, st>>>(a, b, nx, ny, ts, talus, (int)strips, (int)ch, (int)items);
  return check_launch("thermal_stream2");
}

template <int P, bool EXTRAP, bool PLANE = false, bool ZD = false>
static int launch_thermal_stream(const float *a,
                                 float       *b,
                                 int          nx,
                                 int          ny,
                                 float        ts,
                                 int          num_cus,
                                 hipStream_t  st,
                                 const float *talus = nullptr)
{
  static const int blocks_per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_thermal_stream<P, EXTRAP, PLANE, ZD>,
                                                     256, 0) !=
            hipSuccess ||
        n < 1)
    {
      (void)hipGetLastError();
      n = 2;
    }
    return n;
  }();
  const long strips = (nx + TH_OUT - 1) / TH_OUT;
  // two rounds of wavefronts: 4096^2 takes 30.6-31.1 us per step with one,
  // two or four rounds (two rows loaded ahead), 32.7 with eight
  const long slots = 2 * 4L * blocks_per_cu * (num_cus > 0 ? num_cus : 256);
  long       ch = (strips * ny + slots - 1) / slots;
  ch = ch < 3 ? 3 : ch;
  const long chunks = (ny + ch - 1) / ch;
  const long items = strips * chunks;
  const int  G = th_grid(items);
  k_thermal_stream<P, EXTRAP, PLANE, ZD>
      <<<G, 256, 0, st>>>(a, b, nx, ny, ts, talus, (int)strips, (int)ch, (int)items);
  return check_launch("thermal_stream");
}

// thermal_auto_bedrock (thermal.cl:102-145), one Jacobi step, streamed as
// k_thermal_stream<P, false, true>: each cell's eight exchanges with its own
// talus (th_cell_step), then the bedrock rule of the interior cells:
// zb = -FLT_MAX on the first step, else the cell's bedrock; where
// z >= zb the height becomes max(max(z + exchanges, z0), zb) and the
// bedrock max(z0, height), else the height stays (and on the first step the
// bedrock becomes zb).  The bedrock plane is read and written in place, by
// its own cell only, so each column has one owner: strip s writes columns
// 62 s .. 62 s + 61 (the last strip, moved left to end at column nx - 1,
// only those past the previous strip's), and chunks do not overlap.
template <int P>
__global__ void __launch_bounds__(256) k_thermal_auto_stream(const float *__restrict__ z,
                                                             float *__restrict__ zn,
                                                             int   nx,
                                                             int   ny,
                                                             float ts,
                                                             const float *__restrict__ talus,
                                                             float *__restrict__ bed,
                                                             const float *__restrict__ zinit,
                                                             int   first,
                                                             int   strips,
                                                             int   chunk_rows,
                                                             int   n_items)
{
  const int item = th_wave_item();
  if (item >= n_items) return;
  const int      lane = threadIdx.x & 63;
  const int      strip = item % strips, chunk = item / strips;
  const int      x0 = nx > TH_OUT ? min(strip * TH_OUT - 1, nx - TH_OUT - 1) : -1;
  const int      gx = x0 + lane;
  const uint32_t off = (uint32_t)min(max(gx, 0), nx - 1); // column loaded
  const uint32_t soff = (uint32_t)max(gx, 0);              // column stored
  const bool     writer =
      lane >= 1 && lane <= TH_OUT && gx < nx && gx >= strip * TH_OUT; // the column's owner
  const bool     c0 = gx == 0, cn = gx == nx - 1;
  const bool     edge_strip = x0 < 0 || x0 + TH_OUT + 1 >= nx;
  const int      j0 = chunk * chunk_rows;
  const int      j1 = min(j0 + chunk_rows, ny);
  auto           load = [&](const float *p, int r) {
    return at_off(p + (size_t)min(max(r, 0), ny - 1) * (size_t)nx, off);
  };
  auto load_t = [&](int r) { return talus ? load(talus, r) : ts; };

  constexpr int U = P + 3; // rows r-1 .. r+P+1
  float         R[U], TR[U];
#pragma unroll
  for (int q = 0; q <= P + 1; q++) R[q] = load(z, j0 - 1 + q);
  TR[0] = ts; // row j0 - 1: no step of it here
#pragma unroll
  for (int q = 1; q <= P + 1; q++) TR[q] = load_t(j0 - 1 + q);

  for (int r0 = j0; r0 < j1; r0 += U)
  {
#pragma unroll
    for (int q = 0; q < U; q++)
    {
      const int r = r0 + q;
      if (r >= j1) break;
      const float zm = R[q % U], zc = R[(q + 1) % U], z1 = R[(q + 2) % U];
      const float t = TR[(q + 1) % U];
      if (r + P + 1 <= j1)
      {
        R[(q + P + 2) % U] = load(z, r + P + 1);
        TR[(q + P + 2) % U] = load_t(r + P + 1);
      }
      float       *brow = bed + (size_t)r * (size_t)nx;
      const float  zb = first ? -FLT_MAX : at_off((const float *)brow, off);
      const float  zi = load(zinit, r);
      const float  step = th_cell_step(zm, zc, z1, t);
      const bool   up = zc >= zb;
      float        out = up ? fmaxf(fmaxf(step, zi), zb) : zc;
      out = th_border(out, zm, zc, z1, r, ny, edge_strip, c0, cn);
      const bool interior = !c0 && !cn && r != 0 && r != ny - 1;
      if (writer)
      {
        at_off(zn + (size_t)r * (size_t)nx, soff) = out;
        if (interior && (up || first)) at_off(brow, soff) = up ? fmaxf(zi, out) : zb;
      }
    }
  }
}

template <int P>
static int launch_thermal_auto_stream(const float *a,
                                      float       *b,
                                      int          nx,
                                      int          ny,
                                      float        ts,
                                      const float *talus,
                                      float       *bed,
                                      const float *zinit,
                                      bool         first,
                                      int          num_cus,
                                      hipStream_t  st)
{
  static const int blocks_per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_thermal_auto_stream<P>, 256, 0) !=
            hipSuccess ||
        n < 1)
    {
      (void)hipGetLastError();
      n = 2;
    }
    return n;
  }();
  const long strips = (nx + TH_OUT - 1) / TH_OUT;
  const long slots = 2 * 4L * blocks_per_cu * (num_cus > 0 ? num_cus : 256);
  long       ch = (strips * ny + slots - 1) / slots;
  ch = ch < 3 ? 3 : ch;
  const long chunks = (ny + ch - 1) / ch;
  const long items = strips * chunks;
  const int  G = th_grid(items);
  k_thermal_auto_stream<P><<<G, 256, 0, st>>>(a, b, nx, ny, ts, talus, bed, zinit, first ? 1 : 0,
                                              (int)strips, (int)ch, (int)items);
  return check_launch("thermal_auto_stream");
}

// Kernel choice for uniform talus: 8, the LDS-blocked kernel (8 steps per
// launch) below 2^18 cells and the streaming kernels, two steps per launch,
// from there; 2, the streaming kernels, two steps per launch, at every
// size; 0, the one-step streaming kernel at every size.
static int g_thermal_block_steps = 8;

// Runs `iterations` Jacobi steps, then (extrapolate) extrapolate_borders.
// z holds the input; zscratch is a second plane of the same size (see the
// declaration for where the result ends).
int launch_thermal(hmg_ctx     *ctx,
                   float       *z,
                   float       *zscratch,
                   int          nx,
                   int          ny,
                   const float *talus,
                   float        talus_scalar,
                   int          iterations,
                   const float *bedrock,
                   hipStream_t  st,
                   float      **result,
                   bool         extrapolate)
{
  float    *a = z, *b = zscratch;
  int       done = 0;
  bool      borders_done = false;
  const int T = g_thermal_block_steps;
  const int cus = ctx && ctx->num_cus > 0 ? ctx->num_cus : 256;
  // The blocked kernel recomputes an 8-cell halo per tile; it pays where
  // launches dominate (below 2^18 cells).  From there the streaming kernels
  // take two steps per launch (one for a last odd step).
  const bool uniform = !talus && !bedrock && talus_scalar >= 0.f;
  const bool blocked_pays = (long)nx * ny < (1L << 18);
  if (uniform && T == 8 && blocked_pays)
  {
    // 32 x 8 tiles: at these sizes they give every CU a tile or more
    constexpr int TT = 8;
    while (iterations - done >= 2)
    {
      int s = iterations - done < TT ? iterations - done : TT;
      k_thermal_blocked<32, 8, TT>
          <<<dim3((nx + 31) / 32, (ny + 7) / 8), 256, 0, st>>>(a, b, nx, ny, talus_scalar, s);
      HMG_TRY(check_launch("thermal_blocked"));
      float *t = a;
      a = b;
      b = t;
      done += s;
    }
  }
  const bool stream = uniform;
  // a talus plane or a negative scalar, without bedrock: the streaming
  // kernel with each cell's own talus
  const bool plane_stream = !uniform && !bedrock && nx >= 4 && ny >= 4;
  dim3       blk(64, 4);
  dim3       grd((nx + 63) / 64, (ny + 3) / 4);
  // two steps per launch while two or more are left (an odd last step, or
  // a lone one, runs below; after an even count the borders are
  // extrapolated by their own launches)
  const bool pairs = (stream || plane_stream) && T != 0 && nx >= 4 && ny >= 4;
  while (pairs && iterations - done >= 2)
  {
    if (plane_stream)
      HMG_TRY((launch_thermal_stream2<2, true>(a, b, nx, ny, talus_scalar, cus, st, talus)));
    else
      HMG_TRY((launch_thermal_stream2<2>(a, b, nx, ny, talus_scalar, cus, st)));
    float *t = a;
    a = b;
    b = t;
    done += 2;
  }
  for (; done < iterations; done++)
  {
    const bool last = done == iterations - 1;
    // two rows loaded ahead: 4096^2 steps take 30.6 us, against 31.9 with
    // four rows and 44-47 with eight or twelve (more registers, no faster
    // memory)
    if (stream && last && extrapolate && nx >= 4 && ny >= 4)
    {
      HMG_TRY((launch_thermal_stream<2, true>(a, b, nx, ny, talus_scalar, cus, st)));
      borders_done = true;
    }
    else if (stream)
      HMG_TRY((launch_thermal_stream<2, false>(a, b, nx, ny, talus_scalar, cus, st)));
    else if (plane_stream && last && extrapolate)
    {
      HMG_TRY((launch_thermal_stream<2, true, true>(a, b, nx, ny, talus_scalar, cus, st, talus)));
      borders_done = true;
    }
    else if (plane_stream)
      HMG_TRY((launch_thermal_stream<2, false, true>(a, b, nx, ny, talus_scalar, cus, st, talus)));
    else
    {
      k_thermal_step<<<grd, blk, 0, st>>>(a, b, nx, ny, talus, talus_scalar, bedrock);
      HMG_TRY(check_launch("thermal_step"));
    }
    float *t = a;
    a = b;
    b = t;
  }
  if (extrapolate && !borders_done) HMG_TRY(launch_extrapolate_borders(a, nx, ny, st));
  if (result)
    *result = a;
  else if (a != z)
    HMG_HIP(hipMemcpyAsync(z, a, (size_t)nx * ny * sizeof(float), hipMemcpyDeviceToDevice, st));
  return HMG_OK;
}

// ------------------------------------------- heights beside water depths
//
// Sessions of FP32 state in the fused form keep the heights and the water
// depths as one float2 plane (one 8-byte stream each way per cell in the
// hydraulic kernel).  Thermal steps on them stream the float2 rows
// (k_thermal_stream<2, ·, false, true>), carrying the depths over; other
// cases split the heights out, run launch_thermal and merge them back.

__global__ void __launch_bounds__(256) k_zd_split(const float2 *__restrict__ zd,
                                                  float *__restrict__ out,
                                                  int    component,
                                                  size_t n)
{
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
  {
    const float2 v = zd[c];
    out[c] = component ? v.y : v.x;
  }
}

__global__ void __launch_bounds__(256) k_zd_merge(float2 *__restrict__ zd,
                                                  const float *__restrict__ in,
                                                  int    component,
                                                  size_t n)
{
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
  {
    float2 v = zd[c];
    if (component)
      v.y = in[c];
    else
      v.x = in[c];
    zd[c] = v;
  }
}

int launch_zd_split(const float2 *zd, float *out, int component, size_t n, hipStream_t st)
{
  const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  if (nb > 0) k_zd_split<<<nb, 256, 0, st>>>(zd, out, component, n);
  return check_launch("zd_split");
}

int launch_zd_merge(float2 *zd, const float *in, int component, size_t n, hipStream_t st)
{
  const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  if (nb > 0) k_zd_merge<<<nb, 256, 0, st>>>(zd, in, component, n);
  return check_launch("zd_merge");
}

int launch_thermal_zd(hmg_ctx     *ctx,
                      float2      *zd,
                      float2      *zd_scratch,
                      int          nx,
                      int          ny,
                      float        talus_scalar,
                      int          iterations,
                      hipStream_t  st,
                      float2     **result)
{
  const int cus = ctx && ctx->num_cus > 0 ? ctx->num_cus : 256;
  // the pair form needs a talus >= 0 (a NaN fails the test)
  if (talus_scalar >= 0.f && nx >= 4 && ny >= 4 && iterations > 0)
  {
    float2 *a = zd, *b = zd_scratch;
    for (int done = 0; done < iterations; done++)
    {
      const float *ia = reinterpret_cast<const float *>(a);
      float       *ob = reinterpret_cast<float *>(b);
      if (done == iterations - 1)
        HMG_TRY((launch_thermal_stream<2, true, false, true>(ia, ob, nx, ny, talus_scalar, cus,
                                                             st)));
      else
        HMG_TRY((launch_thermal_stream<2, false, false, true>(ia, ob, nx, ny, talus_scalar, cus,
                                                              st)));
      std::swap(a, b);
    }
    *result = a;
    return HMG_OK;
  }
  const size_t n = (size_t)nx * ny;
  DevBuf       h, hs;
  HMG_TRY(h.get(ctx->pool, n * 4));
  HMG_TRY(hs.get(ctx->pool, n * 4));
  HMG_TRY(launch_zd_split(zd, h.as<float>(), 0, n, st));
  float *res = nullptr;
  HMG_TRY(launch_thermal(ctx, h.as<float>(), hs.as<float>(), nx, ny, nullptr, talus_scalar,
                         iterations, nullptr, st, &res, true));
  HMG_TRY(launch_zd_merge(zd, res, 0, n, st));
  // the pool planes return when this call ends: wait for the merge first
  HMG_HIP(hipStreamSynchronize(st));
  *result = zd;
  return HMG_OK;
}

// thermal_auto_bedrock steps: z holds the input (it is also z0, copied to
// z0buf first) and receives the result; zscratch and bed are planes of the
// same size.
int launch_thermal_auto(float       *z,
                        float       *zscratch,
                        float       *bed,
                        float       *z0buf,
                        int          nx,
                        int          ny,
                        const float *talus,
                        float        talus_scalar,
                        int          iterations,
                        hipStream_t  st,
                        int          num_cus)
{
  const size_t bytes = (size_t)nx * ny * sizeof(float);
  HMG_HIP(hipMemcpyAsync(z0buf, z, bytes, hipMemcpyDeviceToDevice, st));
  float *a = z, *b = zscratch;
  dim3   blk(64, 4);
  dim3   grd((nx + 63) / 64, (ny + 3) / 4);
  // the streaming step from 4 x 4 (one thread per cell below)
  const bool stream = nx >= 4 && ny >= 4;
  for (int it = 0; it < iterations; it++)
  {
    if (stream)
      HMG_TRY((launch_thermal_auto_stream<2>(a, b, nx, ny, talus_scalar, talus, bed, z0buf,
                                             it == 0, num_cus, st)));
    else
    {
      k_thermal_auto_step<<<grd, blk, 0, st>>>(a, b, nx, ny, talus, talus_scalar, bed, z0buf,
                                               it == 0 ? 1 : 0);
      HMG_TRY(check_launch("thermal_auto_step"));
    }
    float *t = a;
    a = b;
    b = t;
  }
  if (a != z) HMG_HIP(hipMemcpyAsync(z, a, bytes, hipMemcpyDeviceToDevice, st));
  return HMG_OK;
}

// ------------------------------------------------- allocation-free scratch
//
// hmg_thermal_dev_noalloc's second Jacobi plane: a static device array, one
// per GPU the library runs on, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (on the host it names a stand-in, not
// device memory): its device address is looked up once per device, with that
// device current, when a context is created on it, and kept in g_scratch_ptr.
constexpr size_t NOALLOC_MAX_CELLS = (size_t)4096 * 4096; // 64 MiB of float
static __device__ __align__(256) float scratch[NOALLOC_MAX_CELLS];

constexpr int     MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float     *g_scratch_ptr[MAX_DEVICES] = {};

int noalloc_scratch_setup(int device)
{
  if (device < 0 || device >= MAX_DEVICES)
    return set_error(HMG_ERR_INVALID, "device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return HMG_OK;
  HMG_HIP(hipSetDevice(device));
  void *p = nullptr;
  HMG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<float *>(p);
  return HMG_OK;
}

static float *noalloc_scratch(int device)
{
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

static int post_blocks(size_t n)
{
  return (int)std::min<size_t>((n + 255) / 256, 4096);
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
  if (steps != 0 && steps != 2 && steps != 8)
    return set_error(HMG_ERR_INVALID, "block steps must be 0, 2 or 8");
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
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_dev) return set_error(HMG_ERR_INVALID, "null z");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  DevBuf tmp;
  HMG_TRY(tmp.get(ctx->pool, (size_t)nx * ny * sizeof(float)));
  HMG_TRY(launch_thermal(ctx, z_dev, tmp.as<float>(), nx, ny, talus_dev, talus_scalar,
                         iterations, bedrock_dev, ctx->stream, nullptr, true));
  // the scratch plane returns to the pool when this call ends
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

int hmg_thermal_dev_noalloc(hmg_ctx     *ctx,
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
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_dev) return set_error(HMG_ERR_INVALID, "null z");
  if ((size_t)nx * (size_t)ny > NOALLOC_MAX_CELLS)
    return set_error(HMG_ERR_INVALID,
                     "hmg_thermal_dev_noalloc: %d x %d is above its %zu-cell scratch plane", nx,
                     ny, (size_t)NOALLOC_MAX_CELLS);
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  float *zs = noalloc_scratch(ctx->device);
  if (!zs) return set_error(HMG_ERR_DEVICE, "no scratch plane on device %d", ctx->device);
  // the Jacobi ping-pong runs through the device's static plane; an odd
  // number of launches copies the result back into z_dev (a device copy)
  return launch_thermal(ctx, z_dev, zs, nx, ny, talus_dev, talus_scalar, iterations,
                        bedrock_dev, ctx->stream, nullptr, true);
}

// Host-pointer form shared by thermal and thermal_auto_bedrock: upload,
// steps, extrapolate_borders, then deposition and mask on the device.
static int thermal_host_call(hmg_ctx     *ctx,
                             bool         auto_bedrock,
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
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  hipStream_t  st = ctx->stream;
  const size_t n = (size_t)nx * ny;
  const size_t bytes = n * sizeof(float);
  DevBuf       zb, z0, tmp, tb, bb, db, mb;
  HMG_TRY(zb.get(ctx->pool, bytes));
  HMG_TRY(z0.get(ctx->pool, bytes));
  HMG_TRY(tmp.get(ctx->pool, bytes));
  const float *pt = nullptr, *pb = nullptr, *pm = nullptr;
  float       *pd = nullptr;
  HMG_HIP(hipMemcpyAsync(zb.ptr, z_host, bytes, hipMemcpyHostToDevice, st));
  if (talus_host)
  {
    HMG_TRY(tb.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(tb.ptr, talus_host, bytes, hipMemcpyHostToDevice, st));
    pt = tb.as<float>();
  }
  if (bedrock_host)
  {
    HMG_TRY(bb.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(bb.ptr, bedrock_host, bytes, hipMemcpyHostToDevice, st));
    pb = bb.as<float>();
  }
  if (mask_host)
  {
    HMG_TRY(mb.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(mb.ptr, mask_host, bytes, hipMemcpyHostToDevice, st));
    pm = mb.as<float>();
  }
  if (deposition_host)
  {
    HMG_TRY(db.get(ctx->pool, bytes));
    pd = db.as<float>();
  }
  if (auto_bedrock)
  {
    DevBuf bed;
    HMG_TRY(bed.get(ctx->pool, bytes));
    HMG_TRY(launch_thermal_auto(zb.as<float>(), tmp.as<float>(), bed.as<float>(), z0.as<float>(),
                                nx, ny, pt, talus_scalar, iterations, st, ctx->num_cus));
    HMG_TRY(launch_extrapolate_borders(zb.as<float>(), nx, ny, st));
    HMG_HIP(hipStreamSynchronize(st)); // bed returns to the pool here
  }
  else
  {
    HMG_HIP(hipMemcpyAsync(z0.ptr, zb.ptr, bytes, hipMemcpyDeviceToDevice, st));
    HMG_TRY(launch_thermal(ctx, zb.as<float>(), tmp.as<float>(), nx, ny, pt, talus_scalar,
                           iterations, pb, st, nullptr, true));
  }
  if (pd || pm)
  {
    k_thermal_post<<<post_blocks(n), 256, 0, st>>>(zb.as<float>(), z0.as<float>(), pd, pm, n);
    HMG_TRY(check_launch("thermal_post"));
  }
  HMG_HIP(hipMemcpyAsync(z_host, zb.ptr, bytes, hipMemcpyDeviceToHost, st));
  if (pd) HMG_HIP(hipMemcpyAsync(deposition_host, pd, bytes, hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  // the result is delivered either way; a non-finite one is reported
  return check_finite_f32(ctx, zb.as<float>(), n);
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
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_host) return set_error(HMG_ERR_INVALID, "null z");
  return thermal_host_call(ctx, false, z_host, nx, ny, talus_host, talus_scalar, iterations,
                           bedrock_host, deposition_host, mask_host);
}

int hmg_thermal_auto_bedrock(hmg_ctx     *ctx,
                             float       *z_host,
                             int          nx,
                             int          ny,
                             const float *talus_host,
                             float        talus_scalar,
                             int          iterations,
                             float       *deposition_host,
                             const float *mask_host)
{
  clear_error();
  HMG_TRY(check_thermal_args(nx, ny, iterations));
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_host) return set_error(HMG_ERR_INVALID, "null z");
  return thermal_host_call(ctx, true, z_host, nx, ny, talus_host, talus_scalar, iterations,
                           nullptr, deposition_host, mask_host);
}

int hmg_thermal_auto_bedrock_dev(hmg_ctx     *ctx,
                                 float       *z_dev,
                                 int          nx,
                                 int          ny,
                                 const float *talus_dev,
                                 float        talus_scalar,
                                 int          iterations)
{
  clear_error();
  HMG_TRY(check_thermal_args(nx, ny, iterations));
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_dev) return set_error(HMG_ERR_INVALID, "null z");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  const size_t bytes = (size_t)nx * ny * sizeof(float);
  DevBuf       tmp, bed, z0;
  HMG_TRY(tmp.get(ctx->pool, bytes));
  HMG_TRY(bed.get(ctx->pool, bytes));
  HMG_TRY(z0.get(ctx->pool, bytes));
  HMG_TRY(launch_thermal_auto(z_dev, tmp.as<float>(), bed.as<float>(), z0.as<float>(), nx, ny,
                              talus_dev, talus_scalar, iterations, ctx->stream, ctx->num_cus));
  HMG_TRY(launch_extrapolate_borders(z_dev, nx, ny, ctx->stream));
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

int hmg_sediment_deposition(hmg_ctx     *ctx,
                            float       *z_host,
                            int          nx,
                            int          ny,
                            const float *talus_host,
                            float        talus_scalar,
                            float       *deposition_host,
                            float        max_deposition,
                            int          iterations,
                            int          thermal_subiterations,
                            const float *mask_host)
{
  clear_error();
  HMG_TRY(check_thermal_args(nx, ny, thermal_subiterations));
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_host) return set_error(HMG_ERR_INVALID, "null z");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  hipStream_t  st = ctx->stream;
  const size_t n = (size_t)nx * ny;
  const size_t bytes = n * sizeof(float);
  DevBuf       zb, z0, smap, ztot, tmp, tb, mb;
  HMG_TRY(zb.get(ctx->pool, bytes));
  HMG_TRY(smap.get(ctx->pool, bytes));
  HMG_TRY(ztot.get(ctx->pool, bytes));
  HMG_TRY(tmp.get(ctx->pool, bytes));
  const float *pt = nullptr, *pm = nullptr;
  HMG_HIP(hipMemcpyAsync(zb.ptr, z_host, bytes, hipMemcpyHostToDevice, st));
  if (talus_host)
  {
    HMG_TRY(tb.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(tb.ptr, talus_host, bytes, hipMemcpyHostToDevice, st));
    pt = tb.as<float>();
  }
  if (mask_host)
  {
    HMG_TRY(mb.get(ctx->pool, bytes));
    HMG_TRY(z0.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(mb.ptr, mask_host, bytes, hipMemcpyHostToDevice, st));
    HMG_HIP(hipMemcpyAsync(z0.ptr, zb.ptr, bytes, hipMemcpyDeviceToDevice, st));
    pm = mb.as<float>();
  }
  HMG_HIP(hipMemsetAsync(smap.ptr, 0, bytes, st));
  // deposition_step = max_deposition / (int)iterations (float division)
  const float step = max_deposition / (float)iterations;
  const int   nb = post_blocks(n);
  for (int it = 0; it < iterations; it++)
  {
    k_sd_pre<<<nb, 256, 0, st>>>(smap.as<float>(), zb.as<float>(), ztot.as<float>(), step, n);
    HMG_TRY(check_launch("sediment_deposition(pre)"));
    HMG_TRY(launch_thermal(ctx, ztot.as<float>(), tmp.as<float>(), nx, ny, pt, talus_scalar,
                           thermal_subiterations, nullptr, st, nullptr, true));
    k_sd_post<<<nb, 256, 0, st>>>(smap.as<float>(), ztot.as<float>(), zb.as<float>(), n);
    HMG_TRY(check_launch("sediment_deposition(post)"));
  }
  k_sd_finish<<<nb, 256, 0, st>>>(zb.as<float>(), smap.as<float>(), n);
  HMG_TRY(check_launch("sediment_deposition(finish)"));
  if (pm)
  {
    // the masked overload: z = lerp(z, z_f, mask) (deposition.cpp:47-62)
    k_thermal_post<<<nb, 256, 0, st>>>(zb.as<float>(), z0.as<float>(), nullptr, pm, n);
    HMG_TRY(check_launch("sediment_deposition(mask)"));
  }
  HMG_HIP(hipMemcpyAsync(z_host, zb.ptr, bytes, hipMemcpyDeviceToHost, st));
  if (deposition_host)
    HMG_HIP(hipMemcpyAsync(deposition_host, smap.ptr, bytes, hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  return HMG_OK;
}

} // extern "C"
