// vpipes.hip -- hmap::gpu::hydraulic_vpipes for gfx950, and the
// device-resident erosion session it runs on.
//
// Reference: host loop HighMap/src/erosion/hydraulic_vpipes_gpu.cpp:15-178,
// kernels HighMap/src/gpu_opencl/kernels/hydraulic_vpipes.cl:5-211.
//
// State kept in HBM between launches (one session = one map):
//   z    FP32            terrain height
//   D    storage type    water depth after the last water pass, before
//                        evaporation (d2 of the reference)
//   F    4 x storage     fluxes fl, fr, ft, fb after the last flow pass
//   Sp   storage type    sediment after the last erosion pass, before
//                        transport
// The reference iteration is flow -> water -> erosion -> transport ->
// evaporation -> top-up.  Transport, evaporation and top-up only feed the
// next iteration, so they are applied at the start of the next launch from
// the stored state; u and v of the previous iteration are recomputed from F
// and D exactly as the water pass formed them.  The result after k
// iterations is identical, value for value, to the reference order.
//
// Water volume: the top-up needs sum(d) over the whole grid every
// iteration.  Each launch adds round(d * 2^64) of its cells into 128-bit
// integer slots (exact, order-independent), and the next launch turns the
// slots into the top-up amount.  Three slot sets rotate: launch k reads set
// k-1, adds into set k and clears set k+1.
//
// Kernels (storage type S in {f32, bf16, f16}; FP32 sessions in the fused
// form keep z and D as one float2 plane, S = StF32ZD):
//   k_vp_stream, k_vp_stream_ga
//                  the fused form, one launch per iteration: each wavefront
//                  streams a strip of 64 columns (60 written) down a chunk
//                  of rows with a rolling register window, and writes z, D,
//                  F and Sp once (HBM: 28 B read + 28 B written per cell in
//                  FP32 storage); _ga issues each row's sediment gathers one
//                  row step before it uses them.
//   k_vp_batch, k_vp_batch_persist
//                  the same work item over many maps: one launch per
//                  iteration of every map, or one cooperative launch for
//                  all iterations with a grid barrier between them.
//   k_flow_stream  flow_simulation: the fused item with the flow and water
//                  passes only.
//   k_vp_prep, k_vp_flow, k_vp_water, k_vp_erosion, k_vp_sum
//                  the plain form: one launch per pass, every intermediate
//                  through HBM, as the reference structures its loop.
//   k_vp_export    forms d, s, u, v as hydraulic_vpipes returns them.
//   k_fill_state, k_nonfinite, k_absmax2, k_flow_init, k_dry_out
//                  session set-up, the finite check, the backtrace bound
//                  and flow_simulation's set-up and dry-out.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "hmg_internal.hpp"

namespace hmg
{

constexpr int NSLOT = 16; // fixed-point accumulators per slot set

// The streaming kernels address planes as a scalar base plus a 32-bit byte
// offset; a 16-byte flux quadruple keeps that below 2^32 up to 2^28 cells.
// Larger grids run the plain passes (64-bit indexing).
constexpr size_t STREAM_MAX_CELLS = (size_t)1 << 28;

struct VpDyn
{
  int    has_prev; // 0 on the first iteration of a session
  int    topup;    // volume top-up pending from the previous iteration
  double W0;       // initial water volume
  float  rf;       // rain volume (float, as the reference divides by it)
};

__device__ __forceinline__ int clampi(int x, int n)
{
  return min(max(x, 0), n - 1);
}

// Top-up amount (W0 - sum(d)) of the previous iteration from its slots.
__device__ __forceinline__ float slots_wta(const Fix128 *set, double W0)
{
  Fix128 acc{0, 0};
  for (int q = 0; q < NSLOT; q++) fix_add(acc, set[q]);
  return (float)(W0 - fix_to_double(acc.lo, (int64_t)acc.hi));
}

// Block-wide reduction of a Fix128 partial, then two atomics into one slot.
// `red` is LDS of at least (blockDim.x / 64) entries.
__device__ __forceinline__ void block_fix_atomic(Fix128 acc, Fix128 *red, Fix128 *set)
{
  for (int o = 32; o > 0; o >>= 1)
  {
    Fix128 other;
    other.lo = __shfl_xor(acc.lo, o, 64);
    other.hi = __shfl_xor(acc.hi, o, 64);
    fix_add(acc, other);
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    for (int q = 1; q < nw; q++) fix_add(acc, red[q]);
    Fix128             *slot = set + (blockIdx.x % NSLOT);
    unsigned long long  old = atomicAdd((unsigned long long *)&slot->lo, (unsigned long long)acc.lo);
    unsigned long long  carry = (old + acc.lo < old) ? 1ull : 0ull;
    atomicAdd((unsigned long long *)&slot->hi, (unsigned long long)acc.hi + carry);
  }
}

__device__ __forceinline__ void clear_slots(Fix128 *set)
{
  if (blockIdx.x == 0 && threadIdx.x < NSLOT) set[threadIdx.x] = Fix128{0, 0};
}

// u, v of the previous iteration at cell (i, j) from its stored fluxes and
// depth (the water pass, hydraulic_vpipes.cl:104-115).
template <class S>
__device__ __forceinline__ void prev_velocity(const Flux4<S> *__restrict__ F,
                                              const typename S::T *__restrict__ D,
                                              int i,
                                              int j,
                                              const VpConst &k,
                                              float &u,
                                              float &v)
{
  const int      nx = k.nx, ny = k.ny;
  size_t         c = (size_t)j * nx + i;
  Flux4<S>       fc = F[c];
  float          frl = S::ld(F[(size_t)j * nx + clampi(i - 1, nx)].r);
  float          flr = S::ld(F[(size_t)j * nx + clampi(i + 1, nx)].l);
  float          ftb = S::ld(F[(size_t)clampi(j - 1, ny) * nx + i].t);
  float          fbt = S::ld(F[(size_t)clampi(j + 1, ny) * nx + i].b);
  vp_velocity(S::ld(D[c]), frl, ftb, flr, fbt, S::ld(fc.l), S::ld(fc.r), S::ld(fc.t),
              S::ld(fc.b), u, v, k);
}

// ------------------------------------------------------------ plain form

// Start of an iteration: transport of the stored sediment with the previous
// velocities (or a copy on the first iteration), and the water depth after
// evaporation, top-up and rain.
template <class S>
__global__ void __launch_bounds__(256) k_vp_prep(const Flux4<S> *__restrict__ F,
                                                 const typename S::T *__restrict__ D,
                                                 const typename S::T *__restrict__ Sp,
                                                 const float *__restrict__ rain,
                                                 float *__restrict__ d1,
                                                 float *__restrict__ s,
                                                 const Fix128 *__restrict__ prev_set,
                                                 VpConst k,
                                                 VpDyn   y)
{
  __shared__ float s_wta;
  if (threadIdx.x == 0 && threadIdx.y == 0)
    s_wta = (y.has_prev && y.topup) ? slots_wta(prev_set, y.W0) : 0.f;
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  size_t c = (size_t)j * k.nx + i;
  float  r = rain ? rain[c] : 1.f;
  d1[c] = vp_d1(S::ld(D[c]), r, y.has_prev, y.topup, s_wta, y.rf, k);
  if (y.has_prev)
  {
    float u, v;
    prev_velocity<S>(F, D, i, j, k, u, v);
    s[c] = vp_transport<S>(Sp, i, j, u, v, k);
  }
  else
    s[c] = S::ld(Sp[c]);
}

template <class S>
__global__ void __launch_bounds__(256) k_vp_flow(const float *__restrict__ z,
                                                 const float *__restrict__ d1,
                                                 const Flux4<S> *__restrict__ F,
                                                 Flux4<S> *__restrict__ Fo,
                                                 VpConst k)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  const int nx = k.nx, ny = k.ny;
  size_t    c = (size_t)j * nx + i;
  size_t    cl = (size_t)j * nx + clampi(i - 1, nx), cr = (size_t)j * nx + clampi(i + 1, nx);
  size_t    ct = (size_t)clampi(j + 1, ny) * nx + i, cb = (size_t)clampi(j - 1, ny) * nx + i;
  Flux4<S>  f = F[c];
  float     l = S::ld(f.l), r = S::ld(f.r), t = S::ld(f.t), b = S::ld(f.b);
  vp_flow(z[c], d1[c], z[cl], d1[cl], z[cr], d1[cr], z[ct], d1[ct], z[cb], d1[cb], l, r, t, b, k);
  Flux4<S> o;
  o.l = S::st(l);
  o.r = S::st(r);
  o.t = S::st(t);
  o.b = S::st(b);
  Fo[c] = o;
}

template <class S>
__global__ void __launch_bounds__(256) k_vp_water(const Flux4<S> *__restrict__ F,
                                                  const float *__restrict__ d1,
                                                  typename S::T *__restrict__ Do,
                                                  float *__restrict__ uo,
                                                  float *__restrict__ vo,
                                                  VpConst k)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  const int nx = k.nx, ny = k.ny;
  size_t    c = (size_t)j * nx + i;
  Flux4<S>  fc = F[c];
  float     frl = S::ld(F[(size_t)j * nx + clampi(i - 1, nx)].r);
  float     flr = S::ld(F[(size_t)j * nx + clampi(i + 1, nx)].l);
  float     ftb = S::ld(F[(size_t)clampi(j - 1, ny) * nx + i].t);
  float     fbt = S::ld(F[(size_t)clampi(j + 1, ny) * nx + i].b);
  float     fl = S::ld(fc.l), fr = S::ld(fc.r), ft = S::ld(fc.t), fb = S::ld(fc.b);
  float     d2 = rnd<S>(vp_depth(d1[c], frl, ftb, flr, fbt, fl, fr, ft, fb, k));
  float     u, v;
  vp_velocity(d2, frl, ftb, flr, fbt, fl, fr, ft, fb, u, v, k);
  Do[c] = S::st(d2);
  uo[c] = u;
  vo[c] = v;
}

template <class S>
__global__ void __launch_bounds__(256) k_vp_erosion(const float *__restrict__ z,
                                                    const float *__restrict__ s,
                                                    const typename S::T *__restrict__ D,
                                                    const float *__restrict__ u,
                                                    const float *__restrict__ v,
                                                    float *__restrict__ zo,
                                                    typename S::T *__restrict__ So,
                                                    VpConst k)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  const int nx = k.nx, ny = k.ny;
  size_t    c = (size_t)j * nx + i;
  size_t    cl = (size_t)j * nx + clampi(i - 1, nx), cr = (size_t)j * nx + clampi(i + 1, nx);
  size_t    ct = (size_t)clampi(j + 1, ny) * nx + i, cb = (size_t)clampi(j - 1, ny) * nx + i;
  float     zn, sn;
  vp_erosion(z[c], s[c], z[cr], z[cl], z[ct], z[cb], s[cr], s[cl], s[ct], s[cb], S::ld(D[c]),
             u[c], v[c], zn, sn, k);
  zo[c] = zn;
  So[c] = S::st(sn);
}

// Sum of round(ld(a) * scale * 2^64) into a slot set.
template <class S>
__global__ void __launch_bounds__(256) k_vp_sum(const typename S::T *__restrict__ a,
                                                size_t n,
                                                float  scale,
                                                int    use_scale,
                                                Fix128 *set,
                                                Fix128 *clear_set)
{
  __shared__ Fix128 red[4];
  if (clear_set) clear_slots(clear_set);
  Fix128 acc{0, 0};
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
  {
    float x = S::ld(a[q]);
    if (use_scale) x = x * scale;
    fix_add(acc, fix_term(x));
  }
  block_fix_atomic(acc, red, set);
}

// ------------------------------------------------------------ fused form
//
// One launch per iteration.  Each wavefront owns a strip of 64 columns
// (60 written, a 2-cell halo on each side) and a chunk of rows, and streams
// down the rows keeping a rolling window in registers:
//   A(r)  staged state of row r: z, water d1 (after evaporation, top-up and
//         rain), stored depth, stored fluxes      -- loaded at row k+2
//   B(r)  sediment after transport               -- computed at row k+1
//   C(r)  new fluxes (flow pass)                 -- computed at row k+1
//   D(k)  water pass + erosion of row k, stores  -- computed at row k
// Left/right neighbours come from the adjacent lanes through DPP wave
// shifts (a VALU operand modifier, no LDS round trip).  Lanes left of
// column 0 and right of column nx-1 load the clamped column, so a staged
// A value fetched across a grid edge already equals the clamp-to-edge read;
// computed values (fluxes C, sediment B) differ in those lanes and are
// replaced by the lane's own value at the edge columns.  The wave's end
// lanes 0 and 63 receive 0; they are halo lanes whose results are unused.
// Above and below, loads use clamped rows, and the computed rows B/C(-1)
// and B/C(ny) are replaced by those of row 0 and ny-1.

#ifndef HMG_GA_ATTR
#define HMG_GA_ATTR
#endif
constexpr int STRIP_OUT = 60; // columns written per wavefront
constexpr int WPB = 4;        // wavefronts (adjacent strips) per workgroup

// lane i receives lane i-1's value (wave_shr:1)
__device__ __forceinline__ float from_left(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

// lane i receives lane i+1's value (wave_shl:1)
__device__ __forceinline__ float from_right(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// The neighbour's value, or the lane's own where `own` holds.  A DPP read
// from a lane that EXEC disables yields 0, and left to itself the compiler
// turns `own ? v : from_left(v)` into a DPP move under EXEC = !own, which
// disables exactly the edge lane the next lane reads.  The empty asm pins
// the exchange to all lanes before the select.
__device__ __forceinline__ float left_or_own(float v, bool own)
{
  float n = from_left(v);
  asm volatile("" : "+v"(n));
  return own ? v : n;
}

__device__ __forceinline__ float right_or_own(float v, bool own)
{
  float n = from_right(v);
  asm volatile("" : "+v"(n));
  return own ? v : n;
}

struct RowA
{
  float z, d1, dr, fl, fr, ft, fb;
};

template <class S, bool URAIN, bool FLOW = false>
__device__ __forceinline__ RowA stream_load(const float *__restrict__ z,
                                            const typename S::T *__restrict__ D,
                                            const Flux4<S> *__restrict__ F,
                                            const float *__restrict__ rain,
                                            uint32_t row_off,
                                            float    topup_add,
                                            float    wta,
                                            const VpConst &k,
                                            const VpDyn   &y)
{
  RowA     a;
  Flux4<S> f = ld_stream(F, row_off);
  if constexpr (S::ZD)
  {
    // the height and the water depth in one 8-byte load
    const float2 zd = ld_stream(reinterpret_cast<const float2 *>(z), row_off);
    a.z = zd.x;
    a.dr = zd.y;
  }
  else
  {
    a.z = ld_stream(z, row_off);
    a.dr = S::ld(ld_stream(D, row_off));
  }
  a.fl = S::ld(f.l);
  a.fr = S::ld(f.r);
  a.ft = S::ld(f.t);
  a.fb = S::ld(f.b);
  if (FLOW)
    a.d1 = a.dr; // flow_simulation: no evaporation, top-up or rain
  else if (URAIN)
  {
    // vp_d1 with rain == 1: (wta * 1) / R and rr * 1 are uniform
    float d = a.dr;
    if (y.has_prev)
    {
      d = a.dr * k.ce;
      if (y.topup) d = d + topup_add;
    }
    if (k.rr != 0.f) d = d + k.rr;
    a.d1 = d;
  }
  else
    a.d1 = vp_d1(a.dr, rain ? at_off(rain, row_off) : 1.f, y.has_prev, y.topup, wta, y.rf, k);
  return a;
}

// XCD-aware work order: workgroup b runs on XCD b % 8; give each XCD a
// contiguous range of work items so halo reuse stays in its L2.  The item
// is the same in all lanes of a wave; readfirstlane tells the compiler so,
// and everything derived from it (strip, rows, map) lives in SGPRs with
// scalar branches instead of EXEC-masked ones.
__device__ __forceinline__ int xcd_wave_item()
{
  const int G = gridDim.x;
  const int b = (G & 7) ? blockIdx.x : (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  return __builtin_amdgcn_readfirstlane(b * WPB + (threadIdx.x >> 6));
}

// Workgroups of a launch of n_items wavefront items, rounded up to a multiple
// of 8 so that the XCD-contiguous order above applies at every size; the
// extra workgroups hold no item and return at once.  At 4096^2 the 70-row
// chunks make 1018 workgroups of items: padded, an FP32 or FP16 launch is
// about 1% faster, and 74-row chunks 5% faster.
static inline int xcd_grid(int n_items) { return ((n_items + WPB - 1) / WPB + 7) & ~7; }

// One wavefront's work item of one iteration of one map: a strip of 60
// output columns over `chunk_rows` rows.  `item` numbers the map's items;
// the wavefront of item 0 clears the slot set of the next iteration.
// FLOW: the flow and water passes only (flow_simulation): d1 is the stored
// depth, no transport, erosion or volume sum; only D and F are written
// (Sp, rain, zo, So and the slot sets are not touched).
// NOSYNC: no per-row workgroup barrier (the persistent batch kernel, whose
// wavefronts of one workgroup may hold items of different lengths).
template <class S, bool URAIN, bool PF, bool KEXP1, bool FLOW = false, bool GA = false,
          bool NOSYNC = false>
__device__ __forceinline__ void vp_stream_item(const float *__restrict__ z,
                                               const typename S::T *__restrict__ D,
                                               const Flux4<S> *__restrict__ F,
                                               const typename S::T *__restrict__ Sp,
                                               const float *__restrict__ rain,
                                               float *__restrict__ zo,
                                               typename S::T *__restrict__ Do,
                                               Flux4<S> *__restrict__ Fo,
                                               typename S::T *__restrict__ So,
                                               const Fix128 *__restrict__ prev_set,
                                               Fix128 *set,
                                               Fix128 *clear_set,
                                               int     item,
                                               int     strips,
                                               int     chunk_rows,
                                               const VpConst &k,
                                               const VpDyn   &y)
{
  const int nx = k.nx, ny = k.ny;
  const int lane = threadIdx.x & 63;
  if (!FLOW && k.maintain && item == 0 && lane < NSLOT) clear_set[lane] = Fix128{0, 0};
  const int strip = item % strips, chunk = item / strips;

  float wta = 0.f, topup_add = 0.f;
  if (!FLOW && y.has_prev && y.topup)
  {
    wta = slots_wta(prev_set, y.W0);
    topup_add = (wta * 1.f) / y.rf;
  }

  const int x0 = strip * STRIP_OUT - 2; // grid column of lane 0
  const int gx = x0 + lane;
  const int gi = clampi(gx, nx);
  const bool ledge = gx == 0, redge = gx == nx - 1;
  const bool writer = lane >= 2 && lane < 62 && gx < nx;

  const int j0 = chunk * chunk_rows;
  const int j1 = min(j0 + chunk_rows, ny);

  auto load = [&](int r) {
    return stream_load<S, URAIN, FLOW>(z, D, F, rain, (uint32_t)(clampi(r, ny) * nx + gi),
                                       topup_add, wta, k, y);
  };
  // B(r): transport of row r from A(r-1), A(r), A(r+1), in two halves
  auto transport_issue = [&](const RowA &am, const RowA &a, const RowA &ap, int r) {
    const int gj = clampi(r, ny);
    float     frl = from_left(a.fr), flr = from_right(a.fl);
    float     u, v;
    vp_velocity(a.dr, frl, am.ft, flr, ap.fb, a.fl, a.fr, a.ft, a.fb, u, v, k);
    // first iteration of a session: the stored sediment at (gi, gj), read
    // without a branch so that the loads in flight do not depend on it
    return vp_transport_issue<S>(Sp, gi, gj, u, v, k, !y.has_prev);
  };
  auto transport_finish = [&](const TexelFetch &t) {
    // first iteration of a session: the stored sediment, untransported
    return y.has_prev ? vp_transport_finish(t) : t.s00;
  };
  auto transport = [&](const RowA &am, const RowA &a, const RowA &ap, int r) {
    return transport_finish(transport_issue(am, a, ap, r));
  };
  // C(r): flow of row r from A(r-1), A(r), A(r+1); rounded to storage
  auto flow = [&](const RowA &am, const RowA &a, const RowA &ap, float c[4]) {
    float zl = from_left(a.z), dl = from_left(a.d1);
    float zr = from_right(a.z), dr = from_right(a.d1);
    float l = a.fl, r = a.fr, t = a.ft, bb = a.fb;
    vp_flow(a.z, a.d1, zl, dl, zr, dr, ap.z, ap.d1, am.z, am.d1, l, r, t, bb, k);
    rnd4<S>(l, r, t, bb);
    c[0] = l;
    c[1] = r;
    c[2] = t;
    c[3] = bb;
  };

  // The window lives in rings of U slots indexed by compile-time constants:
  // the row loop is unrolled U times by hand (the compiler will not unroll
  // a loop holding DPP exchanges, which are convergent), so rows rotate
  // through the slots by renaming instead of register moves.  Row r + q of
  // an unrolled pass finds A(r+q-1) in slot q, A(r+q) in q+1, ..., and
  // B/C(r+q-1), B/C(r+q), B/C(r+q+1) in slots q, q+1, q+2 (mod U).  With
  // the load-ahead row, five A rows are live at once.  GA (gathers ahead):
  // rows are loaded two rows ahead and each row's sediment gathers are
  // issued one row before they are consumed, so every load has a whole row
  // step to arrive; six A rows and one row of gathered texels are live.
  static_assert(!GA || (PF && !FLOW), "gathers ahead imply the load-ahead row");
  static_assert(!(S::ZD && FLOW), "flow_simulation keeps separate planes");
  constexpr int  U = GA ? 6 : (PF ? 5 : 4);
  constexpr bool ROW_SYNC = !NOSYNC && !std::is_same<S, StBF16>::value;
  RowA       R[U];
  float      Sr[U], Cr[U][4];
  TexelFetch TFr[2]; // GA: gathers of the next row to finish, by parity

  // pre-roll: A(j0-2..j0+1), B/C(j0-1), B/C(j0)
  {
    RowA am2 = load(j0 - 2);
    R[0] = load(j0 - 1);
    R[1] = load(j0);
    R[2] = load(j0 + 1);
    if (!FLOW) Sr[0] = transport(am2, R[0], R[1], j0 - 1);
    flow(am2, R[0], R[1], Cr[0]);
    if (!FLOW) Sr[1] = transport(R[0], R[1], R[2], j0);
    flow(R[0], R[1], R[2], Cr[1]);
    if (PF) R[3] = load(j0 + 2);
    if constexpr (GA)
    {
      TFr[0] = transport_issue(R[1], R[2], R[3], j0 + 1);
      R[4] = load(j0 + 3);
    }
  }

  Fix128 acc{0, 0};
  for (int r0 = j0; r0 < j1; r0 += U)
  {
#pragma unroll
    for (int q = 0; q < U; q++)
    {
      const int r = r0 + q;
      if (r >= j1) break;
      // The wavefronts of a workgroup (adjacent strips) step through the
      // rows together, so the cache lines two strips share at their edge
      // are fetched from memory once rather than once per strip.  A
      // wavefront that has left the loop or the kernel no longer counts at
      // the barrier.  Measured at 4096^2: FP32 and FP16 state 2-7% faster
      // with it, BF16 about 3% slower.
      if (ROW_SYNC) __builtin_amdgcn_s_barrier();
      RowA &am1 = R[q % U], &a0 = R[(q + 1) % U], &ap1 = R[(q + 2) % U];
      RowA &ap2 = R[(q + 3) % U];
      // without the load-ahead the row is loaded here; with it, after this
      // row's gathers (below)
      if (!PF) ap2 = load(r + 2);
      float &s_m1 = Sr[q % U], &s_0 = Sr[(q + 1) % U], &s_p1 = Sr[(q + 2) % U];
      float *c_m1 = Cr[q % U], *c_0 = Cr[(q + 1) % U], *c_p1 = Cr[(q + 2) % U];

      if (FLOW)
      {
        // flow_simulation: the water pass of row r, then D and F
        if (PF) R[(q + 4) % U] = load(r + 3);
        flow(a0, ap1, ap2, c_p1);
        const bool top = r == 0, bot = r == ny - 1;
        float      ftb = top ? c_0[2] : c_m1[2], fbt = bot ? c_0[3] : c_p1[3];
        float      frl = left_or_own(c_0[1], ledge);
        float      flr = right_or_own(c_0[0], redge);
        float d2 = rnd<S>(vp_depth(a0.d1, frl, ftb, flr, fbt, c_0[0], c_0[1], c_0[2], c_0[3], k));
        if (writer)
        {
          const uint32_t c = (uint32_t)(r * nx + gx);
          st_stream(Do, c, S::st_rounded(d2));
          Flux4<S> o;
          o.l = S::st_rounded(c_0[0]);
          o.r = S::st_rounded(c_0[1]);
          o.t = S::st_rounded(c_0[2]);
          o.b = S::st_rounded(c_0[3]);
          st_stream(Fo, c, o);
        }
        continue;
      }
      // Memory order within the row: the sediment gathers of row r+1 first,
      // then the load-ahead row r+3; the flow of row r+1 runs while both are
      // in flight, then the gathers are consumed.  Loads complete in issue
      // order for the wait counter, so with the row loads issued after the
      // gathers, waiting for the gathers leaves the row loads in flight (to
      // be consumed one row later).  The empty asm statements fix this order
      // against the scheduler: loads cannot cross a memory clobber, the flow
      // starts from values redefined after the loads, and the gathered
      // samples are first used after the flow results.
      if (GA)
      {
        // gathers of row r+2 (rows r+1..r+3 are in), then the row r+4
        // loads, then the flow of row r+1; the gathers of row r+1 were
        // issued one step ago
        RowA &ap3 = R[(q + 4) % U];
        TFr[(q + 1) % 2] = transport_issue(ap1, ap2, ap3, r + 2);
        asm volatile("" ::: "memory");
        R[(q + 5) % U] = load(r + 4);
        asm volatile("" : "+v"(ap1.z), "+v"(ap1.d1)::"memory");
        flow(a0, ap1, ap2, c_p1);
        s_p1 = transport_finish(TFr[q % 2]);
      }
      else
      {
        TexelFetch tf = transport_issue(a0, ap1, ap2, r + 1);
        asm volatile("" ::: "memory");
        if (PF) R[(q + 4) % U] = load(r + 3);
        asm volatile("" : "+v"(ap1.z), "+v"(ap1.d1)::"memory");
        flow(a0, ap1, ap2, c_p1);
        asm volatile(""
                     : "+v"(c_p1[0]), "+v"(c_p1[1]), "+v"(c_p1[2]), "+v"(c_p1[3]), "+v"(tf.s00),
                       "+v"(tf.s10), "+v"(tf.s01), "+v"(tf.s11));
        s_p1 = transport_finish(tf);
      }

      // row r: water pass and erosion
      const bool top = r == 0, bot = r == ny - 1;
      float      s_lo = top ? s_0 : s_m1, s_hi = bot ? s_0 : s_p1;
      float      ftb = top ? c_0[2] : c_m1[2], fbt = bot ? c_0[3] : c_p1[3];
      float      frl = left_or_own(c_0[1], ledge);
      float      flr = right_or_own(c_0[0], redge);
      float      d2 = rnd<S>(vp_depth(a0.d1, frl, ftb, flr, fbt, c_0[0], c_0[1], c_0[2], c_0[3], k));
      float      u, v;
      vp_velocity(d2, frl, ftb, flr, fbt, c_0[0], c_0[1], c_0[2], c_0[3], u, v, k);
      float zl = from_left(a0.z), zr = from_right(a0.z);
      float sl = left_or_own(s_0, ledge), sr = right_or_own(s_0, redge);
      float zn, sn;
      vp_erosion<KEXP1>(a0.z, s_0, zr, zl, ap1.z, am1.z, sr, sl, s_hi, s_lo, d2, u, v, zn, sn, k);
      if (writer)
      {
        const uint32_t c = (uint32_t)(r * nx + gx);
        // z, D and F are read once by the next launch, and the sediment
        // many times (its gathers): the streamed planes carry the
        // nontemporal hint, the sediment does not, so that it stays in
        // the caches for the gathers.  At 4096^2 FP32 this takes a launch
        // from 226 to 212 us; the hint on the loads alone makes it 235, on
        // every store (the sediment's too) 229.
        if constexpr (S::ZD)
          st_stream(reinterpret_cast<float2 *>(zo), c, make_float2(zn, d2));
        else
          st_stream(zo, c, zn);
        at_off(So, c) = S::st(sn);
        if constexpr (!S::ZD) st_stream(Do, c, S::st_rounded(d2));
        Flux4<S> o;
        o.l = S::st_rounded(c_0[0]);
        o.r = S::st_rounded(c_0[1]);
        o.t = S::st_rounded(c_0[2]);
        o.b = S::st_rounded(c_0[3]);
        st_stream(Fo, c, o);
        if (k.maintain) fix_accumulate(acc, d2 * k.ce);
      }
    }
  }

  if (!FLOW && k.maintain)
  {
    for (int o = 32; o > 0; o >>= 1)
    {
      Fix128 other;
      other.lo = __shfl_xor(acc.lo, o, 64);
      other.hi = __shfl_xor(acc.hi, o, 64);
      fix_add(acc, other);
    }
    if (lane == 0)
    {
      Fix128            *slot = set + (item % NSLOT);
      unsigned long long old = atomicAdd((unsigned long long *)&slot->lo, (unsigned long long)acc.lo);
      unsigned long long carry = (old + acc.lo < old) ? 1ull : 0ull;
      atomicAdd((unsigned long long *)&slot->hi, (unsigned long long)acc.hi + carry);
    }
  }
}

// One iteration of one session.
template <class S, bool URAIN, bool PF, bool KEXP1, bool GA = false>
__global__ void __launch_bounds__(64 * WPB) k_vp_stream(const float *__restrict__ z,
                                                   const typename S::T *__restrict__ D,
                                                   const Flux4<S> *__restrict__ F,
                                                   const typename S::T *__restrict__ Sp,
                                                   const float *__restrict__ rain,
                                                   float *__restrict__ zo,
                                                   typename S::T *__restrict__ Do,
                                                   Flux4<S> *__restrict__ Fo,
                                                   typename S::T *__restrict__ So,
                                                   const Fix128 *__restrict__ prev_set,
                                                   Fix128 *set,
                                                   Fix128 *clear_set,
                                                   int     strips,
                                                   int     chunk_rows,
                                                   int     n_items,
                                                   VpConst k,
                                                   VpDyn   y)
{
  const int item = xcd_wave_item();
  if (item >= n_items) return;
  vp_stream_item<S, URAIN, PF, KEXP1, false, GA>(z, D, F, Sp, rain, zo, Do, Fo, So, prev_set, set,
                                                 clear_set, item, strips, chunk_rows, k, y);
}

// The gathers-ahead form (see vp_stream_item).
template <class S, bool URAIN, bool KEXP1>
__global__ void __launch_bounds__(64 * WPB) HMG_GA_ATTR k_vp_stream_ga(const float *__restrict__ z,
                                                           const typename S::T *__restrict__ D,
                                                           const Flux4<S> *__restrict__ F,
                                                           const typename S::T *__restrict__ Sp,
                                                           const float *__restrict__ rain,
                                                           float *__restrict__ zo,
                                                           typename S::T *__restrict__ Do,
                                                           Flux4<S> *__restrict__ Fo,
                                                           typename S::T *__restrict__ So,
                                                           const Fix128 *__restrict__ prev_set,
                                                           Fix128 *set,
                                                           Fix128 *clear_set,
                                                           int     strips,
                                                           int     chunk_rows,
                                                           int     n_items,
                                                           VpConst k,
                                                           VpDyn   y)
{
  const int item = xcd_wave_item();
  if (item >= n_items) return;
  vp_stream_item<S, URAIN, true, KEXP1, false, true>(z, D, F, Sp, rain, zo, Do, Fo, So, prev_set,
                                                     set, clear_set, item, strips, chunk_rows, k,
                                                     y);
}

// ------------------------------------------------------------ batched form
//
// One launch advances every map of a batch by one iteration.  The maps are
// described by a table in device memory; each map keeps its own planes (both
// ping-pong halves), slot sets and constants, so its result does not depend
// on which other maps share the launch.  Map m runs its iteration
// it0 + (t - t_start) at launch t while that is below it0 + iters; the
// parity of its planes and its slot sets follow from that iteration number.

struct VpMapDesc
{
  void        *z[2], *D[2], *F[2], *Sp[2];
  const float *rain;
  Fix128      *slots; // 3 sets of NSLOT
  VpConst      k;
  double       W0;
  float        rf;
  int          topup;
  int          strips, chunk_rows, n_items, item_base;
  int          t_start, it0, iters;
};

// Item `item` of batched launch t: finds its map (maps are ordered by
// item_base) and runs that map's iteration it0 + (t - t_start) on it, if
// that is inside the map's window.
template <class S, bool URAIN, bool PF, bool KEXP1, bool GA, bool NOSYNC>
__device__ __forceinline__ void vp_batch_item(const VpMapDesc *__restrict__ maps,
                                              int n_maps,
                                              int item,
                                              int t)
{
  using T = typename S::T;
  int lo = 0, hi = n_maps - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (maps[mid].item_base <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  const VpMapDesc &m = maps[lo];
  const int        rel = t - m.t_start;
  if (rel < 0 || rel >= m.iters) return;
  const int mi = m.it0 + rel;
  const int a = mi & 1, b = a ^ 1;
  VpDyn     y;
  y.has_prev = mi > 0;
  y.topup = m.topup;
  y.W0 = m.W0;
  y.rf = m.rf;
  Fix128 *prev = m.slots + ((mi + 2) % 3) * NSLOT;
  Fix128 *cur = m.slots + (mi % 3) * NSLOT;
  Fix128 *nxt = m.slots + ((mi + 1) % 3) * NSLOT;
  vp_stream_item<S, URAIN, PF, KEXP1, false, GA, NOSYNC>(
      (const float *)m.z[a], (const T *)m.D[a], (const Flux4<S> *)m.F[a], (const T *)m.Sp[a],
      m.rain, (float *)m.z[b], (T *)m.D[b], (Flux4<S> *)m.F[b], (T *)m.Sp[b], prev, cur, nxt,
      item - m.item_base, m.strips, m.chunk_rows, m.k, y);
}

template <class S, bool URAIN, bool PF, bool KEXP1, bool GA = false>
__global__ void __launch_bounds__(64 * WPB) k_vp_batch(const VpMapDesc *__restrict__ maps,
                                                  int n_maps,
                                                  int n_items,
                                                  int t)
{
  const int item = xcd_wave_item();
  if (item >= n_items) return;
  vp_batch_item<S, URAIN, PF, KEXP1, GA, false>(maps, n_maps, item, t);
}

// ------------------------------------------------ persistent batched form
//
// One cooperative launch runs launches t = 0 .. T-1 of the batch: every
// wavefront takes items of launch t, then the whole grid meets at a barrier
// before launch t + 1 (an iteration reads its neighbours' previous state,
// and the sediment gathers reach any distance, so no item of t + 1 may start
// before every item of t is done).  The grid is no larger than what the
// device holds at once, which the cooperative launch guarantees.

struct GridBar
{
  unsigned count; // workgroups arrived at the current barrier
  unsigned gen;   // barriers completed
  unsigned err;   // set when a wait gave up
  unsigned limit; // polls of a wait before it gives up
};

// polls of a barrier wait (an atomic load and an s_sleep each, about a
// microsecond) before it gives up and flags the batch as failed: seconds,
// far beyond one launch's work.  HMG_GRID_SPIN_LIMIT in the environment
// replaces it (tests force a give-up with a limit of 1).
constexpr unsigned GRID_SPIN_LIMIT = 1u << 22;

// All workgroups of the grid meet; the writes of each before it are visible
// to all after it (agent-scope release before arriving, acquire after).
// Returns false, in every thread of the workgroup, once any wait has given
// up (the flag is polled while waiting, so the other workgroups stop soon
// after instead of each waiting out its own limit).
__device__ __forceinline__ bool grid_barrier(GridBar *bar)
{
  __shared__ unsigned ok;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    __threadfence();
    const unsigned g = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned       good = 1u;
    if (atomicAdd(&bar->count, 1u) == gridDim.x - 1)
    {
      __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    else
    {
      unsigned spins = 0;
      while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g)
      {
        if (__hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
            ++spins >= bar->limit)
        {
          atomicOr(&bar->err, 1u);
          good = 0u;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
    ok = good;
  }
  __syncthreads();
  return ok != 0u;
}

// items_at[t], active_at[t]: items and active maps of launch t.  Each XCD
// (workgroup b runs on XCD b % 8) takes a contiguous eighth of a launch's
// items, so halo rows shared by neighbouring items stay in its L2.
template <class S, bool URAIN, bool PF, bool KEXP1, bool GA = false>
__global__ void __launch_bounds__(64 * WPB) k_vp_batch_persist(const VpMapDesc *__restrict__ maps,
                                                          const int *__restrict__ items_at,
                                                          const int *__restrict__ active_at,
                                                          int      T,
                                                          GridBar *bar)
{
  const int  G = gridDim.x, wave = threadIdx.x >> 6;
  const bool by_xcd = (G & 7) == 0;
  const int  parts = by_xcd ? 8 : 1;
  const int  part = by_xcd ? (int)(blockIdx.x & 7) : 0;
  const int  slot = __builtin_amdgcn_readfirstlane((by_xcd ? (int)(blockIdx.x >> 3) : (int)blockIdx.x) *
                                                       WPB + wave);
  const int  slots = (G / parts) * WPB;
  for (int t = 0; t < T; t++)
  {
    const int n = items_at[t], active = active_at[t];
    const int i0 = (int)((long)n * part / parts), i1 = (int)((long)n * (part + 1) / parts);
    for (int item = i0 + slot; item < i1; item += slots)
      vp_batch_item<S, URAIN, PF, KEXP1, GA, true>(maps, active, item, t);
    if (!grid_barrier(bar)) return;
  }
}

// ---------------------------------------------------------------- export

template <class S>
__global__ void __launch_bounds__(256) k_vp_export(const Flux4<S> *__restrict__ F,
                                                   const typename S::T *__restrict__ D,
                                                   const typename S::T *__restrict__ Sp,
                                                   const float *__restrict__ rain,
                                                   float *__restrict__ d_out,
                                                   float *__restrict__ s_out,
                                                   float *__restrict__ u_out,
                                                   float *__restrict__ v_out,
                                                   const Fix128 *__restrict__ prev_set,
                                                   VpConst k,
                                                   VpDyn   y)
{
  __shared__ float s_wta;
  if (threadIdx.x == 0 && threadIdx.y == 0)
    s_wta = (y.has_prev && y.topup) ? slots_wta(prev_set, y.W0) : 0.f;
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= k.nx || j >= k.ny) return;
  size_t c = (size_t)j * k.nx + i;
  float  dr = S::ld(D[c]);
  if (!y.has_prev)
  {
    d_out[c] = dr;
    s_out[c] = S::ld(Sp[c]);
    u_out[c] = 0.f;
    v_out[c] = 0.f;
    return;
  }
  VpConst k0 = k;
  k0.rr = 0.f; // state d excludes the rain of the next iteration
  d_out[c] = vp_d1(dr, rain ? rain[c] : 1.f, 1, y.topup, s_wta, y.rf, k0);
  float u, v;
  prev_velocity<S>(F, D, i, j, k, u, v);
  s_out[c] = vp_transport<S>(Sp, i, j, u, v, k);
  u_out[c] = u;
  v_out[c] = v;
}

template <class S>
__global__ void __launch_bounds__(256) k_fill_state(typename S::T *__restrict__ D,
                                                    typename S::T *__restrict__ Sp,
                                                    Flux4<S> *__restrict__ F,
                                                    const float *__restrict__ rain,
                                                    float wh,
                                                    size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
  {
    // hydraulic_vpipes_gpu.cpp:39,53: d = water_height * rain
    D[q] = S::st(wh * (rain ? rain[q] : 1.f));
    Sp[q] = S::st(0.f);
    Flux4<S> f;
    f.l = f.r = f.t = f.b = S::st(0.f);
    F[q] = f;
  }
}

__global__ void k_zero_fix(Fix128 *p, int n)
{
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = Fix128{0, 0};
}

} // namespace hmg

// ---------------------------------------------------------------- session

using namespace hmg;

namespace hmg
{

static VpDyn make_dyn(const hmg_session *s)
{
  VpDyn y;
  y.has_prev = s->iter > 0;
  y.topup = s->p.maintain_water_volume ? 1 : 0;
  y.W0 = s->W0;
  y.rf = s->rf;
  return y;
}

static Fix128 *slot_set(hmg_session *s, long it)
{
  return s->slots.as<Fix128>() + (size_t)(((it % 3) + 3) % 3) * NSLOT;
}

// D0: the initial water depth plane (the session's D[0], or for the float2
// layout a plane the caller then merges into z[0]).
template <class S>
static int session_init_state(hmg_session *s, void *D0)
{
  size_t      n = (size_t)s->nx * s->ny;
  hipStream_t st = s->ctx->stream;
  int         nb = (int)((n + 255) / 256);
  if (nb > 4096) nb = 4096;
  k_fill_state<S><<<nb, 256, 0, st>>>((typename S::T *)D0,
                                      (typename S::T *)s->Sp[0].ptr,
                                      (Flux4<S> *)s->F[0].ptr,
                                      s->rain.as<float>(),
                                      s->p.water_height,
                                      n);
  HMG_TRY(check_launch("fill_state"));
  k_zero_fix<<<1, 64, 0, st>>>(s->slots.as<Fix128>(), 4 * NSLOT);
  HMG_TRY(check_launch("zero_fix"));
  // W0 = sum(d) and R = sum(rain) (hydraulic_vpipes_gpu.cpp:54-55), exact
  Fix128 *w_set = s->slots.as<Fix128>() + 3 * NSLOT;
  k_vp_sum<S><<<nb, 256, 0, st>>>((const typename S::T *)D0, n, 1.f, 0, w_set, nullptr);
  HMG_TRY(check_launch("vp_sum(W0)"));
  Fix128 h[NSLOT];
  HMG_HIP(hipMemcpyAsync(h, w_set, sizeof(h), hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  Fix128 acc{0, 0};
  for (int q = 0; q < NSLOT; q++)
  {
    uint64_t lo = acc.lo + h[q].lo;
    uint64_t c = lo < acc.lo ? 1 : 0;
    acc.lo = lo;
    acc.hi = (int64_t)((uint64_t)acc.hi + (uint64_t)h[q].hi + c);
  }
  s->W0 = fix_to_double(acc.lo, acc.hi);
  if (s->rain.ptr)
  {
    k_zero_fix<<<1, 64, 0, st>>>(w_set, NSLOT);
    HMG_TRY(check_launch("zero_fix"));
    k_vp_sum<StF32><<<nb, 256, 0, st>>>(s->rain.as<float>(), n, 1.f, 0, w_set, nullptr);
    HMG_TRY(check_launch("vp_sum(R)"));
    HMG_HIP(hipMemcpyAsync(h, w_set, sizeof(h), hipMemcpyDeviceToHost, st));
    HMG_HIP(hipStreamSynchronize(st));
    acc = Fix128{0, 0};
    for (int q = 0; q < NSLOT; q++)
    {
      uint64_t lo = acc.lo + h[q].lo;
      uint64_t c = lo < acc.lo ? 1 : 0;
      acc.lo = lo;
      acc.hi = (int64_t)((uint64_t)acc.hi + (uint64_t)h[q].hi + c);
    }
    s->R = fix_to_double(acc.lo, acc.hi);
  }
  else
    s->R = (double)n;
  s->rf = (float)s->R;
  return HMG_OK;
}

template <class S>
static int session_step_plain(hmg_session *s)
{
  using T = typename S::T;
  const int   nx = s->nx, ny = s->ny;
  hipStream_t st = s->ctx->stream;
  dim3        blk(64, 4), grd((nx + 63) / 64, (ny + 3) / 4);
  const int   a = s->cur, b = s->cur ^ 1;
  VpDyn       y = make_dyn(s);
  k_vp_prep<S><<<grd, blk, 0, st>>>((const Flux4<S> *)s->F[a].ptr, (const T *)s->D[a].ptr,
                                    (const T *)s->Sp[a].ptr, s->rain.as<float>(),
                                    s->d1.as<float>(), s->s.as<float>(),
                                    slot_set(s, s->iter - 1), s->k, y);
  HMG_TRY(check_launch("vp_prep"));
  k_vp_flow<S><<<grd, blk, 0, st>>>(s->z[a].as<float>(), s->d1.as<float>(),
                                    (const Flux4<S> *)s->F[a].ptr, (Flux4<S> *)s->F[b].ptr,
                                    s->k);
  HMG_TRY(check_launch("vp_flow"));
  k_vp_water<S><<<grd, blk, 0, st>>>((const Flux4<S> *)s->F[b].ptr, s->d1.as<float>(),
                                     (T *)s->D[b].ptr, s->u.as<float>(), s->v.as<float>(),
                                     s->k);
  HMG_TRY(check_launch("vp_water"));
  k_vp_erosion<S><<<grd, blk, 0, st>>>(s->z[a].as<float>(), s->s.as<float>(),
                                       (const T *)s->D[b].ptr, s->u.as<float>(),
                                       s->v.as<float>(), s->z[b].as<float>(),
                                       (T *)s->Sp[b].ptr, s->k);
  HMG_TRY(check_launch("vp_erosion"));
  if (s->p.maintain_water_volume)
  {
    size_t n = (size_t)nx * ny;
    int    nb = (int)((n + 255) / 256);
    if (nb > 2048) nb = 2048;
    k_vp_sum<S><<<nb, 256, 0, st>>>((const T *)s->D[b].ptr, n, s->k.ce, 1,
                                    slot_set(s, s->iter), slot_set(s, s->iter + 1));
    HMG_TRY(check_launch("vp_sum"));
  }
  s->cur = b;
  s->iter++;
  return HMG_OK;
}

template <class S>
static int session_step_fused(hmg_session *s)
{
  using T = typename S::T;
  const int   nx = s->nx, ny = s->ny;
  hipStream_t st = s->ctx->stream;
  const int   a = s->cur, b = s->cur ^ 1;
  const int   strips = (nx + STRIP_OUT - 1) / STRIP_OUT;
  const int   ch = s->chunk_rows;
  const int   n_items = strips * ((ny + ch - 1) / ch);
  const int   G = xcd_grid(n_items);
  VpDyn       y = make_dyn(s);
  const T    *D = (const T *)s->D[a].ptr;
  const T    *Sp = (const T *)s->Sp[a].ptr;
  const float *rain = s->rain.as<float>();
  float       *zo = s->z[b].as<float>();
  T           *Do = (T *)s->D[b].ptr, *So = (T *)s->Sp[b].ptr;
  Flux4<S>    *Fo = (Flux4<S> *)s->F[b].ptr;
  const Flux4<S> *F = (const Flux4<S> *)s->F[a].ptr;
  Fix128      *pset = slot_set(s, s->iter - 1), *aset = slot_set(s, s->iter);
  Fix128      *cset = slot_set(s, s->iter + 1);
  const float *z = s->z[a].as<float>();
#define HMG_STREAM(UR, PFV, KE)                                                \
  k_vp_stream<S, UR, PFV, KE><<<G, 64 * WPB, 0, st>>>(z, D, F, Sp, rain, zo, Do, Fo, So, pset, \
                                                 aset, cset, strips, ch, n_items, s->k, y)
#define HMG_STREAM_GA(UR, KE)                                                  \
  k_vp_stream_ga<S, UR, KE><<<G, 64 * WPB, 0, st>>>(z, D, F, Sp, rain, zo, Do, Fo, So, pset,     \
                                                  aset, cset, strips, ch, n_items, s->k, y)
  const bool ke1 = s->k.kexp == 1.f;
  if (s->prefetch == 2)
  {
    if (rain)
    {
      if (ke1)
        HMG_STREAM_GA(false, true);
      else
        HMG_STREAM_GA(false, false);
    }
    else if (ke1)
      HMG_STREAM_GA(true, true);
    else
      HMG_STREAM_GA(true, false);
  }
  else if (rain)
  {
    if (ke1)
      HMG_STREAM(false, true, true);
    else
      HMG_STREAM(false, true, false);
  }
  else if (!s->prefetch)
  {
    if (ke1)
      HMG_STREAM(true, false, true);
    else
      HMG_STREAM(true, false, false);
  }
  else
  {
    if (ke1)
      HMG_STREAM(true, true, true);
    else
      HMG_STREAM(true, true, false);
  }
#undef HMG_STREAM
#undef HMG_STREAM_GA
  HMG_TRY(check_launch("vp_stream"));
  s->cur = b;
  s->iter++;
  return HMG_OK;
}

template <class S>
static int session_step(hmg_session *s)
{
  if (s->p.mode == HMG_VP_PLAIN) return session_step_plain<S>(s);
  return session_step_fused<S>(s);
}

template <class S>
static int session_export(hmg_session *s, float *d, float *sd, float *u, float *v)
{
  using T = typename S::T;
  const int   nx = s->nx, ny = s->ny;
  hipStream_t st = s->ctx->stream;
  dim3        blk(64, 4), grd((nx + 63) / 64, (ny + 3) / 4);
  VpDyn       y = make_dyn(s);
  const void *Dp = s->D[s->cur].ptr;
  DevBuf      Dt; // float2 layout: the depths split out
  if (s->zd)
  {
    const size_t n = (size_t)nx * ny;
    HMG_TRY(Dt.get(s->ctx->pool, n * 4));
    HMG_TRY(launch_zd_split((const float2 *)s->z[s->cur].ptr, Dt.as<float>(), 1, n, st));
    Dp = Dt.ptr;
  }
  k_vp_export<S><<<grd, blk, 0, st>>>((const Flux4<S> *)s->F[s->cur].ptr, (const T *)Dp,
                                      (const T *)s->Sp[s->cur].ptr, s->rain.as<float>(), d,
                                      sd, u, v, slot_set(s, s->iter - 1), s->k, y);
  HMG_TRY(check_launch("vp_export"));
  // the split plane returns to the pool when this call ends
  if (s->zd) HMG_HIP(hipStreamSynchronize(st));
  return HMG_OK;
}

// The session's current heights as a float plane: its z plane, or for the
// float2 layout the heights split into `tmp` (n floats, from the pool).
static int session_heights(hmg_session *s, DevBuf &tmp, const float **out)
{
  if (!s->zd)
  {
    *out = s->z[s->cur].as<float>();
    return HMG_OK;
  }
  const size_t n = (size_t)s->nx * s->ny;
  HMG_TRY(tmp.get(s->ctx->pool, n * 4));
  HMG_TRY(launch_zd_split((const float2 *)s->z[s->cur].ptr, tmp.as<float>(), 0, n,
                          s->ctx->stream));
  *out = tmp.as<float>();
  return HMG_OK;
}

static int dispatch_step(hmg_session *s)
{
  switch (s->storage)
  {
  case HMG_STORAGE_BF16: return session_step<StBF16>(s);
  case HMG_STORAGE_F16: return session_step<StF16>(s);
  default: return s->zd ? session_step_fused<StF32ZD>(s) : session_step<StF32>(s);
  }
}

// Workgroups of the fused kernel variant (1: one-row load-ahead, 2: gathers
// ahead) one CU holds at once for this session's storage and options, from
// the kernel's register use (0 if the runtime cannot tell).
template <class S>
static int stream_blocks_per_cu_t(const hmg_session *s, int variant)
{
  const bool  rain = s->rain.ptr != nullptr, ke1 = s->k.kexp == 1.f;
  const void *f = nullptr;
  if (variant == 2)
    f = rain ? (ke1 ? (const void *)k_vp_stream_ga<S, false, true> : (const void *)k_vp_stream_ga<S, false, false>)
             : (ke1 ? (const void *)k_vp_stream_ga<S, true, true> : (const void *)k_vp_stream_ga<S, true, false>);
  else
    f = rain ? (ke1 ? (const void *)k_vp_stream<S, false, true, true> : (const void *)k_vp_stream<S, false, true, false>)
             : (ke1 ? (const void *)k_vp_stream<S, true, true, true> : (const void *)k_vp_stream<S, true, true, false>);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64 * WPB, 0) != hipSuccess)
  {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

static int stream_blocks_per_cu(const hmg_session *s, int variant)
{
  if (s->storage == HMG_STORAGE_BF16) return stream_blocks_per_cu_t<StBF16>(s, variant);
  if (s->storage == HMG_STORAGE_F16) return stream_blocks_per_cu_t<StF16>(s, variant);
  return s->zd ? stream_blocks_per_cu_t<StF32ZD>(s, variant) : stream_blocks_per_cu_t<StF32>(s, variant);
}

// ------------------------------------------------ online shape tuning
//
// The fused kernel's speed depends on rows per wavefront and load-ahead in
// steps that follow how the wavefronts fill the CUs, which no formula
// predicts well.  A session of at least 2^20 cells in fused mode therefore
// runs its first iterations as probes: each launches one candidate shape
// between two events (two rounds over the candidates, after one warm-up
// launch); at the first iteration after the probes the events are read,
// the shape with the least time (minimum over its two rounds) is kept and
// cached per (storage, nx, ny, rain plane, k_discharge_exp == 1) for later
// sessions of the process.  The probes are ordinary iterations.  They start
// at iteration TUNE_AFTER: the first iterations, while the water spreads
// from rest, have short backtraces whose sediment gathers hit the cache,
// and ranked the shapes differently from the rest of a run (4096^2 BF16:
// 199 us per launch for the shape chosen from iterations 1-29 against 182
// for the best later).

namespace
{
struct TuneKey
{
  int  storage, nx, ny, rain, ke1;
  bool operator<(const TuneKey &o) const
  {
    return std::tie(storage, nx, ny, rain, ke1) < std::tie(o.storage, o.nx, o.ny, o.rain, o.ke1);
  }
};
std::mutex                                   g_shape_mu;
std::map<TuneKey, std::pair<int, int>>       g_shape_cache;
// (rows per wavefront, variant): 0 rows loaded when needed, 1 one row
// ahead, 2 two rows ahead with the gathers one row ahead
constexpr int TUNE_CAND[][2] = {{8, 2},  {12, 1}, {16, 2}, {24, 1}, {32, 1}, {40, 1},
                                {47, 1}, {56, 1}, {64, 1}, {32, 2}, {40, 2}, {47, 2},
                                {72, 2}, {94, 2}};
constexpr int                                N_CAND = sizeof(TUNE_CAND) / sizeof(TUNE_CAND[0]);
constexpr long                               TUNE_AFTER = 20;

TuneKey tune_key(const hmg_session *s)
{
  return TuneKey{s->storage, s->nx, s->ny, s->rain.ptr ? 1 : 0, s->k.kexp == 1.f ? 1 : 0};
}
} // namespace

static void tune_init(hmg_session *s)
{
  s->tune_pos = -1;
  if (s->p.mode != HMG_VP_FUSED || (size_t)s->nx * s->ny < ((size_t)1 << 20)) return;
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto                        it = g_shape_cache.find(tune_key(s));
    if (it != g_shape_cache.end())
    {
      s->chunk_rows = it->second.first;
      s->prefetch = it->second.second;
      return;
    }
  }
  // the candidates that give every SIMD at least one wavefront (the
  // shortest chunks if none does): longer chunks leave most of the GPU idle
  // at this size and only slow the probes down
  s->tune_cand.clear();
  const int  strips = (s->nx + STRIP_OUT - 1) / STRIP_OUT;
  const int  cus = s->ctx->num_cus > 0 ? s->ctx->num_cus : 256;
  const long simds = 4L * cus;
  auto       items = [&](int rows) { return (long)strips * ((s->ny + rows - 1) / rows); };
  for (int c = 0; c < N_CAND; c++)
    if (items(TUNE_CAND[c][0]) >= simds) s->tune_cand.push_back({TUNE_CAND[c][0], TUNE_CAND[c][1]});
  if (s->tune_cand.empty())
    for (int c = 0; c < N_CAND; c++)
      if (TUNE_CAND[c][0] <= 16) s->tune_cand.push_back({TUNE_CAND[c][0], TUNE_CAND[c][1]});
  // Per variant, the shortest chunks whose work items all fit the
  // wavefront slots the kernel's occupancy gives (one round of wavefronts,
  // no nearly empty second round): at 4096^2 the gathers-ahead kernel holds
  // 4 wavefronts per SIMD and 69 strips x 57 chunks of 72 rows fill 96% of
  // them, while 68 rows (61 chunks) overflow into a second round and run
  // 17% slower.
  for (int v = 1; v <= 2; v++)
  {
    const int per_cu = stream_blocks_per_cu(s, v);
    if (per_cu <= 0) continue;
    const long slots = (long)per_cu * WPB * cus;
    const long chunks = slots / strips;
    if (chunks < 1) continue;
    const int rows = (int)((s->ny + chunks - 1) / chunks);
    if (rows < 8 || rows > 128) continue;
    const std::pair<int, int> c{rows, v};
    if (std::find(s->tune_cand.begin(), s->tune_cand.end(), c) == s->tune_cand.end())
      s->tune_cand.push_back(c);
  }
  const int n_probe = 1 + 2 * (int)s->tune_cand.size(); // warm-up + 2 rounds
  s->tune_ev.assign(2 * n_probe, nullptr);
  for (hipEvent_t &e : s->tune_ev)
    if (hipEventCreate(&e) != hipSuccess)
    {
      (void)hipGetLastError();
      for (hipEvent_t &f : s->tune_ev)
        if (f) (void)hipEventDestroy(f);
      s->tune_ev.clear();
      return; // no tuning: keep the default shape
    }
  s->tune_pos = 0;
}

static void tune_stop(hmg_session *s)
{
  for (hipEvent_t e : s->tune_ev) (void)hipEventDestroy(e);
  s->tune_ev.clear();
  s->tune_pos = -1;
}

// Picks the best probed shape from the probes recorded so far (blocks until
// they have run): all of them when the probes are complete, or at least one
// round when a session ends while still probing (a drop-in call shorter
// than the probes), so that the next session of the shape starts tuned.
static void tune_finish(hmg_session *s)
{
  float     best = 0.f;
  int       bi = -1;
  const int nc = (int)s->tune_cand.size();
  const int done = s->tune_pos; // probes recorded: 0 .. done-1
  if (done > 0 && hipEventSynchronize(s->tune_ev[2 * (done - 1) + 1]) == hipSuccess)
    for (int k = 0; k < nc; k++)
    {
      float t = -1.f;
      for (int r = 0; r < 2; r++)
      {
        const int q = 1 + r * nc + k;
        if (q >= done) continue;
        float     ms = 0.f;
        if (hipEventElapsedTime(&ms, s->tune_ev[2 * q], s->tune_ev[2 * q + 1]) == hipSuccess)
          t = t < 0.f ? ms : std::min(t, ms);
      }
      if (t >= 0.f && (bi < 0 || t < best))
      {
        best = t;
        bi = k;
      }
    }
  (void)hipGetLastError();
  if (bi >= 0)
  {
    s->chunk_rows = s->tune_cand[bi].first;
    s->prefetch = s->tune_cand[bi].second;
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shape_cache[tune_key(s)] = {s->chunk_rows, s->prefetch};
  }
  tune_stop(s);
}

static int dispatch_step(hmg_session *s);

// One hydraulic iteration, as a probe while the session is tuning.
static int tuned_step(hmg_session *s)
{
  if (s->tune_pos < 0 || s->iter < TUNE_AFTER) return dispatch_step(s);
  const int nc = (int)s->tune_cand.size();
  if (s->tune_pos >= 1 + 2 * nc)
  {
    tune_finish(s);
    return dispatch_step(s);
  }
  const int q = s->tune_pos++;
  if (q > 0)
  {
    const std::pair<int, int> &c = s->tune_cand[(q - 1) % nc];
    s->chunk_rows = c.first;
    s->prefetch = c.second;
  }
  HMG_HIP(hipEventRecord(s->tune_ev[2 * q], s->ctx->stream));
  HMG_TRY(dispatch_step(s));
  HMG_HIP(hipEventRecord(s->tune_ev[2 * q + 1], s->ctx->stream));
  return HMG_OK;
}

static int session_create_common(hmg_ctx                 *ctx,
                                 const float             *z,
                                 bool                     z_on_dev,
                                 int                      nx,
                                 int                      ny,
                                 const hmg_vpipes_params *p,
                                 const float             *rain,
                                 bool                     rain_on_dev,
                                 int                      storage,
                                 hmg_session            **out,
                                 bool                     allow_zd = true)
{
  clear_error();
  if (!out) return set_error(HMG_ERR_INVALID, "null output pointer");
  *out = nullptr;
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z) return set_error(HMG_ERR_INVALID, "null z");
  if (nx < 1 || ny < 1) return set_error(HMG_ERR_INVALID, "shape must be positive");
  if ((size_t)nx * (size_t)ny > ((size_t)1 << 31))
    return set_error(HMG_ERR_INVALID, "grid larger than 2^31 cells");
  if (storage != HMG_STORAGE_F32 && storage != HMG_STORAGE_BF16 && storage != HMG_STORAGE_F16)
    return set_error(HMG_ERR_INVALID, "unknown storage %d", storage);
  hmg_vpipes_params prm;
  if (p)
    prm = *p;
  else
    hmg_vpipes_defaults(&prm);
  HMG_TRY(validate_vp_params(prm));
  if ((size_t)nx * (size_t)ny > STREAM_MAX_CELLS) prm.mode = HMG_VP_PLAIN;

  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  hmg_session *s = new hmg_session();
  s->ctx = ctx;
  s->nx = nx;
  s->ny = ny;
  s->storage = storage;
  s->p = prm;
  s->k = make_vp_const(prm, nx, ny);
  s->esize = storage == HMG_STORAGE_F32 ? 4 : 2;
  // FP32 state in the fused form: heights and water depths in one float2
  // plane (one 8-byte stream each way per cell instead of two 4-byte ones)
  s->zd = allow_zd && storage == HMG_STORAGE_F32 && prm.mode == HMG_VP_FUSED;
  // Launch shape of the fused kernel: at 4096^2 the measured best is 64 rows
  // per wavefront for FP32 state and 48 for 16-bit state (loads one row
  // ahead).  Smaller grids get shorter chunks so that a launch still has
  // about 16 wavefronts per CU (never fewer than 8 rows per wavefront).
  {
    const long strips = (nx + STRIP_OUT - 1) / STRIP_OUT;
    const long target = 16L * (ctx->num_cus > 0 ? ctx->num_cus : 256);
    long       ch = (ny * strips + target - 1) / target;
    const long cap = storage == HMG_STORAGE_F32 ? 64 : 48;
    s->chunk_rows = (int)(ch < 8 ? 8 : (ch > cap ? cap : ch));
  }
  s->prefetch = 1;
  const size_t n = (size_t)nx * ny;
  int          rc = HMG_OK;
  DevBuf       zt, Dt; // float2 layout: the input heights and depths before the merge
  for (int q = 0; q < 2 && rc == HMG_OK; q++)
  {
    rc = s->z[q].get(ctx->pool, n * (s->zd ? 8 : 4));
    if (rc == HMG_OK && !s->zd) rc = s->D[q].get(ctx->pool, n * s->esize);
    if (rc == HMG_OK) rc = s->Sp[q].get(ctx->pool, n * s->esize);
    if (rc == HMG_OK) rc = s->F[q].get(ctx->pool, n * s->esize * 4);
  }
  if (rc == HMG_OK && s->zd) rc = zt.get(ctx->pool, n * 4);
  if (rc == HMG_OK && s->zd) rc = Dt.get(ctx->pool, n * 4);
  if (rc == HMG_OK) rc = s->slots.get(ctx->pool, sizeof(Fix128) * 4 * NSLOT);
  if (rc == HMG_OK && rain) rc = s->rain.get(ctx->pool, n * 4);
  if (rc == HMG_OK && prm.mode == HMG_VP_PLAIN)
  {
    rc = s->d1.get(ctx->pool, n * 4);
    if (rc == HMG_OK) rc = s->s.get(ctx->pool, n * 4);
    if (rc == HMG_OK) rc = s->u.get(ctx->pool, n * 4);
    if (rc == HMG_OK) rc = s->v.get(ctx->pool, n * 4);
  }
  if (rc != HMG_OK)
  {
    delete s;
    return rc;
  }
  hipStream_t st = ctx->stream;
  hipError_t  e = hipMemcpyAsync(s->zd ? zt.ptr : s->z[0].ptr, z, n * 4,
                                 z_on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st);
  if (e == hipSuccess && rain)
    e = hipMemcpyAsync(s->rain.ptr, rain, n * 4,
                       rain_on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st);
  if (e != hipSuccess)
  {
    delete s;
    return set_error(HMG_ERR_DEVICE, "state upload failed: %s", hipGetErrorString(e));
  }
  switch (storage)
  {
  case HMG_STORAGE_BF16: rc = session_init_state<StBF16>(s, s->D[0].ptr); break;
  case HMG_STORAGE_F16: rc = session_init_state<StF16>(s, s->D[0].ptr); break;
  default: rc = session_init_state<StF32>(s, s->zd ? Dt.ptr : s->D[0].ptr); break;
  }
  if (rc == HMG_OK && s->zd)
  {
    rc = launch_zd_merge((float2 *)s->z[0].ptr, zt.as<float>(), 0, n, st);
    if (rc == HMG_OK) rc = launch_zd_merge((float2 *)s->z[0].ptr, Dt.as<float>(), 1, n, st);
    // the input planes return to the pool when this call ends
    if (rc == HMG_OK && hipStreamSynchronize(st) != hipSuccess)
      rc = set_error(HMG_ERR_DEVICE, "state merge failed");
  }
  if (rc != HMG_OK)
  {
    delete s;
    return rc;
  }
  tune_init(s);
  *out = s;
  return HMG_OK;
}

// ------------------------------------------------------------ finite check
//
// A NaN or an infinity in a state (a non-finite input cell, for one) would
// otherwise come back as an ordinary result: a NaN height stays NaN through
// every pass while max/min clamps keep its neighbours finite.  Export, the
// sums and the server's job completion therefore scan the state once.  A
// word is non-finite when its exponent bits are all ones (FP32 0x7f800000,
// BF16 0x7f80, FP16 0x7c00).  The flag only goes from 0 to 1, by plain
// stores of the same value.

__global__ void k_flag_clear(unsigned int *f)
{
  *f = 0u;
}

template <class W>
__global__ void __launch_bounds__(256) k_nonfinite(const W *__restrict__ a,
                                                   size_t n,
                                                   W      mask,
                                                   unsigned int *flag)
{
  bool bad = false;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    bad |= (a[c] & mask) == mask;
  if (bad) *flag = 1u;
}

template <class W>
static int launch_nonfinite(const void *p, size_t n, W mask, unsigned int *flag, hipStream_t st)
{
  const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  if (nb > 0) k_nonfinite<W><<<nb, 256, 0, st>>>((const W *)p, n, mask, flag);
  return check_launch("nonfinite");
}

// HMG_ERR_NONFINITE if a float plane of n values on the context's device
// holds a NaN or an infinity.  Waits for the context stream.
int check_finite_f32(hmg_ctx *ctx, const float *p, size_t n)
{
  hipStream_t st = ctx->stream;
  DevBuf      flag;
  HMG_TRY(flag.get(ctx->pool, sizeof(unsigned int)));
  unsigned int *f = flag.as<unsigned int>();
  k_flag_clear<<<1, 1, 0, st>>>(f);
  HMG_TRY(check_launch("flag_clear"));
  HMG_TRY(launch_nonfinite<uint32_t>(p, n, 0x7f800000u, f, st));
  unsigned int h = 0;
  HMG_HIP(hipMemcpyAsync(&h, f, sizeof(h), hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  if (h) return set_error(HMG_ERR_NONFINITE, "the result holds a NaN or an infinity");
  return HMG_OK;
}

// HMG_ERR_NONFINITE if z, the water, the sediment or a flux of the
// session's current state holds a NaN or an infinity.  Waits for the stream.
int session_check_finite(hmg_session *s)
{
  const size_t n = (size_t)s->nx * s->ny;
  hipStream_t  st = s->ctx->stream;
  DevBuf       flag;
  HMG_TRY(flag.get(s->ctx->pool, sizeof(unsigned int)));
  unsigned int *f = flag.as<unsigned int>();
  k_flag_clear<<<1, 1, 0, st>>>(f);
  HMG_TRY(check_launch("flag_clear"));
  const int c = s->cur;
  // the float2 layout: heights and depths in one plane of 2n floats
  HMG_TRY(launch_nonfinite<uint32_t>(s->z[c].ptr, s->zd ? 2 * n : n, 0x7f800000u, f, st));
  if (s->storage == HMG_STORAGE_F32)
  {
    if (!s->zd) HMG_TRY(launch_nonfinite<uint32_t>(s->D[c].ptr, n, 0x7f800000u, f, st));
    HMG_TRY(launch_nonfinite<uint32_t>(s->Sp[c].ptr, n, 0x7f800000u, f, st));
    HMG_TRY(launch_nonfinite<uint32_t>(s->F[c].ptr, 4 * n, 0x7f800000u, f, st));
  }
  else
  {
    const uint16_t m = s->storage == HMG_STORAGE_BF16 ? 0x7f80u : 0x7c00u;
    HMG_TRY(launch_nonfinite<uint16_t>(s->D[c].ptr, n, m, f, st));
    HMG_TRY(launch_nonfinite<uint16_t>(s->Sp[c].ptr, n, m, f, st));
    HMG_TRY(launch_nonfinite<uint16_t>(s->F[c].ptr, 4 * n, m, f, st));
  }
  unsigned int h = 0;
  HMG_HIP(hipMemcpyAsync(&h, f, sizeof(h), hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  if (h)
    return set_error(HMG_ERR_NONFINITE,
                     "the %dx%d state holds a NaN or an infinity after %ld iterations", s->nx,
                     s->ny, s->iter);
  return HMG_OK;
}

} // namespace hmg

extern "C" {

size_t hmg_session_bytes(int nx, int ny, int storage)
{
  size_t n = (size_t)(nx > 0 ? nx : 0) * (size_t)(ny > 0 ? ny : 0);
  size_t e = storage == HMG_STORAGE_F32 ? 4 : 2;
  return 2 * n * (4 + e * 6) + sizeof(Fix128) * 4 * NSLOT;
}

int hmg_session_create(hmg_ctx                 *ctx,
                       const float             *z_host,
                       int                      nx,
                       int                      ny,
                       const hmg_vpipes_params *p,
                       const float             *rain_host,
                       int                      storage,
                       hmg_session            **out)
{
  return session_create_common(ctx, z_host, false, nx, ny, p, rain_host, false, storage, out);
}

int hmg_session_create_dev(hmg_ctx                 *ctx,
                           const float             *z_dev,
                           int                      nx,
                           int                      ny,
                           const hmg_vpipes_params *p,
                           const float             *rain_dev,
                           int                      storage,
                           hmg_session            **out)
{
  return session_create_common(ctx, z_dev, true, nx, ny, p, rain_dev, true, storage, out);
}

} // extern "C"

namespace hmg
{

int session_create_separate(hmg_ctx                 *ctx,
                            const float             *z,
                            bool                     z_on_dev,
                            int                      nx,
                            int                      ny,
                            const hmg_vpipes_params *p,
                            int                      storage,
                            hmg_session            **out)
{
  return session_create_common(ctx, z, z_on_dev, nx, ny, p, nullptr, false, storage, out, false);
}

} // namespace hmg

extern "C" {

int hmg_session_vpipes(hmg_session *s, int iterations)
{
  clear_error();
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  HMG_TRY(ctx_enter(s->ctx));
  for (int it = 0; it < iterations; it++) HMG_TRY(tuned_step(s));
  return HMG_OK;
}

int hmg_session_thermal(hmg_session *s, float talus, int iterations)
{
  clear_error();
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  if (s->nx < 3 || s->ny < 3) return set_error(HMG_ERR_INVALID, "thermal needs nx, ny >= 3");
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  HMG_TRY(ctx_enter(s->ctx));
  // the other z plane is free between iterations: after an odd number of
  // launches the result is there, and the two planes trade places
  if (s->zd)
  {
    float2 *zc = (float2 *)s->z[s->cur].ptr, *res = zc;
    HMG_TRY(launch_thermal_zd(s->ctx, zc, (float2 *)s->z[s->cur ^ 1].ptr, s->nx, s->ny, talus,
                              iterations, s->ctx->stream, &res));
    if (res != zc) std::swap(s->z[s->cur].ptr, s->z[s->cur ^ 1].ptr);
    return HMG_OK;
  }
  float *zc = s->z[s->cur].as<float>();
  float *zt = s->z[s->cur ^ 1].as<float>();
  float *res = zc;
  HMG_TRY(launch_thermal(s->ctx, zc, zt, s->nx, s->ny, nullptr, talus, iterations, nullptr,
                         s->ctx->stream, &res, true));
  if (res != zc) std::swap(s->z[s->cur].ptr, s->z[s->cur ^ 1].ptr);
  return HMG_OK;
}

int hmg_session_export(hmg_session *s, float *z_host, float *d_host, float *s_host,
                       float *u_host, float *v_host)
{
  clear_error();
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  HMG_TRY(ctx_enter(s->ctx));
  const size_t n = (size_t)s->nx * s->ny;
  hipStream_t  st = s->ctx->stream;
  DevBuf       zh; // float2 layout: the heights split out
  if (z_host)
  {
    const float *zp = nullptr;
    HMG_TRY(session_heights(s, zh, &zp));
    HMG_HIP(hipMemcpyAsync(z_host, zp, n * 4, hipMemcpyDeviceToHost, st));
  }
  if (d_host || s_host || u_host || v_host)
  {
    DevBuf out;
    HMG_TRY(out.get(s->ctx->pool, n * 16));
    float *o = out.as<float>();
    int    rc;
    switch (s->storage)
    {
    case HMG_STORAGE_BF16: rc = session_export<StBF16>(s, o, o + n, o + 2 * n, o + 3 * n); break;
    case HMG_STORAGE_F16: rc = session_export<StF16>(s, o, o + n, o + 2 * n, o + 3 * n); break;
    default: rc = session_export<StF32>(s, o, o + n, o + 2 * n, o + 3 * n); break;
    }
    HMG_TRY(rc);
    float *dst[4] = {d_host, s_host, u_host, v_host};
    for (int q = 0; q < 4; q++)
      if (dst[q]) HMG_HIP(hipMemcpyAsync(dst[q], o + q * n, n * 4, hipMemcpyDeviceToHost, st));
    HMG_HIP(hipStreamSynchronize(st));
  }
  HMG_HIP(hipStreamSynchronize(st));
  // the planes are delivered either way; a non-finite state is reported
  return session_check_finite(s);
}

// max(|u|, |v|) over n cells as the bit pattern of a non-negative float: for
// those, unsigned order is float order (NaN above +inf), so the result does
// not depend on the order of the atomic maxima.
__global__ void __launch_bounds__(256) k_absmax2(const float *__restrict__ u,
                                                 const float *__restrict__ v,
                                                 size_t n,
                                                 unsigned int *out)
{
  unsigned int m = 0;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
  {
    m = max(m, __float_as_uint(u[c]) & 0x7fffffffu);
    m = max(m, __float_as_uint(v[c]) & 0x7fffffffu);
  }
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_zero_u32(unsigned int *p)
{
  *p = 0u;
}

int hmg_session_max_backtrace(hmg_session *s, double *max_backtrace)
{
  clear_error();
  if (!s || !max_backtrace) return set_error(HMG_ERR_INVALID, "null argument");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  HMG_TRY(ctx_enter(s->ctx));
  const size_t n = (size_t)s->nx * s->ny;
  hipStream_t  st = s->ctx->stream;
  DevBuf       out, mx;
  HMG_TRY(out.get(s->ctx->pool, n * 16));
  HMG_TRY(mx.get(s->ctx->pool, sizeof(unsigned int)));
  float *o = out.as<float>();
  int    rc;
  switch (s->storage)
  {
  case HMG_STORAGE_BF16: rc = session_export<StBF16>(s, o, o + n, o + 2 * n, o + 3 * n); break;
  case HMG_STORAGE_F16: rc = session_export<StF16>(s, o, o + n, o + 2 * n, o + 3 * n); break;
  default: rc = session_export<StF32>(s, o, o + n, o + 2 * n, o + 3 * n); break;
  }
  HMG_TRY(rc);
  unsigned int *m = mx.as<unsigned int>();
  k_zero_u32<<<1, 1, 0, st>>>(m);
  HMG_TRY(check_launch("zero_u32"));
  int nb = (int)((n + 255) / 256);
  if (nb > 2048) nb = 2048;
  k_absmax2<<<nb, 256, 0, st>>>(o + 2 * n, o + 3 * n, n, m);
  HMG_TRY(check_launch("absmax2"));
  unsigned int h = 0;
  HMG_HIP(hipMemcpyAsync(&h, m, sizeof(h), hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  float mf;
  std::memcpy(&mf, &h, sizeof(mf));
  // dt * max|u| in FP64: the product of two floats is exact there, and the
  // maximum of dt |u| over the cells is dt times the maximum of |u| (dt > 0)
  *max_backtrace = (double)s->k.dt * (double)mf;
  return HMG_OK;
}

int hmg_session_sums(hmg_session *s, double *sum_z, double *sum_s, double *sum_d)
{
  clear_error();
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  HMG_TRY(ctx_enter(s->ctx));
  const size_t n = (size_t)s->nx * s->ny;
  hipStream_t  st = s->ctx->stream;
  DevBuf       out, acc;
  HMG_TRY(out.get(s->ctx->pool, n * 16));
  HMG_TRY(acc.get(s->ctx->pool, sizeof(Fix128) * 3 * NSLOT));
  float *o = out.as<float>();
  int    rc;
  switch (s->storage)
  {
  case HMG_STORAGE_BF16: rc = session_export<StBF16>(s, o, o + n, o + 2 * n, o + 3 * n); break;
  case HMG_STORAGE_F16: rc = session_export<StF16>(s, o, o + n, o + 2 * n, o + 3 * n); break;
  default: rc = session_export<StF32>(s, o, o + n, o + 2 * n, o + 3 * n); break;
  }
  HMG_TRY(rc);
  Fix128 *sets = acc.as<Fix128>();
  k_zero_fix<<<1, 64, 0, st>>>(sets, 3 * NSLOT);
  HMG_TRY(check_launch("zero_fix"));
  int nb = (int)((n + 255) / 256);
  if (nb > 2048) nb = 2048;
  DevBuf       zh;
  const float *zp = nullptr;
  HMG_TRY(session_heights(s, zh, &zp));
  const float *src[3] = {zp, o + n, o};
  for (int q = 0; q < 3; q++)
  {
    k_vp_sum<StF32><<<nb, 256, 0, st>>>(src[q], n, 1.f, 0, sets + q * NSLOT, nullptr);
    HMG_TRY(check_launch("vp_sum"));
  }
  Fix128 h[3 * NSLOT];
  HMG_HIP(hipMemcpyAsync(h, sets, sizeof(h), hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  double *dst[3] = {sum_z, sum_s, sum_d};
  for (int q = 0; q < 3; q++)
  {
    Fix128 a{0, 0};
    for (int r = 0; r < NSLOT; r++)
    {
      const Fix128 &b = h[q * NSLOT + r];
      uint64_t      lo = a.lo + b.lo;
      uint64_t      c = lo < a.lo ? 1 : 0;
      a.lo = lo;
      a.hi = (int64_t)((uint64_t)a.hi + (uint64_t)b.hi + c);
    }
    if (dst[q]) *dst[q] = fix_to_double(a.lo, a.hi);
  }
  // a fixed-point sum of non-finite values is meaningless
  return session_check_finite(s);
}

int hmg_session_set_tuning(hmg_session *s, int chunk_rows, int prefetch)
{
  clear_error();
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  if (chunk_rows < 1 || chunk_rows > 65536)
    return set_error(HMG_ERR_INVALID, "chunk_rows must be in [1, 65536]");
  if (prefetch < 0 || prefetch > 2)
    return set_error(HMG_ERR_INVALID, "prefetch must be 0, 1 or 2");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  tune_stop(s); // the caller's shape wins over online tuning
  s->chunk_rows = chunk_rows;
  s->prefetch = prefetch;
  return HMG_OK;
}

int hmg_session_get_tuning(hmg_session *s, int *chunk_rows, int *prefetch, int *tuning)
{
  clear_error();
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  if (chunk_rows) *chunk_rows = s->chunk_rows;
  if (prefetch) *prefetch = s->prefetch;
  if (tuning) *tuning = s->tune_pos >= 0 ? 1 : 0;
  return HMG_OK;
}

float *hmg_session_z_dev(hmg_session *s)
{
  if (!s) return nullptr;
  if (!s->zd) return s->z[s->cur].as<float>();
  // float2 layout: a copy of the heights in the session's scratch plane,
  // made in stream order (valid until the session's next call)
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  const size_t n = (size_t)s->nx * s->ny;
  if (ctx_enter(s->ctx) != HMG_OK || s->zscratch.get(s->ctx->pool, n * 4) != HMG_OK ||
      launch_zd_split((const float2 *)s->z[s->cur].ptr, s->zscratch.as<float>(), 0, n,
                      s->ctx->stream) != HMG_OK)
    return nullptr;
  return s->zscratch.as<float>();
}

int hmg_session_sync(hmg_session *s)
{
  if (!s) return set_error(HMG_ERR_INVALID, "null session");
  return hmg_sync(s->ctx);
}

int hmg_session_destroy(hmg_session *s)
{
  if (!s) return HMG_OK;
  std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  // a session that ends after a full round of probes keeps their result
  if (s->tune_pos > (int)s->tune_cand.size()) tune_finish(s);
  delete s;
  return HMG_OK;
}

int hmg_vpipes(hmg_ctx                 *ctx,
               float                   *z_host,
               int                      nx,
               int                      ny,
               const hmg_vpipes_params *p,
               const float             *rain_host,
               float                   *d_out_host,
               float                   *s_out_host,
               float                   *u_out_host,
               float                   *v_out_host)
{
  hmg_vpipes_params prm;
  if (p)
    prm = *p;
  else
    hmg_vpipes_defaults(&prm);
  hmg_session *s = nullptr;
  HMG_TRY(hmg_session_create(ctx, z_host, nx, ny, &prm, rain_host, HMG_STORAGE_F32, &s));
  int rc = hmg_session_vpipes(s, prm.iterations);
  if (rc == HMG_OK) rc = hmg_session_export(s, z_host, d_out_host, s_out_host, u_out_host,
                                            v_out_host);
  hmg_session_destroy(s);
  return rc;
}

} // extern "C"

// ------------------------------------------------------------------ batch

namespace hmg
{

// Advances sessions[i] by iters[i] iterations with batched launches (one
// launch per iteration for all maps still running).  All sessions share one
// storage type.  With use_graph the whole launch sequence is captured in a
// hipGraph and submitted as one graph.  Returns the elapsed device time of
// the launches in *ms (if not null).
// The batch's launches t = 0 .. T-1 as one cooperative launch of
// k_vp_batch_persist (d: the map table in host memory, ordered by iteration
// count; prefix[k]: items of the first k maps).  Events e0/e1 (may be null)
// bracket the launch.  Waits for it.  HMG_ERR_UNSUPPORTED: the runtime
// refused the launch, nothing ran; HMG_ERR_DEVICE: a barrier wait gave up
// (the maps' state is then not valid).
template <class S, bool PF, bool KEXP1, bool GA>
static int batch_persistent(hmg_ctx                      *ctx,
                            const VpMapDesc              *table,
                            const std::vector<VpMapDesc> &d,
                            const std::vector<int>       &prefix,
                            int                           T,
                            bool                          any_rain,
                            hipEvent_t                    e0,
                            hipEvent_t                    e1)
{
  hipStream_t      st = ctx->stream;
  std::vector<int> sched(2 * (size_t)(T > 0 ? T : 1), 0); // items, then active maps, per t
  int              active = (int)d.size(), max_items = 0;
  for (int t = 0; t < T; t++)
  {
    while (active > 0 && d[active - 1].iters <= t) active--;
    sched[t] = prefix[active];
    sched[T + t] = active;
    max_items = std::max(max_items, prefix[active]);
  }
  if (max_items == 0)
  {
    if (e0) HMG_HIP(hipEventRecord(e0, st));
    if (e1) HMG_HIP(hipEventRecord(e1, st));
    HMG_HIP(hipStreamSynchronize(st));
    return HMG_OK;
  }
  const void *f = any_rain ? (const void *)k_vp_batch_persist<S, false, PF, KEXP1, GA>
                           : (const void *)k_vp_batch_persist<S, true, PF, KEXP1, GA>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * WPB, 0) != hipSuccess ||
      per_cu < 1)
    return set_error(HMG_ERR_DEVICE, "persistent batch kernel: no resident workgroup");
  long G = (long)per_cu * ctx->num_cus;
  const long need = (max_items + WPB - 1) / WPB;
  if (G > need) G = need;
  if (G >= 8) G -= G % 8;
  DevBuf buf;
  HMG_TRY(buf.get(ctx->pool, sizeof(GridBar) + sizeof(int) * sched.size()));
  GridBar *bar = buf.as<GridBar>();
  int     *items_at = (int *)((char *)buf.ptr + sizeof(GridBar));
  int     *active_at = items_at + T;
  GridBar  init{0u, 0u, 0u, GRID_SPIN_LIMIT};
  if (const char *env = std::getenv("HMG_GRID_SPIN_LIMIT"))
    init.limit = (unsigned)std::max(1L, std::strtol(env, nullptr, 10));
  HMG_HIP(hipMemcpyAsync(bar, &init, sizeof(GridBar), hipMemcpyHostToDevice, st));
  HMG_HIP(hipMemcpyAsync(items_at, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice,
                         st));
  void *args[] = {(void *)&table, (void *)&items_at, (void *)&active_at, (void *)&T, (void *)&bar};
  if (e0) HMG_HIP(hipEventRecord(e0, st));
  const hipError_t e = hipLaunchCooperativeKernel(f, dim3((unsigned)G), dim3(64 * WPB), args, 0, st);
  if (e != hipSuccess)
  {
    // nothing ran: the maps' state is unchanged (a second process on the
    // device has been refused the cooperative launch this way)
    (void)hipGetLastError();
    (void)hipStreamSynchronize(st);
    return set_error(HMG_ERR_UNSUPPORTED, "persistent batch launch refused: %s",
                     hipGetErrorString(e));
  }
  if (e1) HMG_HIP(hipEventRecord(e1, st));
  GridBar h{};
  HMG_HIP(hipMemcpyAsync(&h, bar, sizeof(GridBar), hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  if (h.err)
    return set_error(HMG_ERR_DEVICE,
                     "persistent batch: a grid barrier wait gave up; the maps' state is not valid");
  return HMG_OK;
}

template <class S, bool PF, bool KEXP1, bool GA = false>
static int batch_advance_t(hmg_ctx           *ctx,
                           hmg_session *const *ss,
                           const int          *iters,
                           int                 n,
                           int                 use_graph,
                           int                 chunk_rows,
                           float              *ms)
{
  hipStream_t      st = ctx->stream;
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return iters[a] > iters[b]; });
  std::vector<VpMapDesc> d(n);
  std::vector<int>       prefix(n + 1, 0); // items of the first k maps
  bool                   any_rain = false;
  int                    T = 0;
  for (int q = 0; q < n; q++)
  {
    hmg_session *s = ss[order[q]];
    VpMapDesc   &m = d[q];
    for (int h = 0; h < 2; h++)
    {
      m.z[h] = s->z[h].ptr;
      m.D[h] = s->D[h].ptr;
      m.F[h] = s->F[h].ptr;
      m.Sp[h] = s->Sp[h].ptr;
    }
    m.rain = s->rain.as<float>();
    any_rain = any_rain || m.rain;
    m.slots = s->slots.as<Fix128>();
    m.k = s->k;
    m.W0 = s->W0;
    m.rf = s->rf;
    m.topup = s->p.maintain_water_volume ? 1 : 0;
    m.strips = (s->nx + STRIP_OUT - 1) / STRIP_OUT;
    m.chunk_rows = chunk_rows < s->ny ? chunk_rows : s->ny;
    m.n_items = m.strips * ((s->ny + m.chunk_rows - 1) / m.chunk_rows);
    m.item_base = prefix[q];
    prefix[q + 1] = prefix[q] + m.n_items;
    m.t_start = 0;
    m.it0 = (int)s->iter;
    m.iters = iters[order[q]];
    if (m.iters > T) T = m.iters;
  }
  DevBuf table;
  HMG_TRY(table.get(ctx->pool, sizeof(VpMapDesc) * (size_t)(n > 0 ? n : 1)));
  if (n > 0)
    HMG_HIP(hipMemcpyAsync(table.ptr, d.data(), sizeof(VpMapDesc) * n, hipMemcpyHostToDevice, st));
  HMG_HIP(hipStreamSynchronize(st)); // d (pageable) may be released after this

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms)
  {
    HMG_HIP(hipEventCreate(&e0));
    HMG_HIP(hipEventCreate(&e1));
  }
  if (use_graph == HMG_BATCH_PERSISTENT)
  {
    HMG_TRY((batch_persistent<S, PF, KEXP1, GA>(ctx, table.as<VpMapDesc>(), d, prefix, T,
                                                any_rain, e0, e1)));
    if (ms)
    {
      HMG_HIP(hipEventElapsedTime(ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    for (int i = 0; i < n; i++)
    {
      ss[i]->iter += iters[i];
      ss[i]->cur = (int)(ss[i]->iter & 1);
    }
    return HMG_OK;
  }
  hipGraph_t     graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph) HMG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  else if (ms)
    HMG_HIP(hipEventRecord(e0, st));
  int active = n;
  for (int t = 0; t < T; t++)
  {
    while (active > 0 && d[active - 1].iters <= t) active--;
    const int items = prefix[active];
    if (items == 0) continue;
    const int G = xcd_grid(items);
    const VpMapDesc *tab = table.as<VpMapDesc>();
    if (any_rain)
      k_vp_batch<S, false, PF, KEXP1, GA><<<G, 64 * WPB, 0, st>>>(tab, active, items, t);
    else
      k_vp_batch<S, true, PF, KEXP1, GA><<<G, 64 * WPB, 0, st>>>(tab, active, items, t);
  }
  int rc = check_launch("vp_batch");
  if (use_graph)
  {
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc != HMG_OK || e != hipSuccess)
    {
      if (graph) (void)hipGraphDestroy(graph);
      return rc != HMG_OK ? rc : set_error(HMG_ERR_DEVICE, "graph capture failed: %s",
                                           hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess && ms) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e != hipSuccess)
    {
      if (exec) (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      return set_error(HMG_ERR_DEVICE, "graph launch failed: %s", hipGetErrorString(e));
    }
  }
  if (ms) HMG_HIP(hipEventRecord(e1, st));
  HMG_HIP(hipStreamSynchronize(st));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ms)
  {
    HMG_HIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  HMG_TRY(rc);
  for (int i = 0; i < n; i++)
  {
    ss[i]->iter += iters[i];
    ss[i]->cur = (int)(ss[i]->iter & 1);
  }
  return HMG_OK;
}

int batch_advance(hmg_ctx           *ctx,
                  hmg_session *const *ss,
                  const int          *iters,
                  int                 n,
                  int                 use_graph,
                  int                 chunk_rows,
                  int                 prefetch,
                  float              *ms)
{
  if (n <= 0) return HMG_OK;
  const int  storage = ss[0]->storage;
  const bool ke1 = ss[0]->k.kexp == 1.f;
  const bool zd = ss[0]->zd;
  for (int i = 0; i < n; i++)
  {
    if (ss[i]->storage != storage || ss[i]->zd != zd)
      return set_error(HMG_ERR_INVALID, "batched maps must share one storage type and layout");
    if ((ss[i]->k.kexp == 1.f) != ke1)
      return set_error(HMG_ERR_INVALID, "batched maps must share k_discharge_exp == 1 or != 1");
    if (iters[i] < 0) return set_error(HMG_ERR_INVALID, "negative iteration count");
    if ((size_t)ss[i]->nx * ss[i]->ny > STREAM_MAX_CELLS)
      return set_error(HMG_ERR_INVALID, "batched maps are limited to 2^28 cells");
    if ((long)ss[i]->iter + iters[i] > (1L << 30))
      return set_error(HMG_ERR_INVALID, "iteration count too large");
  }
  if (chunk_rows < 1) return set_error(HMG_ERR_INVALID, "chunk_rows must be >= 1");
  if (prefetch < 0 || prefetch > 2) return set_error(HMG_ERR_INVALID, "prefetch must be 0, 1 or 2");
#define HMG_BATCH(ST)                                                          \
  (prefetch == 2                                                               \
       ? (ke1 ? batch_advance_t<ST, true, true, true>(ctx, ss, iters, n, use_graph, chunk_rows, ms) \
              : batch_advance_t<ST, true, false, true>(ctx, ss, iters, n, use_graph, chunk_rows, ms)) \
   : prefetch                                                                  \
       ? (ke1 ? batch_advance_t<ST, true, true>(ctx, ss, iters, n, use_graph, chunk_rows, ms) \
              : batch_advance_t<ST, true, false>(ctx, ss, iters, n, use_graph, chunk_rows, ms)) \
       : (ke1 ? batch_advance_t<ST, false, true>(ctx, ss, iters, n, use_graph, chunk_rows, ms) \
              : batch_advance_t<ST, false, false>(ctx, ss, iters, n, use_graph, chunk_rows, ms)))
  switch (storage)
  {
  case HMG_STORAGE_BF16: return HMG_BATCH(StBF16);
  case HMG_STORAGE_F16: return HMG_BATCH(StF16);
  default: return zd ? HMG_BATCH(StF32ZD) : HMG_BATCH(StF32);
  }
#undef HMG_BATCH
}

} // namespace hmg

// Autotuner cache: best (chunk_rows, prefetch, form) per (storage, size
// class), size class = floor(log2(total cells of the batch)).
namespace
{
struct BatchTune
{
  int rows, prefetch, form;
};
std::mutex                                g_tune_mu;
std::map<std::pair<int, int>, BatchTune> g_tune;
} // namespace

struct hmg_batch
{
  hmg_ctx                  *ctx = nullptr;
  int                       storage = HMG_STORAGE_F32;
  hmg_vpipes_params         p{};
  std::vector<hmg_session *> maps;
  std::vector<int>          iters;
  int                       chunk_rows = 48;
  int                       prefetch = 1;
  int                       form = HMG_BATCH_GRAPH; // what HMG_BATCH_AUTO runs
  float                     last_ms = 0.f;
};

extern "C" {

int hmg_batch_create(hmg_ctx *ctx, int storage, const hmg_vpipes_params *p, hmg_batch **out)
{
  clear_error();
  if (!ctx || !out) return set_error(HMG_ERR_INVALID, "null argument");
  if (storage != HMG_STORAGE_F32 && storage != HMG_STORAGE_BF16 && storage != HMG_STORAGE_F16)
    return set_error(HMG_ERR_INVALID, "unknown storage %d", storage);
  hmg_batch *b = new hmg_batch();
  b->ctx = ctx;
  b->storage = storage;
  if (p)
    b->p = *p;
  else
    hmg_vpipes_defaults(&b->p);
  b->p.mode = HMG_VP_FUSED;
  int rc = validate_vp_params(b->p);
  if (rc != HMG_OK)
  {
    delete b;
    return rc;
  }
  b->chunk_rows = storage == HMG_STORAGE_F32 ? 64 : 48;
  *out = b;
  return HMG_OK;
}

int hmg_batch_add(hmg_batch   *b,
                  const float *z_host,
                  int          nx,
                  int          ny,
                  const float *rain_host,
                  int          iterations,
                  int         *map_id)
{
  clear_error();
  if (!b) return set_error(HMG_ERR_INVALID, "null batch");
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  if (nx > 0 && ny > 0 && (size_t)nx * (size_t)ny > STREAM_MAX_CELLS)
    return set_error(HMG_ERR_INVALID, "batched maps are limited to 2^28 cells");
  hmg_session *s = nullptr;
  HMG_TRY(hmg_session_create(b->ctx, z_host, nx, ny, &b->p, rain_host, b->storage, &s));
  b->maps.push_back(s);
  b->iters.push_back(iterations);
  if (map_id) *map_id = (int)b->maps.size() - 1;
  return HMG_OK;
}

int hmg_batch_set_tuning(hmg_batch *b, int chunk_rows, int prefetch)
{
  clear_error();
  if (!b) return set_error(HMG_ERR_INVALID, "null batch");
  if (chunk_rows < 1 || prefetch < 0 || prefetch > 2)
    return set_error(HMG_ERR_INVALID, "bad tuning values");
  b->chunk_rows = chunk_rows;
  b->prefetch = prefetch;
  return HMG_OK;
}

} // extern "C"

// Workgroups of the batched kernel variant (prefetch 0: rows loaded when
// needed, 1: one row ahead, 2: gathers ahead) one CU holds at once, from the
// kernel's register use (0 if the runtime cannot tell).
template <class S>
static int batch_blocks_per_cu_t(bool rain, bool ke1, int pf)
{
  const void *f = nullptr;
#define HMG_BATCH_KERNEL(UR, PF, KE, GA) (const void *)k_vp_batch<S, UR, PF, KE, GA>
  if (pf == 2)
    f = rain ? (ke1 ? HMG_BATCH_KERNEL(false, true, true, true) : HMG_BATCH_KERNEL(false, true, false, true))
             : (ke1 ? HMG_BATCH_KERNEL(true, true, true, true) : HMG_BATCH_KERNEL(true, true, false, true));
  else if (pf == 1)
    f = rain ? (ke1 ? HMG_BATCH_KERNEL(false, true, true, false) : HMG_BATCH_KERNEL(false, true, false, false))
             : (ke1 ? HMG_BATCH_KERNEL(true, true, true, false) : HMG_BATCH_KERNEL(true, true, false, false));
  else
    f = rain ? (ke1 ? HMG_BATCH_KERNEL(false, false, true, false) : HMG_BATCH_KERNEL(false, false, false, false))
             : (ke1 ? HMG_BATCH_KERNEL(true, false, true, false) : HMG_BATCH_KERNEL(true, false, false, false));
#undef HMG_BATCH_KERNEL
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64 * WPB, 0) != hipSuccess)
  {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

static int batch_blocks_per_cu(const hmg_batch *b, int pf)
{
  bool rain = false;
  for (const hmg_session *s : b->maps) rain = rain || s->rain.ptr != nullptr;
  const bool ke1 = b->maps[0]->k.kexp == 1.f;
  if (b->storage == HMG_STORAGE_BF16) return batch_blocks_per_cu_t<StBF16>(rain, ke1, pf);
  if (b->storage == HMG_STORAGE_F16) return batch_blocks_per_cu_t<StF16>(rain, ke1, pf);
  return b->maps[0]->zd ? batch_blocks_per_cu_t<StF32ZD>(rain, ke1, pf)
                        : batch_blocks_per_cu_t<StF32>(rain, ke1, pf);
}

// Work items of one launch over every map of the batch at chunk length
// `rows` (capped at each map's height, as batch_advance_t does).
static long batch_items(const hmg_batch *b, int rows)
{
  long n = 0;
  for (const hmg_session *s : b->maps)
  {
    const int cr = rows < s->ny ? rows : s->ny;
    n += (long)((s->nx + STRIP_OUT - 1) / STRIP_OUT) * ((s->ny + cr - 1) / cr);
  }
  return n;
}

extern "C" {

// Times each candidate launch shape on `probe` iterations of the live maps
// (results do not depend on the shape, so the probe iterations are part of
// the run) and keeps the fastest; the choice is cached per storage type and
// size class for later batches.  The candidates are a fixed list of chunk
// lengths for each variant plus, per variant, the shortest chunks whose
// work items fill one, two or three rounds of the wavefront slots the
// kernel's occupancy gives (a nearly empty last round leaves most of the
// GPU idle for a whole item's time).  When every map has iterations to
// spare, the probes start after 20 ordinary iterations in the default
// shape: while the water spreads from rest the shapes rank differently from
// the rest of a run.
int hmg_batch_autotune(hmg_batch *b, int probe, int *chunk_rows_out, int *prefetch_out)
{
  clear_error();
  if (!b) return set_error(HMG_ERR_INVALID, "null batch");
  if (b->maps.empty()) return HMG_OK;
  std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
  HMG_TRY(ctx_enter(b->ctx));
  double cells = 0;
  for (hmg_session *s : b->maps) cells += (double)s->nx * s->ny;
  const int key_size = (int)std::log2(cells > 1 ? cells : 1);
  const auto key = std::make_pair(b->storage, key_size);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto                        it = g_tune.find(key);
    if (it != g_tune.end())
    {
      b->chunk_rows = it->second.rows;
      b->prefetch = it->second.prefetch;
      b->form = it->second.form;
      if (chunk_rows_out) *chunk_rows_out = b->chunk_rows;
      if (prefetch_out) *prefetch_out = b->prefetch;
      return HMG_OK;
    }
  }
  if (probe < 1) probe = 1;
  static const int                 cand_rows[] = {16, 32, 48, 64, 96};
  std::vector<std::pair<int, int>> cand; // (rows, prefetch)
  for (int pf = 0; pf < 3; pf++)         // rows loaded when needed, one ahead, gathers ahead
    for (int rows : cand_rows) cand.push_back({rows, pf});
  const int cus = b->ctx->num_cus > 0 ? b->ctx->num_cus : 256;
  for (int pf = 0; pf < 3; pf++)
  {
    const int per_cu = batch_blocks_per_cu(b, pf);
    if (per_cu <= 0) continue;
    const long slots = (long)per_cu * WPB * cus;
    for (int rounds = 1; rounds <= 3; rounds++)
    {
      int rows = 8;
      while (rows <= 128 && batch_items(b, rows) > rounds * slots) rows++;
      if (rows > 128) continue;
      const std::pair<int, int> c{rows, pf};
      if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
    }
  }
  float            best = 1e30f;
  int              best_rows = b->chunk_rows, best_pf = b->prefetch;
  std::vector<int> step(b->maps.size());
  {
    // 20 ordinary iterations first when every map can spare them besides
    // the probes (both forms included)
    const int need = probe * ((int)cand.size() + 2) + 20;
    int       min_left = 1 << 30;
    for (int it : b->iters) min_left = it < min_left ? it : min_left;
    if (min_left >= need)
    {
      for (size_t i = 0; i < step.size(); i++) step[i] = 20;
      HMG_TRY(batch_advance(b->ctx, b->maps.data(), step.data(), (int)step.size(), HMG_BATCH_GRAPH,
                            b->chunk_rows, b->prefetch, nullptr));
      for (size_t i = 0; i < step.size(); i++) b->iters[i] -= 20;
    }
  }
  // runs `probe` iterations of every map in one form and shape; false when
  // a map has fewer iterations left (every map's count stays exact)
  auto probe_run = [&](int form, int rows, int pf, float &ms, bool &ran) -> int
  {
    ran = false;
    int min_left = 1 << 30;
    for (int it : b->iters) min_left = it < min_left ? it : min_left;
    if (min_left < probe) return HMG_OK;
    for (size_t i = 0; i < step.size(); i++) step[i] = probe;
    HMG_TRY(batch_advance(b->ctx, b->maps.data(), step.data(), (int)step.size(), form, rows, pf,
                          &ms));
    for (size_t i = 0; i < step.size(); i++) b->iters[i] -= probe;
    ran = true;
    return HMG_OK;
  };
  for (const auto &c : cand)
  {
    float ms = 0.f;
    bool  ran = false;
    HMG_TRY(probe_run(HMG_BATCH_STREAM, c.first, c.second, ms, ran));
    if (ran && ms < best)
    {
      best = ms;
      best_rows = c.first;
      best_pf = c.second;
    }
  }
  // the form, with the chosen shape: the captured launch sequence against
  // the persistent launch
  int   best_form = HMG_BATCH_GRAPH;
  float ms_graph = 0.f, ms_persist = 0.f;
  bool  ran_graph = false, ran_persist = false;
  HMG_TRY(probe_run(HMG_BATCH_GRAPH, best_rows, best_pf, ms_graph, ran_graph));
  {
    // a refused cooperative launch leaves the maps as they were: keep the
    // graph form
    const int rc = probe_run(HMG_BATCH_PERSISTENT, best_rows, best_pf, ms_persist, ran_persist);
    if (rc == HMG_ERR_UNSUPPORTED)
    {
      clear_error();
      ran_persist = false;
    }
    else
      HMG_TRY(rc);
  }
  if (ran_graph && ran_persist && ms_persist < ms_graph) best_form = HMG_BATCH_PERSISTENT;
  b->chunk_rows = best_rows;
  b->prefetch = best_pf;
  b->form = best_form;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[key] = BatchTune{best_rows, best_pf, best_form};
  }
  if (chunk_rows_out) *chunk_rows_out = best_rows;
  if (prefetch_out) *prefetch_out = best_pf;
  return HMG_OK;
}

int hmg_batch_tuned_form(hmg_batch *b, int *form)
{
  clear_error();
  if (!b || !form) return set_error(HMG_ERR_INVALID, "null argument");
  *form = b->form;
  return HMG_OK;
}

int hmg_batch_run(hmg_batch *b, int use_graph, float *device_ms)
{
  clear_error();
  if (!b) return set_error(HMG_ERR_INVALID, "null batch");
  std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
  HMG_TRY(ctx_enter(b->ctx));
  float     ms = 0.f;
  const int form = use_graph == HMG_BATCH_AUTO ? b->form
                   : use_graph == HMG_BATCH_PERSISTENT ? HMG_BATCH_PERSISTENT
                   : use_graph ? HMG_BATCH_GRAPH
                               : HMG_BATCH_STREAM;
  HMG_TRY(batch_advance(b->ctx, b->maps.data(), b->iters.data(), (int)b->maps.size(), form,
                        b->chunk_rows, b->prefetch, &ms));
  for (int &it : b->iters) it = 0;
  b->last_ms = ms;
  if (device_ms) *device_ms = ms;
  return HMG_OK;
}

int hmg_batch_export(hmg_batch *b, int map_id, float *z, float *d, float *s, float *u, float *v)
{
  clear_error();
  if (!b || map_id < 0 || map_id >= (int)b->maps.size())
    return set_error(HMG_ERR_INVALID, "bad batch or map id");
  return hmg_session_export(b->maps[map_id], z, d, s, u, v);
}

int hmg_batch_destroy(hmg_batch *b)
{
  if (!b) return HMG_OK;
  for (hmg_session *s : b->maps) hmg_session_destroy(s);
  delete b;
  return HMG_OK;
}

} // extern "C"

// -------------------------------------------------------- flow_simulation
//
// hmap::gpu::flow_simulation (HighMap/src/hydrology/flow_simulation.cpp:26-110):
// d = water_height * depth_map, fluxes 0; per iteration the flow pass and the
// water pass (d <- d2; the velocities are not kept); then, when
// dry_out_ratio != 0, water_depth_dry_out (water_depth.cpp:40-64) with
// depth_max = max(d).  One fused launch per iteration (the streaming item
// without transport and erosion): z is read, d and the fluxes ping-pong.

namespace hmg
{

template <bool PF>
__global__ void __launch_bounds__(64 * WPB) k_flow_stream(const float *__restrict__ z,
                                                     const float *__restrict__ D,
                                                     const Flux4<StF32> *__restrict__ F,
                                                     float *__restrict__ Do,
                                                     Flux4<StF32> *__restrict__ Fo,
                                                     int     strips,
                                                     int     chunk_rows,
                                                     int     n_items,
                                                     VpConst k)
{
  const int item = xcd_wave_item();
  if (item >= n_items) return;
  const VpDyn y{0, 0, 0.0, 1.f}; // not read by the flow-only item
  vp_stream_item<StF32, true, PF, true, true>(z, D, F, nullptr, nullptr, nullptr, Do, Fo, nullptr,
                                              nullptr, nullptr, nullptr, item, strips,
                                              chunk_rows, k, y);
}

__global__ void __launch_bounds__(256) k_flow_init(const float *__restrict__ depth_map,
                                                   float                wh,
                                                   float *__restrict__ d,
                                                   Flux4<StF32> *__restrict__ F,
                                                   size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride)
  {
    d[c] = wh * depth_map[c]; // Array d = water_height * depth_map
    F[c] = Flux4<StF32>{0.f, 0.f, 0.f, 0.f};
  }
}

// water_depth_dry_out without a mask: d = max(0, d - ratio * depth_max),
// depth_max from the per-block maxima of launch_minmax.
__global__ void __launch_bounds__(256) k_dry_out(const float *__restrict__ d,
                                                 float *__restrict__ out,
                                                 size_t n,
                                                 const float *__restrict__ mm,
                                                 int   nblk,
                                                 float ratio)
{
  __shared__ float s_sub;
  if (threadIdx.x == 0)
  {
    float hi = mm[3];
    for (int b = 1; b < nblk; b++) hi = fmaxf(hi, mm[3 + 2 * b]);
    s_sub = ratio * hi;
  }
  __syncthreads();
  const float sub = s_sub;
  size_t      stride = (size_t)gridDim.x * blockDim.x;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride)
    out[c] = fmaxf(0.f, d[c] - sub);
}

} // namespace hmg

extern "C" {

int hmg_flow_simulation_dev(hmg_ctx     *ctx,
                            const float *z_dev,
                            int          nx,
                            int          ny,
                            float        water_height,
                            const float *depth_map_dev,
                            int          iterations,
                            float        dt,
                            int          flux_diffusion,
                            float        flux_diffusion_strength,
                            float        dry_out_ratio,
                            float       *d_out_dev)
{
  clear_error();
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_dev || !depth_map_dev || !d_out_dev) return set_error(HMG_ERR_INVALID, "null array");
  if (nx < 1 || ny < 1) return set_error(HMG_ERR_INVALID, "shape must be positive");
  if ((size_t)nx * (size_t)ny > STREAM_MAX_CELLS)
    return set_error(HMG_ERR_INVALID, "flow_simulation is limited to 2^28 cells");
  if (iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  if (!(dt > 0.f)) return set_error(HMG_ERR_INVALID, "dt must be > 0");
  hmg_vpipes_params prm;
  hmg_vpipes_defaults(&prm);
  prm.water_height = water_height;
  prm.dt = dt;
  prm.flux_diffusion = flux_diffusion ? 1 : 0;
  prm.flux_diffusion_strength = flux_diffusion_strength;
  prm.maintain_water_volume = 0;
  prm.evap_rate = 0.f;
  const VpConst k = make_vp_const(prm, nx, ny);

  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  hipStream_t  st = ctx->stream;
  const size_t n = (size_t)nx * ny;
  DevBuf       D[2], F[2];
  for (int h = 0; h < 2; h++)
  {
    HMG_TRY(D[h].get(ctx->pool, n * sizeof(float)));
    HMG_TRY(F[h].get(ctx->pool, n * sizeof(Flux4<StF32>)));
  }
  const int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_flow_init<<<nb, 256, 0, st>>>(depth_map_dev, water_height, D[0].as<float>(),
                                  (Flux4<StF32> *)F[0].ptr, n);
  HMG_TRY(check_launch("flow_init"));

  // launch shape: the chunk length whose work items just fit the wavefront
  // slots the kernel's occupancy gives (one full round of wavefronts), 8 to
  // 64 rows per wavefront; about 16 wavefronts per CU if the runtime cannot
  // tell the occupancy
  const int  strips = (nx + STRIP_OUT - 1) / STRIP_OUT;
  const int  cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
  int        per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_flow_stream<true>,
                                                    64 * WPB, 0) != hipSuccess)
  {
    (void)hipGetLastError();
    per_cu = 0;
  }
  const long slots = per_cu > 0 ? (long)per_cu * WPB * cus : 16L * cus;
  const long chunks = std::max(1L, slots / strips);
  long       ch = (ny + chunks - 1) / chunks;
  ch = std::max(8L, std::min(64L, ch));
  const int n_items = strips * (int)((ny + ch - 1) / ch);
  const int G = xcd_grid(n_items);
  int       a = 0;
  for (int it = 0; it < iterations; it++, a ^= 1)
  {
    k_flow_stream<true><<<G, 64 * WPB, 0, st>>>(z_dev, D[a].as<float>(),
                                           (const Flux4<StF32> *)F[a].ptr, D[a ^ 1].as<float>(),
                                           (Flux4<StF32> *)F[a ^ 1].ptr, strips, (int)ch,
                                           n_items, k);
    HMG_TRY(check_launch("flow_stream"));
  }
  if (dry_out_ratio != 0.f)
  {
    int nblk = (int)std::min<size_t>((n + 255) / 256, 1024);
    DevBuf mm;
    HMG_TRY(mm.get(ctx->pool, sizeof(float) * (2 + 2 * (size_t)nblk)));
    HMG_TRY(launch_minmax(D[a].as<float>(), n, mm.as<float>(), nblk, st));
    k_dry_out<<<nblk, 256, 0, st>>>(D[a].as<float>(), d_out_dev, n, mm.as<float>(), nblk,
                                    dry_out_ratio);
    HMG_TRY(check_launch("dry_out"));
    // the pooled buffers are reused by later calls: finish before they return
    HMG_HIP(hipStreamSynchronize(st));
    return HMG_OK;
  }
  HMG_HIP(hipMemcpyAsync(d_out_dev, D[a].ptr, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  HMG_HIP(hipStreamSynchronize(st));
  return HMG_OK;
}

int hmg_flow_simulation(hmg_ctx     *ctx,
                        const float *z_host,
                        int          nx,
                        int          ny,
                        float        water_height,
                        const float *depth_map_host,
                        int          iterations,
                        float        dt,
                        int          flux_diffusion,
                        float        flux_diffusion_strength,
                        float        dry_out_ratio,
                        float       *d_out_host)
{
  clear_error();
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!z_host || !depth_map_host || !d_out_host) return set_error(HMG_ERR_INVALID, "null array");
  if (nx < 1 || ny < 1) return set_error(HMG_ERR_INVALID, "shape must be positive");
  if ((size_t)nx * (size_t)ny > STREAM_MAX_CELLS)
    return set_error(HMG_ERR_INVALID, "flow_simulation is limited to 2^28 cells");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  hipStream_t  st = ctx->stream;
  const size_t bytes = (size_t)nx * ny * sizeof(float);
  DevBuf       z, dm, d;
  HMG_TRY(z.get(ctx->pool, bytes));
  HMG_TRY(dm.get(ctx->pool, bytes));
  HMG_TRY(d.get(ctx->pool, bytes));
  HMG_HIP(hipMemcpyAsync(z.ptr, z_host, bytes, hipMemcpyHostToDevice, st));
  HMG_HIP(hipMemcpyAsync(dm.ptr, depth_map_host, bytes, hipMemcpyHostToDevice, st));
  HMG_TRY(hmg_flow_simulation_dev(ctx, z.as<float>(), nx, ny, water_height, dm.as<float>(),
                                  iterations, dt, flux_diffusion, flux_diffusion_strength,
                                  dry_out_ratio, d.as<float>()));
  HMG_HIP(hipMemcpyAsync(d_out_host, d.ptr, bytes, hipMemcpyDeviceToHost, st));
  HMG_HIP(hipStreamSynchronize(st));
  return HMG_OK;
}

} // extern "C"
