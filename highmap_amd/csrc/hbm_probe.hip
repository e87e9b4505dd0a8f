// hbm_probe.hip -- HBM stream kernels that measure what this GPU's memory
// delivers in practice (libhmg_probe.so, used by bench.py's ceiling and
// headline workloads; not part of the drop-in library).
//
// Three access patterns over buffers far larger than every cache:
//   read   each lane loads U float4 per trip (U loads in flight), sums them;
//   write  each lane stores U float4 per trip;
//   copy   each lane loads U float4, then stores them to the second buffer.
// Options: nontemporal loads and stores (__builtin_nontemporal_*), the grid
// either one pass (each lane handles U float4 once) or `blocks` workgroups
// of 256 that stride over the buffer.  The best of the sweep is the
// practical ceiling a bandwidth-bound kernel here can be held against.
//
// A fourth kernel, k_vp_pattern, makes the memory operations of one fused
// vpipes iteration (k_vp_stream_ga) with its work decomposition and almost
// no arithmetic: a wavefront per strip of 64 lanes (60 columns written, two
// halo lanes each side) and chunk of rows, four wavefronts of adjacent
// strips per workgroup stepping through the rows together (s_barrier),
// items spread over the XCDs in contiguous runs; per row the nontemporal
// loads of z, the water depth and the flux quadruple two rows ahead, four
// cached loads of the sediment (the bilinear texels at zero velocity, the
// cheapest backtrace), and the stores of z, depth and fluxes (nontemporal)
// and sediment (cached).  Its time is what this access pattern costs on
// the device, the ceiling the fused kernel can be held against.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace
{

typedef float v4f __attribute__((ext_vector_type(4)));

template <int MODE, int U, bool NT>
__global__ void __launch_bounds__(256) k_stream(const v4f *__restrict__ a,
                                                v4f *__restrict__ b,
                                                size_t n4,
                                                float *sink)
{
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t       i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float        acc = 0.f;
  for (; i + (size_t)(U - 1) * stride < n4; i += (size_t)U * stride)
  {
    v4f v[U];
    if (MODE != 1)
    {
#pragma unroll
      for (int u = 0; u < U; u++)
        v[u] = NT ? __builtin_nontemporal_load(&a[i + u * stride]) : a[i + u * stride];
    }
    if (MODE == 0)
    {
#pragma unroll
      for (int u = 0; u < U; u++) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    else
    {
#pragma unroll
      for (int u = 0; u < U; u++)
      {
        v4f w = v[u];
        if (MODE == 1) w = v4f{(float)i, 1.f, 2.f, (float)u};
        if (NT)
          __builtin_nontemporal_store(w, &b[i + u * stride]);
        else
          b[i + u * stride] = w;
      }
    }
  }
  for (; i < n4; i += stride) // the tail, one float4 at a time
  {
    if (MODE == 0)
    {
      const v4f v = a[i];
      acc += (v.x + v.y) + (v.z + v.w);
    }
    else
      b[i] = MODE == 1 ? v4f{0.f, 1.f, 2.f, 3.f} : a[i];
  }
  if (MODE == 0 && acc == 1234.5f) *sink = acc; // keeps the loads
}

template <int MODE, int U, bool NT>
hipError_t launch(const void *a, void *b, size_t bytes, int blocks, float *sink)
{
  const size_t n4 = bytes / 16;
  size_t       g = blocks > 0 ? (size_t)blocks : (n4 + 256 * (size_t)U - 1) / (256 * (size_t)U);
  if (g < 1) g = 1;
  k_stream<MODE, U, NT><<<(unsigned)g, 256>>>((const v4f *)a, (v4f *)b, n4, sink);
  return hipGetLastError();
}

template <int MODE>
hipError_t dispatch(int unroll, int nt, const void *a, void *b, size_t bytes, int blocks,
                    float *sink)
{
  switch (unroll * 2 + (nt ? 1 : 0))
  {
  case 2: return launch<MODE, 1, false>(a, b, bytes, blocks, sink);
  case 3: return launch<MODE, 1, true>(a, b, bytes, blocks, sink);
  case 4: return launch<MODE, 2, false>(a, b, bytes, blocks, sink);
  case 5: return launch<MODE, 2, true>(a, b, bytes, blocks, sink);
  case 8: return launch<MODE, 4, false>(a, b, bytes, blocks, sink);
  case 9: return launch<MODE, 4, true>(a, b, bytes, blocks, sink);
  case 16: return launch<MODE, 8, false>(a, b, bytes, blocks, sink);
  case 17: return launch<MODE, 8, true>(a, b, bytes, blocks, sink);
  default: return hipErrorInvalidValue;
  }
}


// ------------------------------------------------ fused vpipes access pattern

constexpr int PAT_WPB = 4;   // wavefronts per workgroup, as the fused kernel
constexpr int PAT_OUT = 60;  // columns written per wavefront

__device__ __forceinline__ int pat_clamp(int x, int n) { return min(max(x, 0), n - 1); }

template <class T>
__device__ __forceinline__ T pat_ld_nt(const T *base, uint32_t idx)
{
  return __builtin_nontemporal_load(base + idx);
}

template <class T>
__device__ __forceinline__ void pat_st_nt(T *base, uint32_t idx, T v)
{
  __builtin_nontemporal_store(v, base + idx);
}

// Storage element T (float, or uint16_t for 16-bit state) and the matching
// flux quadruple Q (16 or 8 bytes).
template <class T, class Q>
struct PatRow
{
  float z;
  T     d;
  Q     f;
};

// ZD: z and zo hold float2 (height, water depth) as the FP32 sessions keep
// them; D and Do are unused.
template <class T, class Q, bool ZD = false>
__global__ void __launch_bounds__(64 * PAT_WPB) k_vp_pattern(const float *__restrict__ z,
                                                            const T *__restrict__ D,
                                                            const Q *__restrict__ F,
                                                            const T *__restrict__ S,
                                                            float *__restrict__ zo,
                                                            T *__restrict__ Do,
                                                            Q *__restrict__ Fo,
                                                            T *__restrict__ So,
                                                            int nx,
                                                            int ny,
                                                            int strips,
                                                            int chunk_rows,
                                                            int n_items)
{
  const int G = gridDim.x;
  const int b = (G & 7) ? blockIdx.x : (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  const int item = __builtin_amdgcn_readfirstlane(b * PAT_WPB + (threadIdx.x >> 6));
  if (item >= n_items) return;
  const int  lane = threadIdx.x & 63;
  const int  strip = item % strips, chunk = item / strips;
  const int  gx = strip * PAT_OUT - 2 + lane;
  const int  gi = pat_clamp(gx, nx), gi1 = pat_clamp(gx + 1, nx);
  const bool writer = lane >= 2 && lane < 62 && gx < nx;
  const int  j0 = chunk * chunk_rows, j1 = min(j0 + chunk_rows, ny);
  typedef float v2f __attribute__((ext_vector_type(2)));
  auto load = [&](int r) {
    const uint32_t o = (uint32_t)(pat_clamp(r, ny) * nx + gi);
    if constexpr (ZD)
    {
      const v2f v = pat_ld_nt(reinterpret_cast<const v2f *>(z), o);
      return PatRow<T, Q>{v.x, __builtin_bit_cast(T, v.y), pat_ld_nt(F, o)};
    }
    else
      return PatRow<T, Q>{pat_ld_nt(z, o), pat_ld_nt(D, o), pat_ld_nt(F, o)};
  };
  PatRow<T, Q> R[5];
  R[0] = load(j0 - 1);
  R[1] = load(j0);
  R[2] = load(j0 + 1);
  R[3] = load(j0 + 2);
  for (int r0 = j0; r0 < j1; r0 += 5)
  {
#pragma unroll
    for (int q = 0; q < 5; q++)
    {
      const int r = r0 + q;
      if (r >= j1) break;
      __builtin_amdgcn_s_barrier();
      PatRow<T, Q> &am1 = R[q % 5], &a0 = R[(q + 1) % 5], &ap1 = R[(q + 2) % 5];
      const uint32_t o0 = (uint32_t)(pat_clamp(r, ny) * nx), o1 = (uint32_t)(pat_clamp(r + 1, ny) * nx);
      const T s00 = S[o0 + gi], s10 = S[o0 + gi1], s01 = S[o1 + gi], s11 = S[o1 + gi1];
      asm volatile("" ::: "memory");
      R[(q + 4) % 5] = load(r + 3);
      asm volatile("" : "+v"(am1.z), "+v"(a0.z), "+v"(ap1.z)::"memory");
      if (writer)
      {
        const uint32_t c = (uint32_t)(r * nx + gx);
        // the loaded values pass through with one operation each so that
        // every load is consumed; no arithmetic of the erosion step
        if constexpr (ZD)
          pat_st_nt(reinterpret_cast<v2f *>(zo), c,
                    v2f{a0.z + (am1.z - ap1.z), __builtin_bit_cast(float, a0.d)});
        else
        {
          pat_st_nt(zo, c, a0.z + (am1.z - ap1.z));
          pat_st_nt(Do, c, a0.d);
        }
        pat_st_nt(Fo, c, a0.f);
        So[c] = (T)(s00 ^ s10 ^ s01 ^ s11);
      }
    }
  }
}

} // namespace

extern "C" {

// mode 0 read, 1 write, 2 copy; unroll 1, 2, 4 or 8 float4 per lane and
// trip; nt: nontemporal; blocks 0: one pass over the buffer, else that many
// workgroups striding.  Runs `reps` launches on the null stream and returns
// the fastest one's time in *ms.  a and b hold `bytes` (a multiple of 16),
// sink one float.  Returns 0 or a hipError_t.
int hmg_probe_stream(int         mode,
                     const void *a,
                     void       *b,
                     size_t      bytes,
                     int         unroll,
                     int         nt,
                     int         blocks,
                     int         reps,
                     float      *sink,
                     float      *ms)
{
  if (mode < 0 || mode > 2 || reps < 1 || !ms || bytes % 16) return (int)hipErrorInvalidValue;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return (int)e;
  e = hipEventCreate(&e1);
  float best = -1.f;
  for (int r = 0; r < reps && e == hipSuccess; r++)
  {
    e = hipEventRecord(e0, nullptr);
    if (e != hipSuccess) break;
    e = mode == 0   ? dispatch<0>(unroll, nt, a, b, bytes, blocks, sink)
        : mode == 1 ? dispatch<1>(unroll, nt, a, b, bytes, blocks, sink)
                    : dispatch<2>(unroll, nt, a, b, bytes, blocks, sink);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess && (best < 0.f || t < best)) best = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return (int)e;
}


// The fused vpipes iteration's memory operations (k_vp_pattern) over an
// nx x ny grid with chunks of `chunk_rows` rows; `elem` is the state
// storage's element size (4: FP32, 2: BF16/FP16), or 8 for FP32 state with
// the heights and water depths in one float2 plane (the sessions' FP32
// layout).  The planes are the caller's: z and zo nx*ny floats (nx*ny
// float2 for elem 8), D, S, Do, So nx*ny elements (D, Do unused for elem 8),
// F and Fo nx*ny flux quadruples (4 elements each).  Runs two warm-up launches and
// then `reps` on the null stream (inputs and outputs swap between launches)
// and returns their average time in *ms.  Returns 0 or a hipError_t.
int hmg_probe_vp_pattern(int    elem,
                         void  *z,
                         void  *D,
                         void  *F,
                         void  *S,
                         void  *zo,
                         void  *Do,
                         void  *Fo,
                         void  *So,
                         int    nx,
                         int    ny,
                         int    chunk_rows,
                         int    reps,
                         float *ms)
{
  if ((elem != 4 && elem != 2 && elem != 8) || nx < 3 || ny < 3 || chunk_rows < 1 || reps < 1 ||
      !ms ||
      (long long)nx * ny > (1ll << 28))
    return (int)hipErrorInvalidValue;
  const int  strips = (nx + PAT_OUT - 1) / PAT_OUT;
  const int  chunks = (ny + chunk_rows - 1) / chunk_rows;
  const long items = (long)strips * chunks;
  if (items > (1l << 30)) return (int)hipErrorInvalidValue;
  // workgroups rounded up to a multiple of 8, as the fused kernel's launches
  const int G = (int)(((items + PAT_WPB - 1) / PAT_WPB + 7) & ~7l);
  auto      launch = [&](int r) {
    const bool sw = r & 1; // ping-pong: the previous launch's outputs are the inputs
    void *iz = sw ? zo : z, *iD = sw ? Do : D, *iF = sw ? Fo : F, *iS = sw ? So : S;
    void *oz = sw ? z : zo, *oD = sw ? D : Do, *oF = sw ? F : Fo, *oS = sw ? S : So;
    typedef uint32_t q4 __attribute__((ext_vector_type(4)));
    if (elem == 8)
      k_vp_pattern<uint32_t, q4, true><<<G, 64 * PAT_WPB>>>(
          (const float *)iz, nullptr, (const q4 *)iF, (const uint32_t *)iS, (float *)oz, nullptr,
          (q4 *)oF, (uint32_t *)oS, nx, ny, strips, chunk_rows, (int)items);
    else if (elem == 4)
    {
      k_vp_pattern<uint32_t, q4><<<G, 64 * PAT_WPB>>>(
          (const float *)iz, (const uint32_t *)iD, (const q4 *)iF, (const uint32_t *)iS, (float *)oz,
          (uint32_t *)oD, (q4 *)oF, (uint32_t *)oS, nx, ny, strips, chunk_rows, (int)items);
    }
    else
    {
      typedef uint32_t q2 __attribute__((ext_vector_type(2)));
      k_vp_pattern<uint16_t, q2><<<G, 64 * PAT_WPB>>>(
          (const float *)iz, (const uint16_t *)iD, (const q2 *)iF, (const uint16_t *)iS, (float *)oz,
          (uint16_t *)oD, (q2 *)oF, (uint16_t *)oS, nx, ny, strips, chunk_rows, (int)items);
    }
    return hipGetLastError();
  };
  // two warm-up launches, then `reps` back to back between two events (the
  // average launch time, as the fused kernel is measured)
  hipError_t e = launch(0);
  if (e == hipSuccess) e = launch(1);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch(r);
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  *ms = t / reps;
  return (int)e;
}

} // extern "C"
