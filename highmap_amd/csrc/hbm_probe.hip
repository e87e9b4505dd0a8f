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
#include <hip/hip_runtime.h>

#include <cstddef>

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

} // extern "C"
