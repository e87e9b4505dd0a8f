// arith_check -- checks the short arithmetic forms of the kernels over every
// float pattern they can receive, on the GPU, and prints the counts of
// mismatches:
//   sqrt_big   equals sqrtf bit for bit for every x in [2^-96, +inf], and
//              stays in [0, 2^-47] for every x in [0, 2^-96);
//   sqrt_sumsq equals sqrtf bit for bit for every x in [0, +inf];
//   fix_term_unit (float form) equals fix_term_unit_int for every x in [0, 1).
// Compiled as HIP by the build.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "hmg_cell.hpp"

struct Counts
{
  unsigned long long sqrt_big, sqrt_small, sqrt_sumsq, fix;
  unsigned int       first_sqrt, first_small, first_sumsq, first_fix;
};

__device__ void note(unsigned long long *n, unsigned int *first, unsigned int v)
{
  if (atomicAdd(n, 1ull) == 0) *first = v;
}

__global__ void k_check(Counts *c)
{
  const unsigned int stride = gridDim.x * blockDim.x;
  for (unsigned int v = blockIdx.x * blockDim.x + threadIdx.x; v <= 0x7f800000u; v += stride)
  {
    const float x = __uint_as_float(v);
    const float ref = sqrtf(x);
    const float big = hmg::sqrt_big(x);
    if (v >= 0x0f800000u)
    {
      if (__float_as_uint(big) != __float_as_uint(ref)) note(&c->sqrt_big, &c->first_sqrt, v);
    }
    else if (!(big >= 0.f && big <= 0x1p-47f))
      note(&c->sqrt_small, &c->first_small, v);
    if (__float_as_uint(hmg::sqrt_sumsq(x)) != __float_as_uint(ref))
      note(&c->sqrt_sumsq, &c->first_sumsq, v);
    if (v < 0x3f800000u && hmg::fix_term_unit(x) != hmg::fix_term_unit_int(x))
      note(&c->fix, &c->first_fix, v);
    if (stride > 0x7f800000u - v) break; // v + stride would pass the end or wrap
  }
}

int main()
{
  Counts *d = nullptr;
  if (hipMalloc(&d, sizeof(Counts)) != hipSuccess) return 1;
  if (hipMemset(d, 0, sizeof(Counts)) != hipSuccess) return 1;
  k_check<<<8192, 256>>>(d);
  Counts h;
  if (hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  std::printf("sqrt_big %llu (first %08x) sqrt_small %llu (first %08x) sqrt_sumsq %llu (first %08x) "
              "fix %llu (first %08x)\n",
              h.sqrt_big, h.first_sqrt, h.sqrt_small, h.first_small, h.sqrt_sumsq, h.first_sumsq, h.fix,
              h.first_fix);
  return 0;
}
