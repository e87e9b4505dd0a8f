// bf16_check -- compares the BF16 storage codec of the kernels
// (StBF16::st, built on the conversion instruction) with its integer form
// (StBF16::st_int) for every 32-bit pattern on the GPU, NaNs included, and
// prints the number of differences.  Compiled as HIP by the build.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "hmg_cell.hpp"

__global__ void k_cmp(unsigned long long *diff, unsigned int *first)
{
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long v = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       v < (1ull << 32); v += stride)
  {
    const float x = __uint_as_float((unsigned int)v);
    if (hmg::StBF16::st(x) != hmg::StBF16::st_int(x))
    {
      const unsigned long long k = atomicAdd(diff, 1ull);
      if (k < 4) first[k] = (unsigned int)v;
    }
  }
}

int main()
{
  unsigned long long *diff = nullptr;
  unsigned int       *first = nullptr;
  if (hipMalloc(&diff, sizeof(*diff)) != hipSuccess || hipMalloc(&first, 4 * sizeof(*first)) != hipSuccess)
    return 1;
  if (hipMemset(diff, 0, sizeof(*diff)) != hipSuccess || hipMemset(first, 0, 4 * sizeof(*first)) != hipSuccess)
    return 1;
  k_cmp<<<8192, 256>>>(diff, first);
  unsigned long long h = 0;
  unsigned int       f[4] = {0, 0, 0, 0};
  if (hipMemcpy(&h, diff, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  if (hipMemcpy(f, first, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  std::printf("differences %llu first %08x %08x %08x %08x\n", h, f[0], f[1], f[2], f[3]);
  return 0;
}
