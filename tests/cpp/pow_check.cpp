// pow_check -- evaluates the kernels' pow (hmg::hm_powf) on the GPU for the
// pairs in a file, for the bit-for-bit comparison with the oracle's copy in
// tests/test_gpu_arith_check.py.
//   pow_check <in> <out>
// <in>: int32 n, then n float32 x, then n float32 y; <out>: n float32.
// Compiled as HIP by the build.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "hmg_cell.hpp"

__global__ void k_pow(const float *x, const float *y, float *out, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = hmg::hm_powf(x[i], y[i]);
}

int main(int argc, char **argv)
{
  if (argc != 3)
  {
    std::fprintf(stderr, "usage: pow_check <in> <out>\n");
    return 2;
  }
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int n = 0;
  if (std::fread(&n, sizeof n, 1, f) != 1 || n < 1 || n > (1 << 26)) return 2;
  std::vector<float> x(n), y(n), out(n);
  if (std::fread(x.data(), sizeof(float), n, f) != (size_t)n) return 2;
  if (std::fread(y.data(), sizeof(float), n, f) != (size_t)n) return 2;
  std::fclose(f);
  float *dx = nullptr, *dy = nullptr, *dout = nullptr;
  const size_t bytes = sizeof(float) * (size_t)n;
  if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
      hipMalloc(&dout, bytes) != hipSuccess)
    return 1;
  if (hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy, y.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
    return 1;
  k_pow<<<(n + 255) / 256, 256>>>(dx, dy, dout, n);
  if (hipGetLastError() != hipSuccess) return 1;
  if (hipMemcpy(out.data(), dout, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  f = std::fopen(argv[2], "wb");
  if (!f || std::fwrite(out.data(), sizeof(float), n, f) != (size_t)n) return 1;
  std::fclose(f);
  std::printf("pow_check %d values\n", n);
  return 0;
}
