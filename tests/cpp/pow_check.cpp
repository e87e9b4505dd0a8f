// pow_check -- evaluates the kernels' pow (hmg::hm_powf), or their sin and
// cos (hmg::hm_sin, hmg::hm_sincos), on the GPU for the values in a file,
// for the bit-for-bit comparison with the oracle's copies in
// tests/test_gpu_arith_check.py.
//   pow_check <in> <out> [pow|sincos]
// <in>: int32 n, then n float32 x, then n float32 y (ignored by sincos);
// <out>: n float32 pow(x, y), or for sincos 3n float32: hm_sin(x), then the
// sine and the cosine of hm_sincos(x).
// Compiled as HIP by the build.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "hmg_cell.hpp"

__global__ void k_pow(const float *x, const float *y, float *out, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = hmg::hm_powf(x[i], y[i]);
}

__global__ void k_sincos(const float *x, float *out, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
  {
    float s, c;
    hmg::hm_sincos(x[i], s, c);
    out[i] = hmg::hm_sin(x[i]);
    out[n + i] = s;
    out[2 * n + i] = c;
  }
}

int main(int argc, char **argv)
{
  const bool sincos = argc == 4 && std::string(argv[3]) == "sincos";
  if (argc != 3 && !(argc == 4 && (sincos || std::string(argv[3]) == "pow")))
  {
    std::fprintf(stderr, "usage: pow_check <in> <out> [pow|sincos]\n");
    return 2;
  }
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int n = 0;
  if (std::fread(&n, sizeof n, 1, f) != 1 || n < 1 || n > (1 << 26)) return 2;
  const int          nout = sincos ? 3 * n : n;
  std::vector<float> x(n), y(n), out(nout);
  if (std::fread(x.data(), sizeof(float), n, f) != (size_t)n) return 2;
  if (std::fread(y.data(), sizeof(float), n, f) != (size_t)n) return 2;
  std::fclose(f);
  float *dx = nullptr, *dy = nullptr, *dout = nullptr;
  const size_t bytes = sizeof(float) * (size_t)n, obytes = sizeof(float) * (size_t)nout;
  if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
      hipMalloc(&dout, obytes) != hipSuccess)
    return 1;
  if (hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dy, y.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
    return 1;
  if (sincos)
    k_sincos<<<(n + 255) / 256, 256>>>(dx, dout, n);
  else
    k_pow<<<(n + 255) / 256, 256>>>(dx, dy, dout, n);
  if (hipGetLastError() != hipSuccess) return 1;
  if (hipMemcpy(out.data(), dout, obytes, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  f = std::fopen(argv[2], "wb");
  if (!f || std::fwrite(out.data(), sizeof(float), nout, f) != (size_t)nout) return 1;
  std::fclose(f);
  std::printf("pow_check %s %d values\n", sincos ? "sincos" : "pow", n);
  return 0;
}
