/*
 * div_const_check.c -- exhaustive check of the correctly rounded division by
 * a constant used in the thermal kernels (highmap_amd/csrc/hmg_cell.hpp,
 * div_1414): for every float x with 2^-100 <= |x| < inf,
 *     q0 = RN(x * r),  q = fma(fma(-q0, b, x), r, q0)
 * equals the IEEE quotient RN(x / b), where r = RN(1 / b).
 *
 * usage: div_const_check <b as %a> <r as %a>
 * Prints the number of mismatches over all 2^32 bit patterns; exit 1 on any.
 * Built with FMA instructions (the GPU's v_fma_f32 is the same IEEE fma).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline float bits_to_f(uint32_t u)
{
  float x;
  memcpy(&x, &u, 4);
  return x;
}

static inline uint32_t f_to_bits(float x)
{
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}

int main(int argc, char **argv)
{
  if (argc < 3) return 2;
  const float b = strtof(argv[1], NULL);
  const float r = strtof(argv[2], NULL);
  volatile float one = 1.f;
  if (f_to_bits(one / b) != f_to_bits(r))
  {
    printf("r is not RN(1/b): %a vs %a\n", r, one / b);
    return 1;
  }
  long long bad = 0, checked = 0;
#pragma omp parallel for reduction(+ : bad, checked) schedule(static)
  for (long long k = 0; k < (1LL << 32); k++)
  {
    const float x = bits_to_f((uint32_t)k);
    const float ax = fabsf(x);
    if (!(ax >= 0x1p-100f && ax < INFINITY)) continue; /* the kernel divides these */
    const float q0 = x * r;
    const float q = __builtin_fmaf(__builtin_fmaf(-q0, b, x), r, q0);
    checked++;
    if (f_to_bits(q) != f_to_bits(x / b)) bad++;
  }
  printf("b=%a r=%a checked=%lld mismatches=%lld\n", b, r, checked, bad);
  return bad != 0;
}
