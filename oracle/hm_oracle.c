/*
 * hm_oracle.c -- CPU oracle for the HighMap grid-erosion hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see hm_oracle.h).  Built with
 * -ffp-contract=off and without -ffast-math so every float operation below
 * is one IEEE-754 operation, rounded to nearest even.
 *
 * Noise, hashing and the project-written sin/cos live in this file; thermal
 * and vpipes are written once in hm_oracle_erosion.inc and instantiated for
 * float (the parity contract) and double (the error oracle).
 */
#include "hm_oracle.h"

#include <math.h>
#include <float.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Integer seed path.                                                        */
/* ------------------------------------------------------------------------ */

/* Thomas Wang's 32-bit integer hash, as used by the reference
 * (HighMap/src/gpu_opencl/kernels/_common_rand.cl:66-74).
 * Parity pinned: known-answer values cross-checked against an independent
 * numpy uint32 restatement (tests/test_oracle_noise.py). */
uint32_t orc_wang_hash(uint32_t seed)
{
  seed = (seed ^ 61u) ^ (seed >> 16);
  seed *= 9u;
  seed = seed ^ (seed >> 4);
  seed *= 0x27d4eb2du;
  seed = seed ^ (seed >> 15);
  return seed;
}

/* Marsaglia xorshift32 with shifts (13, 17, 5)
 * (HighMap/src/gpu_opencl/kernels/_common_rand.cl:76-82).
 * Parity pinned: full period 2^32-1 checked by brute force, plus numpy. */
uint32_t orc_xorshift32(uint32_t x)
{
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}

/* fseed = rand(wang_hash(seed)): one xorshift step, then u32 * 2^-32 in float
 * (HighMap/src/gpu_opencl/kernels/_common_rand.cl:84-88, noise.cl:483-484). */
float orc_seed_to_fseed(uint32_t seed)
{
  uint32_t st = orc_xorshift32(orc_wang_hash(seed));
  return (float)st * (1.f / 4294967296.f);
}

/* ------------------------------------------------------------------------ */
/* Project-written sin / cos.                                                */
/*                                                                           */
/* Method: Cody-Waite argument reduction by pi/2 with a three-part constant  */
/* (each part the float nearest the remaining tail of pi/2), every step an   */
/* explicit fused multiply-add, then the Taylor polynomials of sin (to r^9)  */
/* and cos (to r^10) on |r| <= pi/4, evaluated by Horner with fmaf.  The    */
/* reference uses the OpenCL built-in sin/cos under -cl-fast-relaxed-math,   */
/* whose results are vendor-defined; a shared, fully specified routine is    */
/* what lets the GPU noise match this restatement bit for bit.               */
/* Parity pinned: error against numpy float64 sin/cos over |x| <= 2^18       */
/* (tests/test_oracle_noise.py).                                             */
/* ------------------------------------------------------------------------ */

#define ORC_TWO_OVER_PI 0x1.45f306p-1f
#define ORC_PIO2_HI 0x1.921fb6p+0f
#define ORC_PIO2_MID -0x1.777a5cp-25f
#define ORC_PIO2_LO -0x1.ee59dap-50f

static float orc_reduce(float x, int *quadrant)
{
  float k = rintf(x * ORC_TWO_OVER_PI);
  float r = fmaf(-k, ORC_PIO2_HI, x);
  r = fmaf(-k, ORC_PIO2_MID, r);
  r = fmaf(-k, ORC_PIO2_LO, r);
  /* k mod 4 for every k: floats of magnitude 2^25 and above are multiples
   * of 4, so clamping there keeps the quadrant and keeps the conversion in
   * range (a NaN k gives 0). */
  *quadrant = ((int)fminf(fmaxf(k, -0x1p25f), 0x1p25f)) & 3;
  return r;
}

static float orc_sin_poly(float r)
{
  float r2 = r * r;
  float p = fmaf(r2, 1.f / 362880.f, -1.f / 5040.f);
  p = fmaf(r2, p, 1.f / 120.f);
  p = fmaf(r2, p, -1.f / 6.f);
  return fmaf(r * r2, p, r);
}

static float orc_cos_poly(float r)
{
  float r2 = r * r;
  float p = fmaf(r2, -1.f / 3628800.f, 1.f / 40320.f);
  p = fmaf(r2, p, -1.f / 720.f);
  p = fmaf(r2, p, 1.f / 24.f);
  p = fmaf(r2, p, -0.5f);
  return fmaf(r2, p, 1.f);
}

float orc_sinf(float x)
{
  int   q;
  float r = orc_reduce(x, &q);
  switch (q)
  {
  case 0: return orc_sin_poly(r);
  case 1: return orc_cos_poly(r);
  case 2: return -orc_sin_poly(r);
  default: return -orc_cos_poly(r);
  }
}

float orc_cosf(float x)
{
  int   q;
  float r = orc_reduce(x, &q);
  switch (q)
  {
  case 0: return orc_cos_poly(r);
  case 1: return -orc_sin_poly(r);
  case 2: return -orc_cos_poly(r);
  default: return orc_sin_poly(r);
  }
}

/* 2^k as a float for k in [-126, 127] (exact). */
static float orc_pow2i(int k)
{
  uint32_t b = (uint32_t)(k + 127) << 23;
  float    f;
  memcpy(&f, &b, sizeof f);
  return f;
}

/* exp(x) of the FP32 contract.  OpenCL's exp (used by snow_simulation.cl)
 * is not correctly rounded, so the kernels and this restatement share one
 * function: x = k ln2 + r with k = rint(x / ln2) (a product by the float
 * nearest 1/ln2) and r from two fma steps (ln2 split into its nearest float
 * and the remainder), e^r from the degree-7 Taylor polynomial in Horner
 * form with fma, then times 2^k as two exact powers of two, so that only
 * the last product rounds (into the subnormal range when the result is
 * that small).  Within about 1 ulp of exp; pinned against libm's double
 * exp over a sweep. */
float orc_expf(float x)
{
  if (x != x) return x;
  if (x > 89.f) return INFINITY;
  if (x < -104.f) return 0.f;
  const float k = rintf(x * 0x1.715476p+0f);
  float       r = fmaf(-k, 0x1.62e430p-1f, x);
  r = fmaf(-k, -0x1.05c610p-29f, r);
  float p = fmaf(r, 1.f / 5040.f, 1.f / 720.f);
  p = fmaf(r, p, 1.f / 120.f);
  p = fmaf(r, p, 1.f / 24.f);
  p = fmaf(r, p, 1.f / 6.f);
  p = fmaf(r, p, 0.5f);
  p = fmaf(r, p, 1.f);
  p = fmaf(r, p, 1.f);
  const int ki = (int)k, k1 = ki / 2, k2 = ki - k1;
  return (p * orc_pow2i(k1)) * orc_pow2i(k2);
}

/* OpenCL fract(): min(x - floor(x), 0x1.fffffep-1f). */
float orc_fractf(float x)
{
  return fminf(x - floorf(x), 0x1.fffffep-1f);
}

/* hash12f (HighMap/src/gpu_opencl/kernels/_common_rand.cl:12-17):
 * fract(sin(dot(p, (127.1, 311.7)) + fseed) * 43758.5453123).  The dot
 * product is two rounded products and one rounded sum (no contraction). */
float orc_hash12f(float px, float py, float fseed)
{
  float dp = px * 127.1f + py * 311.7f;
  return orc_fractf(orc_sinf(dp + fseed) * 43758.5453123f);
}

/* wrap_lattice (noise.cl:12-17). */
static float orc_wrap(float c, int period)
{
  if (period > 0) c = c - (float)period * floorf(c / (float)period);
  return c;
}

/* scale_period (noise.cl:21-26), the float-to-int conversion saturating
 * (NaN to 0, beyond the int range to INT_MIN / INT_MAX) so that periods
 * past 2^31 after many octaves are defined. */
static int orc_scale_period(int period, float lacunarity)
{
  if (period > 0)
  {
    const float p = (float)period * lacunarity + 0.5f;
    period = p != p ? 0 : p >= 0x1p31f ? INT_MAX : p < -0x1p31f ? INT_MIN : (int)p;
  }
  return period;
}

/* grad22f (_common_rand.cl:53-57) at a (wrapped) lattice corner, then the
 * dot product with the offset (dx, dy). */
static float orc_corner(float cx, float cy, float fseed, float dx, float dy)
{
  float angle = 6.2831853f * orc_hash12f(cx, cy, fseed);
  float gx = orc_cosf(angle);
  float gy = orc_sinf(angle);
  return gx * dx + gy * dy;
}

/* grad22f alone (for the known-answer tests). */
void orc_grad22f(float px, float py, float fseed, float *gx, float *gy)
{
  float angle = 6.2831853f * orc_hash12f(px, py, fseed);
  *gx = orc_cosf(angle);
  *gy = orc_sinf(angle);
}

static float orc_fade(float x)
{
  /* smoothstep5 (_common_math.cl:126-129): x*x*x*(x*(6x-15)+10) */
  return x * x * x * (x * (6.f * x - 15.f) + 10.f);
}

static float orc_lerp(float a, float b, float t)
{
  /* lerp (_common_math.cl:47-50) */
  return (1.f - t) * a + t * b;
}

/* base_perlin (noise.cl:28-63).
 * Parity pinned: zero at lattice points, bound 1.42857*sqrt(2)/2,
 * continuity across cell edges, tiling with a period, numpy restatement. */
float orc_base_perlin(float px, float py, float fseed, int per_x, int per_y)
{
  float ix = floorf(px), iy = floorf(py);
  float fx = orc_fractf(px), fy = orc_fractf(py);

  float x0 = orc_wrap(ix, per_x), x1 = orc_wrap(ix + 1.f, per_x);
  float y0 = orc_wrap(iy, per_y), y1 = orc_wrap(iy + 1.f, per_y);

  float n00 = orc_corner(x0, y0, fseed, fx, fy);
  float n10 = orc_corner(x1, y0, fseed, fx - 1.f, fy);
  float n01 = orc_corner(x0, y1, fseed, fx, fy - 1.f);
  float n11 = orc_corner(x1, y1, fseed, fx - 1.f, fy - 1.f);

  float ux = orc_fade(fx), uy = orc_fade(fy);
  float nx0 = orc_lerp(n00, n10, ux);
  float nx1 = orc_lerp(n01, n11, ux);
  return 1.42857f * orc_lerp(nx0, nx1, uy);
}

/* hash22f (_common_rand.cl:28-34): two hashes of the lattice point, with
 * the weight pairs (127.1, 311.7) and (269.5, 183.3). */
static void orc_hash22f(float px, float py, float fseed, float *hx, float *hy)
{
  float qx = px * 127.1f + py * 311.7f;
  float qy = px * 269.5f + py * 183.3f;
  *hx = orc_fractf(orc_sinf(qx + fseed) * 43758.5453123f);
  *hy = orc_fractf(orc_sinf(qy + fseed) * 43758.5453123f);
}

/* base_simplex2 (noise.cl:65-99): 2-D simplex noise on the skewed lattice
 * with K1 = (sqrt(3)-1)/2 and K2 = (3-sqrt(3))/6 as float literals, the
 * input scaled by 0.5, contributions (0.5 - |d|^2)^4 (g . d), scale 100.
 * No period. */
static float orc_base_simplex2(float px, float py, float fseed)
{
  const float K1 = 0.366025403f, K2 = 0.211324865f;
  float       x = 0.5f * px, y = 0.5f * py;
  float       s = x * K1 + y * K1;
  float       ix = floorf(x + s), iy = floorf(y + s);
  float       t = ix * K2 + iy * K2;
  float       ax = (x - ix) + t, ay = (y - iy) + t;
  float       m = ax < ay ? 0.f : 1.f; /* step(a.y, a.x) */
  float       ox = m, oy = 1.f - m;
  float       bx = (ax - ox) + K2, by = (ay - oy) + K2;
  float       cx = (ax - 1.f) + 2.f * K2, cy = (ay - 1.f) + 2.f * K2;
  float       n[3];
  const float vx[3] = {ax, bx, cx}, vy[3] = {ay, by, cy};
  const float lx[3] = {ix, ix + ox, ix + 1.f}, ly[3] = {iy, iy + oy, iy + 1.f};
  for (int k = 0; k < 3; k++)
  {
    float w = fmaxf(0.5f - (vx[k] * vx[k] + vy[k] * vy[k]), 0.f);
    /* gradient: grad22f at the vertex, dotted with the offset */
    n[k] = w * w * w * w * orc_corner(lx[k], ly[k], fseed, vx[k], vy[k]);
  }
  return 100.f * (n[0] + n[1] + n[2]);
}

/* base_value (noise.cl:101-119) and base_value_linear (:150-166): bilinear
 * blend of lattice hashes, weights smoothstep3 (x*x*(3 - 2x)) or linear,
 * mapped to [-1, 1]. */
static float orc_base_value(float px, float py, float fseed, int per_x, int per_y, int cubic_fade)
{
  float ix = floorf(px), iy = floorf(py);
  float fx = orc_fractf(px), fy = orc_fractf(py);
  float x0 = orc_wrap(ix, per_x), x1 = orc_wrap(ix + 1.f, per_x);
  float y0 = orc_wrap(iy, per_y), y1 = orc_wrap(iy + 1.f, per_y);
  float v00 = orc_hash12f(x0, y0, fseed), v10 = orc_hash12f(x1, y0, fseed);
  float v01 = orc_hash12f(x0, y1, fseed), v11 = orc_hash12f(x1, y1, fseed);
  float ux = fx, uy = fy;
  if (cubic_fade)
  {
    ux = fx * fx * (3.f - 2.f * fx);
    uy = fy * fy * (3.f - 2.f * fy);
  }
  float nx0 = orc_lerp(v00, v10, ux);
  float nx1 = orc_lerp(v01, v11, ux);
  return 2.f * orc_lerp(nx0, nx1, uy) - 1.f;
}

/* cubic_interp (_common_math.cl:30-37), evaluated as written:
 * a t^3 + b t^2 + c t + d with a = p3 - p2 - p0 + p1, b = p0 - p1 - a,
 * c = p2 - p0, d = p1 (left to right). */
static float orc_cubic(float p0, float p1, float p2, float p3, float t)
{
  float a = p3 - p2 - p0 + p1;
  float b = p0 - p1 - a;
  float c = p2 - p0;
  return a * t * t * t + b * t * t + c * t + p1;
}

/* base_value_cubic (noise.cl:121-148): 4x4 lattice hashes, cubic rows then
 * the column, scaled 1.43 * (v - 0.5). */
static float orc_base_value_cubic(float px, float py, float fseed, int per_x, int per_y)
{
  float ix = floorf(px), iy = floorf(py);
  float fx = orc_fractf(px), fy = orc_fractf(py);
  float row[4];
  for (int dy = -1; dy <= 2; dy++)
  {
    float v[4];
    for (int dx = -1; dx <= 2; dx++)
      v[dx + 1] = orc_hash12f(orc_wrap(ix + (float)dx, per_x), orc_wrap(iy + (float)dy, per_y),
                              fseed);
    row[dy + 1] = orc_cubic(v[0], v[1], v[2], v[3], fx);
  }
  return 1.43f * (orc_cubic(row[0], row[1], row[2], row[3], fy) - 0.5f);
}

/* base_worley (noise.cl:168-190): squared distance to the nearest feature
 * point (one per lattice cell, hash22f offset) over the 3x3 cells around,
 * scanned dx outer, dy inner; 1.66 * d^2 - 1. */
static float orc_base_worley(float px, float py, float fseed, int per_x, int per_y)
{
  float ix = floorf(px), iy = floorf(py);
  float fx = orc_fractf(px), fy = orc_fractf(py);
  float dmin = FLT_MAX;
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
    {
      float hx, hy;
      orc_hash22f(orc_wrap(ix + (float)dx, per_x), orc_wrap(iy + (float)dy, per_y), fseed, &hx,
                  &hy);
      float ex = fx - ((float)dx + hx), ey = fy - ((float)dy + hy);
      dmin = fminf(dmin, ex * ex + ey * ey);
    }
  return 1.66f * dmin - 1.f;
}

/* max_smooth(a, b, k) (_common_math.cl:52-56) with pow_int(h, 3)
 * (_common_math.cl:10-18): ((1*h)*h)*h. */
static float orc_max_smooth(float a, float b, float k)
{
  float h = fmaxf(k - fabsf(a - b), 0.f) / k;
  float h3 = 1.f;
  for (int q = 0; q < 3; q++) h3 *= h;
  return fmaxf(a, b) + h3 * k / 6.f;
}

/* The single-octave value of each hmap::NoiseType the kernels implement
 * (noise.cl:385-454): PERLIN 1, PERLIN_BILLOW 2 (2|P| - 1), PERLIN_HALF 3
 * (max_smooth(P, 0, 0.5)), SIMPLEX2 4, VALUE 6, VALUE_CUBIC 7, VALUE_LINEAR
 * 9, WORLEY 10.  Other types are 0, as the kernels write. */
float orc_base_noise(int type, float px, float py, float fseed, int per_x, int per_y)
{
  switch (type)
  {
  case 1: return orc_base_perlin(px, py, fseed, per_x, per_y);
  case 2: return 2.f * fabsf(orc_base_perlin(px, py, fseed, per_x, per_y)) - 1.f;
  case 3: return orc_max_smooth(orc_base_perlin(px, py, fseed, per_x, per_y), 0.f, 0.5f);
  case 4: return orc_base_simplex2(px, py, fseed);
  case 6: return orc_base_value(px, py, fseed, per_x, per_y, 1);
  case 7: return orc_base_value_cubic(px, py, fseed, per_x, per_y);
  case 9: return orc_base_value(px, py, fseed, per_x, per_y, 0);
  case 10: return orc_base_worley(px, py, fseed, per_x, per_y);
  default: return 0.f;
  }
}

int orc_noise_type_supported(int type)
{
  return type == 1 || type == 2 || type == 3 || type == 4 || type == 6 || type == 7 ||
         type == 9 || type == 10;
}

/* The fBm of each type (noise.cl:194-383): amplitude 0.6, feedback
 * na *= (1 - w) + w * min(v + 1, 2) * 0.5, then na *= persistence; the
 * frequency grows by lacunarity and the period by scale_period (simplex has
 * none).  Parity pinned: octaves=1 gives 0.6 * base; weight=0 gives
 * amplitudes 0.6 * pers^k (closed forms in tests/test_oracle_noise.py). */
static float orc_fbm(int   type,
                     float px,
                     float py,
                     int   octaves,
                     float weight,
                     float persistence,
                     float lacunarity,
                     float fseed,
                     int   per_x,
                     int   per_y)
{
  float n = 0.f, nf = 1.f, na = 0.6f;
  for (int k = 0; k < octaves; k++)
  {
    float v = orc_base_noise(type, px * nf, py * nf, fseed, per_x, per_y);
    n += v * na;
    na *= (1.f - weight) + weight * fminf(v + 1.f, 2.f) * 0.5f;
    na *= persistence;
    nf *= lacunarity;
    per_x = orc_scale_period(per_x, lacunarity);
    per_y = orc_scale_period(per_y, lacunarity);
  }
  return n;
}

/* g_to_xy (_common_index.cl:8-24). */
static void orc_g_to_xy(int          i,
                        int          j,
                        int          nx,
                        int          ny,
                        float        kx,
                        float        ky,
                        float        dx,
                        float        dy,
                        const float *bbox4,
                        float       *x,
                        float       *y)
{
  float u = (float)i / (float)nx;
  float w = (float)j / (float)ny;
  *x = kx * (u * (bbox4[1] - bbox4[0]) + bbox4[0]) + kx * dx;
  *y = ky * (w * (bbox4[3] - bbox4[2]) + bbox4[2]) + ky * dy;
}

/* noise_fbm kernel (noise.cl:456-577) with the host defaults of
 * HighMap/include/highmap/primitives/coherent_noise.hpp:821-834.  Types the
 * kernel does not implement give 0 everywhere (returns 1 then). */
int orc_noise_fbm_f32(float       *out,
                      int          nx,
                      int          ny,
                      int          noise_type,
                      float        kx,
                      float        ky,
                      uint32_t     seed,
                      int          octaves,
                      float        weight,
                      float        persistence,
                      float        lacunarity,
                      const float *ctrl,
                      const float *noise_x,
                      const float *noise_y,
                      const float *bbox4,
                      const int   *period2)
{
  const float bbox_def[4] = {0.f, 1.f, 0.f, 1.f};
  const int   per_def[2] = {0, 0};
  if (!bbox4) bbox4 = bbox_def;
  if (!period2) period2 = per_def;

  const float fseed = orc_seed_to_fseed(seed);
  /* simplex ignores the period (noise.cl:266-285) */
  const int px = noise_type == 4 ? 0 : period2[0];
  const int py = noise_type == 4 ? 0 : period2[1];

#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
    {
      size_t idx = (size_t)j * nx + i;
      float  ct = ctrl ? ctrl[idx] : 1.f;
      float  dx = noise_x ? noise_x[idx] : 0.f;
      float  dy = noise_y ? noise_y[idx] : 0.f;
      float  w = (1.f - ct) + ct * weight;
      float  x, y;
      orc_g_to_xy(i, j, nx, ny, kx, ky, dx, dy, bbox4, &x, &y);
      out[idx] = orc_fbm(noise_type, x, y, octaves, w, persistence, lacunarity, fseed, px, py);
    }
  return orc_noise_type_supported(noise_type) ? 0 : 1;
}

/* noise kernel (noise.cl:385-454), the single-octave hmap::gpu::noise
 * (HighMap/src/primitives/coherent_noise/primitives_gpu.cpp:504-541). */
int orc_noise_f32(float       *out,
                  int          nx,
                  int          ny,
                  int          noise_type,
                  float        kx,
                  float        ky,
                  uint32_t     seed,
                  const float *noise_x,
                  const float *noise_y,
                  const float *bbox4,
                  const int   *period2)
{
  const float bbox_def[4] = {0.f, 1.f, 0.f, 1.f};
  const int   per_def[2] = {0, 0};
  if (!bbox4) bbox4 = bbox_def;
  if (!period2) period2 = per_def;
  const float fseed = orc_seed_to_fseed(seed);

#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
    {
      size_t idx = (size_t)j * nx + i;
      float  x, y;
      orc_g_to_xy(i, j, nx, ny, kx, ky, noise_x ? noise_x[idx] : 0.f,
                  noise_y ? noise_y[idx] : 0.f, bbox4, &x, &y);
      out[idx] = orc_base_noise(noise_type, x, y, fseed, period2[0], period2[1]);
    }
  return orc_noise_type_supported(noise_type) ? 0 : 1;
}

/* remap to [vmin, vmax] (HighMap/src/range/range.cpp:342-361). */
void orc_remap_f32(float *a, size_t n, float vmin, float vmax)
{
  if (n == 0) return;
  float mn = a[0], mx = a[0];
  for (size_t k = 1; k < n; k++)
  {
    if (a[k] < mn) mn = a[k];
    if (a[k] > mx) mx = a[k];
  }
  if (mn == mx)
  {
    for (size_t k = 0; k < n; k++) a[k] = vmin;
    return;
  }
  const float scale = (vmax - vmin) / (mx - mn);
  const float offset = vmin - mn * scale;
  for (size_t k = 0; k < n; k++) a[k] = a[k] * scale + offset;
}

/* ------------------------------------------------------------------------ */
/* Order-independent water-volume sum.                                       */
/*                                                                           */
/* Each float term x is mapped to the integer round(x * 2^64) (exact for    */
/* |x| >= 2^-41, round-to-nearest-even below) and the terms are summed as    */
/* 128-bit two's-complement integers, so the result does not depend on the   */
/* summation order, the tile shape or the thread count.  The 128-bit sum is  */
/* turned into a double by explicit 32-bit pieces so that the device code    */
/* can repeat the same roundings.  The reference sums in FP32 with           */
/* std::accumulate (HighMap/src/array/methods.cpp:428-431).                  */
/* Parity pinned: exact against Python integer arithmetic on random and      */
/* edge-case inputs (tests/test_oracle_vpipes.py).                           */
/* ------------------------------------------------------------------------ */

typedef __int128          i128;
typedef unsigned __int128 u128;

static i128 orc_fix_term(float x)
{
  uint32_t b;
  memcpy(&b, &x, 4);
  uint32_t e = (b >> 23) & 0xffu;
  uint32_t m = b & 0x7fffffu;
  if (e == 0) return 0; /* zero or subnormal: |x| * 2^64 < 2^-62 */
  m |= 0x800000u;
  int  sh = (int)e - 86; /* x = m * 2^(e-150); x * 2^64 = m * 2^(e-86) */
  u128 mag;
  if (sh >= 0)
    mag = (u128)m << (sh > 103 ? 103 : sh);
  else if (sh > -25)
  {
    uint32_t r = (uint32_t)(-sh);
    uint32_t q = m >> r;
    uint32_t rem = m & ((1u << r) - 1u);
    uint32_t half = 1u << (r - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    mag = q;
  }
  else
    mag = 0;
  return (b >> 31) ? -(i128)mag : (i128)mag;
}

static double orc_u32d(uint32_t x)
{
  return (double)x;
}

static double orc_fix_to_double(i128 s)
{
  u128     us = (u128)s;
  uint64_t lo = (uint64_t)us;
  int64_t  hi = (int64_t)(uint64_t)(us >> 64);
  /* hi as double: (hi >> 32) * 2^32 + (hi & 0xffffffff), one rounding */
  double hid = (double)(int32_t)(hi >> 32) * 4294967296.0 +
               orc_u32d((uint32_t)(uint64_t)hi);
  double lod = orc_u32d((uint32_t)(lo >> 32)) * 4294967296.0 +
               orc_u32d((uint32_t)lo);
  return hid + lod * 0x1p-64;
}

double orc_fixsum_f32(const float *a, size_t n)
{
  i128 s = 0;
  for (size_t k = 0; k < n; k++) s += orc_fix_term(a[k]);
  return orc_fix_to_double(s);
}

/* The 128-bit sum itself, as two's-complement limbs (test helper). */
void orc_fixsum_parts(const float *a, size_t n, uint64_t *lo, int64_t *hi)
{
  i128 s = 0;
  for (size_t k = 0; k < n; k++) s += orc_fix_term(a[k]);
  *lo = (uint64_t)(u128)s;
  *hi = (int64_t)(uint64_t)((u128)s >> 64);
}

/* Number of xorshift32 steps from `start` back to `start` (test helper:
 * 2^32 - 1 for a full-period shift triple). */
uint64_t orc_xorshift_period(uint32_t start)
{
  uint32_t x = start;
  uint64_t n = 0;
  do
  {
    x = orc_xorshift32(x);
    n++;
  } while (x != start && n < ((uint64_t)1 << 33));
  return n;
}

/* ------------------------------------------------------------------------ */
/* Storage rounding (round to nearest even), written from the IEEE-754       */
/* binary16 and bfloat16 formats.  Parity pinned: exhaustive comparison with */
/* numpy's float16 conversion and a bit-level bfloat16 restatement over all  */
/* 2^32 float32 patterns in tests/test_oracle_vpipes.py (strided sample).    */
/* ------------------------------------------------------------------------ */

float orc_round_bf16(float x)
{
  uint32_t b;
  memcpy(&b, &x, 4);
  if ((b & 0x7f800000u) == 0x7f800000u) return x; /* inf, nan */
  b += 0x7fffu + ((b >> 16) & 1u);
  b &= 0xffff0000u;
  float r;
  memcpy(&r, &b, 4);
  return r;
}

float orc_round_f16(float x)
{
  uint32_t b;
  memcpy(&b, &x, 4);
  uint32_t sign = b & 0x80000000u, a = b & 0x7fffffffu;
  float    r;
  if (a >= 0x7f800000u) return x;             /* inf, nan */
  if (a >= 0x477ff000u)                       /* >= 65520: rounds to inf */
    return sign ? -INFINITY : INFINITY;
  if (a < 0x38800000u)                        /* below 2^-14: subnormal half */
  {
    float ax;
    memcpy(&ax, &a, 4);
    r = rintf(ax * 0x1p24f) * 0x1p-24f;       /* quantum 2^-24, exact scaling */
  }
  else
  {
    a += 0x0fffu + ((a >> 13) & 1u);          /* keep 10 stored mantissa bits */
    a &= ~0x1fffu;
    memcpy(&r, &a, 4);
  }
  return sign ? -r : r;
}

static float orc_store_f32(float x, int storage)
{
  return storage == 1 ? orc_round_bf16(x) : storage == 2 ? orc_round_f16(x) : x;
}

static double orc_store_f64(double x, int storage)
{
  (void)storage; /* the FP64 oracle keeps every plane in FP64 */
  return x;
}

/* ------------------------------------------------------------------------ */
/* Thermal and vpipes, instantiated for float and double.                    */
/* ------------------------------------------------------------------------ */

static float orc_hypotf(float x, float y)
{
  /* hypot convention: sqrt(x*x + y*y), three IEEE operations. */
  return sqrtf(x * x + y * y);
}

static double orc_hypotd(double x, double y)
{
  return sqrt(x * x + y * y);
}

/* p * 2^k for k in [-151, 128]: two exact powers of two, so that only the
 * last product rounds (into the subnormal range or to infinity). */
static float orc_scale2(float p, int k)
{
  const int k1 = k / 2, k2 = k - k1;
  return (p * orc_pow2i(k1)) * orc_pow2i(k2);
}

/* The contract's pow (k_discharge_exp != 1 in the erosion pass,
 * hydraulic_vpipes.cl:167).  The reference calls OpenCL pow under
 * -cl-fast-relaxed-math, whose result is the vendor's; the kernels and this
 * restatement share one FP32-only function instead:
 *   - the special cases of C99 pow (F.9.4.4) first;
 *   - |x| = m 2^e with m in [sqrt(1/2), sqrt(2)) (subnormals scaled by 2^24
 *     first);
 *   - ln m = 2 atanh(t), t = (m - 1) / (m + 1): m - 1 is exact, m + 1 is
 *     held as two floats (Fast2Sum), t as two floats (the division's exact
 *     fma residual corrected for the low part), and 2 atanh(t) = 2t +
 *     t^3 P(t^2) with P the series 2/3 + 2/5 s + ... + 2/15 s^6 (|t| <
 *     0.172, the next term below 2^-41), summed into two floats;
 *   - log2 |x| = e + ln m / ln 2 and w = y log2 |x| as two floats (products
 *     split exactly by fma, the integer part added by TwoSum);
 *   - 2^w = 2^k 2^f with k = rint(w), f = (w - k) + w_lo, 2^f by the
 *     degree-8 Taylor polynomial of exp(f ln 2) in Horner form with fma,
 *     times 2^k as two exact powers of two.
 * Every step is an IEEE float operation (explicit fmaf, correctly rounded
 * division), so the kernels' copy gives the same bits.  Within about 1 ulp
 * of pow; pinned against Python's double pow over a sweep. */
float orc_powf(float x, float y)
{
  if (y == 0.f) return 1.f;
  if (x == 1.f) return 1.f;
  if (x != x || y != y) return x + y;
  const float ay = fabsf(y);
  const int   yint = ay >= 0x1p24f || rintf(y) == y;
  const int   yodd = yint && ay < 0x1p24f && (((int)y) & 1);
  const int   xneg = signbit(x) != 0;
  float       ax = fabsf(x);
  if (xneg && !yint && ax != 0.f && ax != INFINITY) return NAN;
  if (ay == INFINITY) return ax == 1.f ? 1.f : ((ax < 1.f) == (y < 0.f) ? INFINITY : 0.f);
  const float sgn = (xneg && yodd) ? -1.f : 1.f;
  if (ax == 0.f) return y > 0.f ? sgn * 0.f : sgn * INFINITY;
  if (ax == INFINITY) return y > 0.f ? sgn * INFINITY : sgn * 0.f;

  int e = 0;
  if (ax < 0x1p-126f)
  {
    ax *= 0x1p24f;
    e = -24;
  }
  uint32_t b;
  memcpy(&b, &ax, sizeof b);
  e += (int)(b >> 23) - 127;
  b = (b & 0x7fffffu) | 0x3f800000u;
  float m;
  memcpy(&m, &b, sizeof m);
  if (m >= 0x1.6a09e6p+0f)
  {
    m *= 0.5f;
    e += 1;
  }
  /* ln m = 2 atanh(t), t = (m - 1) / (m + 1) */
  const float u = m - 1.f;
  const float vh = m + 1.f;
  const float vl = m - (vh - 1.f);
  const float th = u / vh;
  const float rho = fmaf(-th, vh, u) - th * vl;
  const float tl = rho / vh;
  const float s = th * th;
  float       p = fmaf(s, 0x1.111112p-3f, 0x1.3b13b2p-3f);
  p = fmaf(s, p, 0x1.745d18p-3f);
  p = fmaf(s, p, 0x1.c71c72p-3f);
  p = fmaf(s, p, 0x1.24924ap-2f);
  p = fmaf(s, p, 0x1.99999ap-2f);
  p = fmaf(s, p, 0x1.555556p-1f);
  const float tail = fmaf(th * s, p, 2.f * tl);
  const float lh = 2.f * th + tail;
  const float ll = tail - (lh - 2.f * th);
  /* log2 |x| = e + (lh + ll) log2(e) */
  const float ph = lh * 0x1.715476p+0f;
  const float pl = fmaf(lh, 0x1.715476p+0f, -ph) + (lh * 0x1.4ae0c0p-26f + ll * 0x1.715476p+0f);
  const float fe = (float)e;
  const float gh = fe + ph;
  const float bv = gh - fe;
  const float gl = ((fe - (gh - bv)) + (ph - bv)) + pl;
  /* w = y log2 |x| */
  const float wh = y * gh;
  if (wh >= 128.5f) return sgn * INFINITY;
  if (wh <= -151.f) return sgn * 0.f;
  const float wl = fmaf(y, gh, -wh) + y * gl;
  const float k = rintf(wh);
  const float f = (wh - k) + wl;
  float       q = fmaf(f, 0x1.62c022p-20f, 0x1.ffcbfcp-17f);
  q = fmaf(f, q, 0x1.430912p-13f);
  q = fmaf(f, q, 0x1.5d87fep-10f);
  q = fmaf(f, q, 0x1.3b2ab6p-7f);
  q = fmaf(f, q, 0x1.c6b08ep-5f);
  q = fmaf(f, q, 0x1.ebfbe0p-3f);
  q = fmaf(f, q, 0x1.62e430p-1f);
  q = fmaf(f, q, 1.f);
  return sgn * orc_scale2(q, (int)k);
}

/* orc_powf over arrays (the tests' sweeps) */
void orc_powf_many(const float *x, const float *y, float *out, long n)
{
  for (long i = 0; i < n; i++) out[i] = orc_powf(x[i], y[i]);
}

void orc_sincos_many(const float *x, float *s, float *c, long n)
{
  for (long i = 0; i < n; i++)
  {
    s[i] = orc_sinf(x[i]);
    c[i] = orc_cosf(x[i]);
  }
}

static float orc_powf_disch(float x, float e)
{
  return e == 1.f ? x : orc_powf(x, e);
}

static double orc_powd(double x, double e)
{
  return e == 1.0 ? x : pow(x, e);
}

static double orc_sum_f64(const double *a, int nx, int ny)
{
  /* per-row partial sums, then rows in order: independent of thread count */
  double *rows = (double *)malloc(sizeof(double) * (size_t)ny);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
  {
    double s = 0.0;
    for (int i = 0; i < nx; i++) s += a[(size_t)j * nx + i];
    rows[j] = s;
  }
  double s = 0.0;
  for (int j = 0; j < ny; j++) s += rows[j];
  free(rows);
  return s;
}

static double orc_sum_f32fix(const float *a, int nx, int ny)
{
#ifdef ORC_BENCH_FLOAT_ACCUMULATE
  /* the CPU-baseline build only: the reference's own volume sum, a serial
   * FP32 std::accumulate over the water plane (methods.cpp:428-431) */
  float acc = 0.f;
  const size_t n = (size_t)nx * (size_t)ny;
  for (size_t k = 0; k < n; k++) acc += a[k];
  return (double)acc;
#else
  return orc_fixsum_f32(a, (size_t)nx * (size_t)ny);
#endif
}

#define REAL float
#define SFX(name) name##_f32
#define RC(x) (x##f)
#define R_SQRT sqrtf
#define R_FLOOR floorf
#define R_RINT rintf
#define R_FABS fabsf
#define R_MAX fmaxf
#define R_MIN fminf
#define R_HYPOT orc_hypotf
#define R_EXP orc_expf
#define R_POW orc_powf_disch
#define R_VOLUME orc_sum_f32fix
#define R_STORE orc_store_f32
#define R_IS_F32 1
#include "hm_oracle_erosion.inc"
#undef REAL
#undef SFX
#undef RC
#undef R_SQRT
#undef R_FLOOR
#undef R_RINT
#undef R_FABS
#undef R_MAX
#undef R_MIN
#undef R_HYPOT
#undef R_EXP
#undef R_POW
#undef R_VOLUME
#undef R_STORE
#undef R_IS_F32

#define REAL double
#define SFX(name) name##_f64
#define RC(x) ((double)(x##f))
#define R_SQRT sqrt
#define R_FLOOR floor
#define R_RINT rint
#define R_FABS fabs
#define R_MAX fmax
#define R_MIN fmin
#define R_HYPOT orc_hypotd
#define R_EXP exp
#define R_POW orc_powd
#define R_VOLUME orc_sum_f64
#define R_STORE orc_store_f64
#define R_IS_F32 0
#include "hm_oracle_erosion.inc"
