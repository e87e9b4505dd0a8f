// noise.hip -- hmap::gpu::noise_fbm and hmap::gpu::noise for gfx950.
//
// Follows the reference kernels noise_fbm and noise
// (HighMap/src/gpu_opencl/kernels/noise.cl:385-577) with the base noises
// (:28-190), their fBm loops (:194-383), wrap_lattice/scale_period (:12-26),
// g_to_xy (_common_index.cl:8-24), the hashes of _common_rand.cl and the
// helpers of _common_math.cl; host wrappers
// HighMap/src/primitives/coherent_noise/primitives_gpu.cpp:504-591.
// Implemented types (hmap::NoiseType, functions.hpp:66-81): PERLIN 1,
// PERLIN_BILLOW 2, PERLIN_HALF 3, SIMPLEX2 4, VALUE 6, VALUE_CUBIC 7,
// VALUE_LINEAR 9, WORLEY 10; the kernels write 0 for the other types and so
// does this file.
//
// The kernels are VALU-bound (hash sines and sincos per lattice point per
// octave), not memory-bound: one 4-byte store per cell.  Without warping
// planes, each wavefront (64 cells of one row) forms every lattice value it
// needs once, one or two per lane, and the cells gather them by cross-lane
// permute (the *_wave forms).  sin/cos are the
// project-written routines of hmg_cell.hpp, so each field matches the FP32
// restatement bit for bit.  One kernel instance per type (template), so the
// per-cell work carries no type dispatch.
#include <cfloat>
#include <climits>

#include "hmg_internal.hpp"

namespace hmg
{

__device__ __forceinline__ float wrap_lattice(float c, int period)
{
  if (period > 0) c = c - (float)period * floorf(c / (float)period);
  return c;
}

// scale_period with a saturating conversion (NaN to 0, beyond the int range
// to INT_MIN / INT_MAX), so that periods past 2^31 after many octaves are
// defined.
__device__ __forceinline__ int scale_period(int period, float lacunarity)
{
  if (period > 0)
  {
    const float p = (float)period * lacunarity + 0.5f;
    period = p != p ? 0 : p >= 0x1p31f ? INT_MAX : p < -0x1p31f ? INT_MIN : (int)p;
  }
  return period;
}

// grad22f at a lattice corner: (cos a, sin a), a = 2 pi hash12f.
__device__ __forceinline__ void perlin_grad(float cx, float cy, float fseed, float &c, float &s)
{
  float angle = 6.2831853f * hash12f(cx, cy, fseed);
  hm_sincos(angle, s, c);
}

// grad22f at a lattice corner dotted with the offset (dx, dy).
__device__ __forceinline__ float perlin_corner(float cx, float cy, float fseed, float dx, float dy)
{
  float s, c;
  perlin_grad(cx, cy, fseed, c, s);
  return c * dx + s * dy;
}

__device__ __forceinline__ float fade5(float x)
{
  return x * x * x * (x * (6.f * x - 15.f) + 10.f);
}

__device__ __forceinline__ float lerp_cl(float a, float b, float t)
{
  return (1.f - t) * a + t * b;
}

__device__ __forceinline__ float base_perlin(float px, float py, float fseed, int per_x, int per_y)
{
  float ix = floorf(px), iy = floorf(py);
  float fx = fract_cl(px), fy = fract_cl(py);
  float x0 = wrap_lattice(ix, per_x), x1 = wrap_lattice(ix + 1.f, per_x);
  float y0 = wrap_lattice(iy, per_y), y1 = wrap_lattice(iy + 1.f, per_y);
  float n00 = perlin_corner(x0, y0, fseed, fx, fy);
  float n10 = perlin_corner(x1, y0, fseed, fx - 1.f, fy);
  float n01 = perlin_corner(x0, y1, fseed, fx, fy - 1.f);
  float n11 = perlin_corner(x1, y1, fseed, fx - 1.f, fy - 1.f);
  float ux = fade5(fx), uy = fade5(fy);
  float a = lerp_cl(n00, n10, ux);
  float b = lerp_cl(n01, n11, ux);
  return 1.42857f * lerp_cl(a, b, uy);
}

// The lattice columns a wavefront needs, for the wave forms below: all lanes
// share iy, and ix is monotonic across the lanes (every step from the cell
// index to ix is), so the columns are those between the first lane's ix and
// the last valid lane's (`last`, uniform).  True, with lo the smaller end,
// when they span at most `span` columns and |ix| < 2^23 at both ends (lattice
// integers within a few columns of them are then exact floats); false for a
// NaN or an infinity at either end.
__device__ __forceinline__ bool wave_lattice_span(float ix, int last, float span, float &lo)
{
  const float ix_first = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(ix)));
  const float ix_last = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ix), last));
  lo = fminf(ix_first, ix_last);
  const float hi = fmaxf(ix_first, ix_last);
  return ix_first == ix_first && ix_last == ix_last && hi - lo <= span && fabsf(lo) < 0x1p23f &&
         fabsf(hi) < 0x1p23f;
}

// base_perlin for a whole wavefront whose 64 lanes are 64 consecutive cells
// of one row, without warping planes.  When the wave's lattice columns span
// at most 31 (wave_lattice_span), lanes 0-31 form the gradients of columns
// lo .. lo+31 on lattice row y0 and lanes 32-63 on row y1, once per lattice
// point, and each lane fetches its four corners from those lanes: the same
// hash12f and sincos of the same wrapped lattice coordinates as base_perlin
// (lo + m is ix and lo + m + 1 is ix + 1 as floats), dotted and blended in
// the same order, so the result is bit for bit base_perlin's.  Otherwise
// the wave takes base_perlin.  `last` is the last valid lane (uniform);
// lanes past it compute but their results are not stored.
__device__ __forceinline__ float base_perlin_wave(float px,
                                                  float py,
                                                  float fseed,
                                                  int   per_x,
                                                  int   per_y,
                                                  int   lane,
                                                  int   last)
{
  const float ix = floorf(px), iy = floorf(py);
  float       lo;
  if (!wave_lattice_span(ix, last, 30.f, lo)) return base_perlin(px, py, fseed, per_x, per_y);
  const float fx = fract_cl(px), fy = fract_cl(py);
  const float y0 = wrap_lattice(iy, per_y), y1 = wrap_lattice(iy + 1.f, per_y);
  float       gc, gs;
  perlin_grad(wrap_lattice(lo + (float)(lane & 31), per_x), lane < 32 ? y0 : y1, fseed, gc, gs);
  const int m = (int)(ix - lo); // 0 .. 30 on valid lanes
  const int a = m & 63, b = (m + 1) & 63, c = (m + 32) & 63, d = (m + 33) & 63;
  const float c00 = __shfl(gc, a, 64), s00 = __shfl(gs, a, 64);
  const float c10 = __shfl(gc, b, 64), s10 = __shfl(gs, b, 64);
  const float c01 = __shfl(gc, c, 64), s01 = __shfl(gs, c, 64);
  const float c11 = __shfl(gc, d, 64), s11 = __shfl(gs, d, 64);
  const float n00 = c00 * fx + s00 * fy;
  const float n10 = c10 * (fx - 1.f) + s10 * fy;
  const float n01 = c01 * fx + s01 * (fy - 1.f);
  const float n11 = c11 * (fx - 1.f) + s11 * (fy - 1.f);
  const float ux = fade5(fx), uy = fade5(fy);
  const float ea = lerp_cl(n00, n10, ux);
  const float eb = lerp_cl(n01, n11, ux);
  return 1.42857f * lerp_cl(ea, eb, uy);
}

// Simplex noise on the skewed triangular lattice: the input is halved, the
// cell found with the skew (sqrt(3)-1)/2 and unskew (3-sqrt(3))/6 constants,
// and each of the three vertices adds (0.5 - |d|^2)^4 (g . d); scale 100.
__device__ __forceinline__ float base_simplex2(float px, float py, float fseed)
{
  constexpr float K1 = 0.366025403f, K2 = 0.211324865f;
  const float     x = 0.5f * px, y = 0.5f * py;
  const float     skew = x * K1 + y * K1;
  const float     ix = floorf(x + skew), iy = floorf(y + skew);
  const float     unskew = ix * K2 + iy * K2;
  const float     ax = (x - ix) + unskew, ay = (y - iy) + unskew;
  const float     m = ax < ay ? 0.f : 1.f; // second vertex: +x or +y
  const float     ox = m, oy = 1.f - m;
  const float     vx[3] = {ax, (ax - ox) + K2, (ax - 1.f) + 2.f * K2};
  const float     vy[3] = {ay, (ay - oy) + K2, (ay - 1.f) + 2.f * K2};
  const float     lx[3] = {ix, ix + ox, ix + 1.f}, ly[3] = {iy, iy + oy, iy + 1.f};
  float           n[3];
#pragma unroll
  for (int k = 0; k < 3; k++)
  {
    const float w = fmaxf(0.5f - (vx[k] * vx[k] + vy[k] * vy[k]), 0.f);
    n[k] = w * w * w * w * perlin_corner(lx[k], ly[k], fseed, vx[k], vy[k]);
  }
  return 100.f * (n[0] + n[1] + n[2]);
}

// base_simplex2 for a wavefront of one row (as base_perlin_wave): along the
// row both skewed lattice indices are monotonic (x + skew and y + skew grow
// or shrink with x), so the first and the last valid lane bound them.  When
// they span at most 15 columns and 3 rows, lane 16 r + c forms the gradient
// of lattice point (lo_x + c, lo_y + r), and each lane fetches those of its
// three vertices (ix, iy), (ix + ox, iy + oy) and (ix + 1, iy + 1); the
// same hash12f and sincos of the same lattice coordinates, dotted, weighted
// and summed in the same order.
__device__ __forceinline__ float base_simplex2_wave(float px, float py, float fseed, int lane,
                                                    int last)
{
  constexpr float K1 = 0.366025403f, K2 = 0.211324865f;
  const float     x = 0.5f * px, y = 0.5f * py;
  const float     skew = x * K1 + y * K1;
  const float     ix = floorf(x + skew), iy = floorf(y + skew);
  float           lox, loy;
  if (!wave_lattice_span(ix, last, 14.f, lox) || !wave_lattice_span(iy, last, 2.f, loy))
    return base_simplex2(px, py, fseed);
  float gc, gs;
  perlin_grad(lox + (float)(lane & 15), loy + (float)(lane >> 4), fseed, gc, gs);
  const float unskew = ix * K2 + iy * K2;
  const float ax = (x - ix) + unskew, ay = (y - iy) + unskew;
  const float m = ax < ay ? 0.f : 1.f; // second vertex: +x or +y
  const float ox = m, oy = 1.f - m;
  const float vx[3] = {ax, (ax - ox) + K2, (ax - 1.f) + 2.f * K2};
  const float vy[3] = {ay, (ay - oy) + K2, (ay - 1.f) + 2.f * K2};
  const int   b = 16 * (int)(iy - loy) + (int)(ix - lox); // 16 by + bx, bx <= 14, by <= 2
  const int   mx = ax < ay ? 0 : 1;
  const int   k3[3] = {b, b + 16 * (1 - mx) + mx, b + 17};
  float       n[3];
#pragma unroll
  for (int k = 0; k < 3; k++)
  {
    const float w = fmaxf(0.5f - (vx[k] * vx[k] + vy[k] * vy[k]), 0.f);
    const float c = __shfl(gc, k3[k] & 63, 64), sn = __shfl(gs, k3[k] & 63, 64);
    n[k] = w * w * w * w * (c * vx[k] + sn * vy[k]);
  }
  return 100.f * (n[0] + n[1] + n[2]);
}

// Value noise: bilinear blend of the four lattice hashes, with the cubic
// fade x^2 (3 - 2x) (VALUE) or linear weights (VALUE_LINEAR), to [-1, 1].
template <bool FADE>
__device__ __forceinline__ float base_value(float px, float py, float fseed, int per_x, int per_y)
{
  const float ix = floorf(px), iy = floorf(py);
  const float fx = fract_cl(px), fy = fract_cl(py);
  const float x0 = wrap_lattice(ix, per_x), x1 = wrap_lattice(ix + 1.f, per_x);
  const float y0 = wrap_lattice(iy, per_y), y1 = wrap_lattice(iy + 1.f, per_y);
  const float v00 = hash12f(x0, y0, fseed), v10 = hash12f(x1, y0, fseed);
  const float v01 = hash12f(x0, y1, fseed), v11 = hash12f(x1, y1, fseed);
  const float ux = FADE ? fx * fx * (3.f - 2.f * fx) : fx;
  const float uy = FADE ? fy * fy * (3.f - 2.f * fy) : fy;
  return 2.f * lerp_cl(lerp_cl(v00, v10, ux), lerp_cl(v01, v11, ux), uy) - 1.f;
}

// base_value for a wavefront of one row (as base_perlin_wave): lanes 0-31
// hash columns lo .. lo+31 of lattice row y0, lanes 32-63 of row y1, and
// each lane fetches its four corners; the same hashes of the same wrapped
// coordinates, blended in the same order.
template <bool FADE>
__device__ __forceinline__ float base_value_wave(float px,
                                                 float py,
                                                 float fseed,
                                                 int   per_x,
                                                 int   per_y,
                                                 int   lane,
                                                 int   last)
{
  const float ix = floorf(px), iy = floorf(py);
  float       lo;
  if (!wave_lattice_span(ix, last, 30.f, lo)) return base_value<FADE>(px, py, fseed, per_x, per_y);
  const float fx = fract_cl(px), fy = fract_cl(py);
  const float y0 = wrap_lattice(iy, per_y), y1 = wrap_lattice(iy + 1.f, per_y);
  const float h = hash12f(wrap_lattice(lo + (float)(lane & 31), per_x), lane < 32 ? y0 : y1, fseed);
  const int   m = (int)(ix - lo); // 0 .. 30 on valid lanes
  const float v00 = __shfl(h, m & 63, 64), v10 = __shfl(h, (m + 1) & 63, 64);
  const float v01 = __shfl(h, (m + 32) & 63, 64), v11 = __shfl(h, (m + 33) & 63, 64);
  const float ux = FADE ? fx * fx * (3.f - 2.f * fx) : fx;
  const float uy = FADE ? fy * fy * (3.f - 2.f * fy) : fy;
  return 2.f * lerp_cl(lerp_cl(v00, v10, ux), lerp_cl(v01, v11, ux), uy) - 1.f;
}

// Cubic through four samples at t in [0, 1] between p1 and p2, evaluated
// left to right as a t^3 + b t^2 + c t + d (cubic_interp,
// _common_math.cl:30-37).
__device__ __forceinline__ float cubic4(float p0, float p1, float p2, float p3, float t)
{
  const float a = p3 - p2 - p0 + p1;
  const float b = p0 - p1 - a;
  const float c = p2 - p0;
  return a * t * t * t + b * t * t + c * t + p1;
}

// VALUE_CUBIC: 4x4 lattice hashes, cubic along x per row, then along y.
__device__ __forceinline__ float base_value_cubic(float px, float py, float fseed, int per_x,
                                                  int per_y)
{
  const float ix = floorf(px), iy = floorf(py);
  const float fx = fract_cl(px), fy = fract_cl(py);
  float       row[4];
#pragma unroll
  for (int dy = 0; dy < 4; dy++)
  {
    const float yy = wrap_lattice(iy + (float)(dy - 1), per_y);
    float       v[4];
#pragma unroll
    for (int dx = 0; dx < 4; dx++) v[dx] = hash12f(wrap_lattice(ix + (float)(dx - 1), per_x), yy, fseed);
    row[dy] = cubic4(v[0], v[1], v[2], v[3], fx);
  }
  return 1.43f * (cubic4(row[0], row[1], row[2], row[3], fy) - 0.5f);
}

// base_value_cubic for a wavefront of one row (as base_perlin_wave): when
// the wave's lattice columns span at most 13, lane 16 r + c hashes column
// lo - 1 + c of lattice row iy - 1 + r (r 0-3, c 0-15), and each lane
// fetches its 4 x 4 block (columns ix - 1 .. ix + 2 are lanes m .. m + 3 of
// each row, m = ix - lo); the same hashes of the same wrapped coordinates,
// interpolated in the same order.
__device__ __forceinline__ float base_value_cubic_wave(float px,
                                                       float py,
                                                       float fseed,
                                                       int   per_x,
                                                       int   per_y,
                                                       int   lane,
                                                       int   last)
{
  const float ix = floorf(px), iy = floorf(py);
  float       lo;
  if (!wave_lattice_span(ix, last, 12.f, lo)) return base_value_cubic(px, py, fseed, per_x, per_y);
  const float fx = fract_cl(px), fy = fract_cl(py);
  const float h = hash12f(wrap_lattice(lo + (float)((lane & 15) - 1), per_x),
                          wrap_lattice(iy + (float)((lane >> 4) - 1), per_y), fseed);
  const int   m = (int)(ix - lo); // 0 .. 12 on valid lanes
  float       row[4];
#pragma unroll
  for (int dy = 0; dy < 4; dy++)
  {
    float v[4];
#pragma unroll
    for (int dx = 0; dx < 4; dx++) v[dx] = __shfl(h, (16 * dy + m + dx) & 63, 64);
    row[dy] = cubic4(v[0], v[1], v[2], v[3], fx);
  }
  return 1.43f * (cubic4(row[0], row[1], row[2], row[3], fy) - 0.5f);
}

// hash22f (_common_rand.cl:28-34).
__device__ __forceinline__ void hash22f(float px, float py, float fseed, float &hx, float &hy)
{
  const float qx = px * 127.1f + py * 311.7f;
  const float qy = px * 269.5f + py * 183.3f;
  hx = fract_cl(hm_sin(qx + fseed) * 43758.5453123f);
  hy = fract_cl(hm_sin(qy + fseed) * 43758.5453123f);
}

// WORLEY: squared distance to the nearest of the feature points of the 3x3
// surrounding cells (x outer, y inner, as the reference scans), 1.66 d^2 - 1.
__device__ __forceinline__ float base_worley(float px, float py, float fseed, int per_x, int per_y)
{
  const float ix = floorf(px), iy = floorf(py);
  const float fx = fract_cl(px), fy = fract_cl(py);
  float       dmin = FLT_MAX;
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
    {
      float hx, hy;
      hash22f(wrap_lattice(ix + (float)dx, per_x), wrap_lattice(iy + (float)dy, per_y), fseed, hx,
              hy);
      const float ex = fx - ((float)dx + hx), ey = fy - ((float)dy + hy);
      dmin = fminf(dmin, ex * ex + ey * ey);
    }
  return 1.66f * dmin - 1.f;
}

// base_worley for a wavefront of one row (as base_perlin_wave): when the
// wave's lattice columns span at most 19, lane 21 r + c forms the feature
// point of column lo - 1 + c in lattice row iy - 1 + r (r 0-2, c 0-20; lane
// 63 forms an unused one), and each lane fetches its 3 x 3 cells in the
// reference's scan order; the same hash22f of the same wrapped coordinates.
__device__ __forceinline__ float base_worley_wave(float px,
                                                  float py,
                                                  float fseed,
                                                  int   per_x,
                                                  int   per_y,
                                                  int   lane,
                                                  int   last)
{
  const float ix = floorf(px), iy = floorf(py);
  float       lo;
  if (!wave_lattice_span(ix, last, 18.f, lo)) return base_worley(px, py, fseed, per_x, per_y);
  const float fx = fract_cl(px), fy = fract_cl(py);
  const int   r = lane >= 42 ? 2 : lane >= 21 ? 1 : 0, c = lane - 21 * r;
  float       hx, hy;
  hash22f(wrap_lattice(lo + (float)(c - 1), per_x), wrap_lattice(iy + (float)(r - 1), per_y),
          fseed, hx, hy);
  const int m = (int)(ix - lo); // 0 .. 18 on valid lanes
  float     dmin = FLT_MAX;
#pragma unroll
  for (int dx = -1; dx <= 1; dx++)
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
    {
      const int   k = (21 * (dy + 1) + m + dx + 1) & 63;
      const float gx = __shfl(hx, k, 64), gy = __shfl(hy, k, 64);
      const float ex = fx - ((float)dx + gx), ey = fy - ((float)dy + gy);
      dmin = fminf(dmin, ex * ex + ey * ey);
    }
  return 1.66f * dmin - 1.f;
}

// max_smooth(a, b, k) (_common_math.cl:52-56), pow_int(h, 3) as ((1 h) h) h.
__device__ __forceinline__ float max_smooth(float a, float b, float k)
{
  const float h = fmaxf(k - fabsf(a - b), 0.f) / k;
  return fmaxf(a, b) + h * h * h * k / 6.f;
}

// The three Perlin types from one Perlin value.
template <int TYPE>
__device__ __forceinline__ float perlin_type(float v)
{
  if constexpr (TYPE == 2) return 2.f * fabsf(v) - 1.f;
  if constexpr (TYPE == 3) return max_smooth(v, 0.f, 0.5f);
  return v;
}

template <int TYPE>
__device__ __forceinline__ float base_noise(float px, float py, float fseed, int per_x, int per_y)
{
  if constexpr (TYPE >= 1 && TYPE <= 3) return perlin_type<TYPE>(base_perlin(px, py, fseed, per_x, per_y));
  if constexpr (TYPE == 4) return base_simplex2(px, py, fseed);
  if constexpr (TYPE == 6) return base_value<true>(px, py, fseed, per_x, per_y);
  if constexpr (TYPE == 7) return base_value_cubic(px, py, fseed, per_x, per_y);
  if constexpr (TYPE == 9) return base_value<false>(px, py, fseed, per_x, per_y);
  if constexpr (TYPE == 10) return base_worley(px, py, fseed, per_x, per_y);
  return 0.f;
}

// The wavefront forms (64 cells of one row, no warping planes).
template <int TYPE>
__device__ __forceinline__ float base_noise_wave(float px,
                                                 float py,
                                                 float fseed,
                                                 int   per_x,
                                                 int   per_y,
                                                 int   lane,
                                                 int   last)
{
  if constexpr (TYPE >= 1 && TYPE <= 3)
    return perlin_type<TYPE>(base_perlin_wave(px, py, fseed, per_x, per_y, lane, last));
  if constexpr (TYPE == 4) return base_simplex2_wave(px, py, fseed, lane, last);
  if constexpr (TYPE == 6) return base_value_wave<true>(px, py, fseed, per_x, per_y, lane, last);
  if constexpr (TYPE == 7) return base_value_cubic_wave(px, py, fseed, per_x, per_y, lane, last);
  if constexpr (TYPE == 9) return base_value_wave<false>(px, py, fseed, per_x, per_y, lane, last);
  if constexpr (TYPE == 10) return base_worley_wave(px, py, fseed, per_x, per_y, lane, last);
  return base_noise<TYPE>(px, py, fseed, per_x, per_y);
}

struct NoiseArgs
{
  int   nx, ny;
  float kx, ky;
  float fseed;
  int   octaves;
  float weight, persistence, lacunarity;
  float b0, b1, b2, b3;
  int   per_x, per_y;
};

// g_to_xy (_common_index.cl:8-24)
__device__ __forceinline__ void noise_xy(int i, int j, const NoiseArgs &a, float dx, float dy,
                                         float &x, float &y)
{
  x = (float)i / (float)a.nx;
  y = (float)j / (float)a.ny;
  x = a.kx * (x * (a.b1 - a.b0) + a.b0) + a.kx * dx;
  y = a.ky * (y * (a.b3 - a.b2) + a.b2) + a.ky * dy;
}

// WAVE: no warping planes, the lattice values shared across each
// wavefront (base_noise_wave).  Blocks of 64 x 4: a wavefront
// is 64 consecutive cells of one row.  With WAVE, lanes past the grid's
// last column stay live (the shared gradients need every lane) and store
// nothing.
template <int TYPE, bool WAVE = false>
__global__ void __launch_bounds__(256) k_noise_fbm(float *__restrict__ out,
                                                   const float *__restrict__ ctrl,
                                                   const float *__restrict__ nxp,
                                                   const float *__restrict__ nyp,
                                                   NoiseArgs a)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.ny) return; // the whole wavefront
  if (!WAVE && i >= a.nx) return;
  const bool   valid = i < a.nx;
  const int    lane = threadIdx.x & 63;
  const int    last = min(63, a.nx - 1 - (int)(blockIdx.x * blockDim.x)); // uniform
  const size_t idx = (size_t)j * a.nx + i;
  float        ct = (ctrl && valid) ? ctrl[idx] : 1.f;
  float        w = (1.f - ct) + ct * a.weight;
  float        x, y;
  if (WAVE)
    noise_xy(i, j, a, 0.f, 0.f, x, y);
  else
    noise_xy(i, j, a, nxp ? nxp[idx] : 0.f, nyp ? nyp[idx] : 0.f, x, y);

  float n = 0.f, nf = 1.f, na = 0.6f;
  int   px = a.per_x, py = a.per_y;
  for (int k = 0; k < a.octaves; k++)
  {
    float v;
    if constexpr (WAVE)
      v = base_noise_wave<TYPE>(x * nf, y * nf, a.fseed, px, py, lane, last);
    else
      v = base_noise<TYPE>(x * nf, y * nf, a.fseed, px, py);
    n += v * na;
    na *= (1.f - w) + w * fminf(v + 1.f, 2.f) * 0.5f;
    na *= a.persistence;
    nf *= a.lacunarity;
    px = scale_period(px, a.lacunarity);
    py = scale_period(py, a.lacunarity);
  }
  if (valid) out[idx] = n;
}

// WAVE as in k_noise_fbm: no warping planes, the lattice values shared
// across each wavefront (64 cells of one row).
template <int TYPE, bool WAVE = false>
__global__ void __launch_bounds__(256) k_noise(float *__restrict__ out,
                                               const float *__restrict__ nxp,
                                               const float *__restrict__ nyp,
                                               NoiseArgs a)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= a.ny) return; // the whole wavefront
  if (!WAVE && i >= a.nx) return;
  const bool   valid = i < a.nx;
  const size_t idx = (size_t)j * a.nx + i;
  float        x, y, v;
  if constexpr (WAVE)
  {
    const int lane = threadIdx.x & 63;
    const int last = min(63, a.nx - 1 - (int)(blockIdx.x * blockDim.x)); // uniform
    noise_xy(i, j, a, 0.f, 0.f, x, y);
    v = base_noise_wave<TYPE>(x, y, a.fseed, a.per_x, a.per_y, lane, last);
  }
  else
  {
    noise_xy(i, j, a, nxp ? nxp[idx] : 0.f, nyp ? nyp[idx] : 0.f, x, y);
    v = base_noise<TYPE>(x, y, a.fseed, a.per_x, a.per_y);
  }
  if (valid) out[idx] = v;
}

static NoiseArgs make_noise_args(int          nx,
                                 int          ny,
                                 float        kx,
                                 float        ky,
                                 uint32_t     seed,
                                 int          octaves,
                                 float        weight,
                                 float        persistence,
                                 float        lacunarity,
                                 const float *bbox4,
                                 const int   *period2)
{
  NoiseArgs a;
  a.nx = nx;
  a.ny = ny;
  a.kx = kx;
  a.ky = ky;
  // fseed = rand(wang_hash(seed)) (noise.cl:483-484, _common_rand.cl:84-88)
  a.fseed = (float)xorshift32(wang_hash(seed)) * (1.f / 4294967296.f);
  a.octaves = octaves;
  a.weight = weight;
  a.persistence = persistence;
  a.lacunarity = lacunarity;
  a.b0 = bbox4 ? bbox4[0] : 0.f;
  a.b1 = bbox4 ? bbox4[1] : 1.f;
  a.b2 = bbox4 ? bbox4[2] : 0.f;
  a.b3 = bbox4 ? bbox4[3] : 1.f;
  a.per_x = period2 ? period2[0] : 0;
  a.per_y = period2 ? period2[1] : 0;
  return a;
}

int launch_noise_fbm(float       *out,
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
                     const int   *period2,
                     hipStream_t  st)
{
  const NoiseArgs a = make_noise_args(nx, ny, kx, ky, seed, octaves, weight, persistence,
                                      lacunarity, bbox4, period2);
  dim3 blk(64, 4);
  dim3 grd((nx + 63) / 64, (ny + 3) / 4);
#define HMG_FBM(T) k_noise_fbm<T><<<grd, blk, 0, st>>>(out, ctrl, noise_x, noise_y, a)
  // without warping planes, every type shares its lattice values across
  // each wavefront
  const bool wave = !noise_x && !noise_y;
#define HMG_FBM_WAVE(T)                                                                \
  if (wave)                                                                            \
    k_noise_fbm<T, true><<<grd, blk, 0, st>>>(out, ctrl, nullptr, nullptr, a);         \
  else                                                                                 \
    HMG_FBM(T)
  switch (noise_type)
  {
  case 1: HMG_FBM_WAVE(1); break;
  case 2: HMG_FBM_WAVE(2); break;
  case 3: HMG_FBM_WAVE(3); break;
  case 4: HMG_FBM_WAVE(4); break;
  case 6: HMG_FBM_WAVE(6); break;
  case 7: HMG_FBM_WAVE(7); break;
  case 9: HMG_FBM_WAVE(9); break;
  case 10: HMG_FBM_WAVE(10); break;
  default: HMG_FBM(0); break; // no kernel branch: zeros
  }
#undef HMG_FBM_WAVE
#undef HMG_FBM
  return check_launch("noise_fbm");
}

int launch_noise(float       *out,
                 int          nx,
                 int          ny,
                 int          noise_type,
                 float        kx,
                 float        ky,
                 uint32_t     seed,
                 const float *noise_x,
                 const float *noise_y,
                 const float *bbox4,
                 const int   *period2,
                 hipStream_t  st)
{
  const NoiseArgs a = make_noise_args(nx, ny, kx, ky, seed, 1, 0.f, 0.f, 0.f, bbox4, period2);
  dim3 blk(64, 4);
  dim3 grd((nx + 63) / 64, (ny + 3) / 4);
  // without warping planes, the lattice values are shared across each
  // wavefront (as in noise_fbm)
  const bool wave = !noise_x && !noise_y;
#define HMG_NOISE(T)                                                                   \
  if (wave)                                                                            \
    k_noise<T, true><<<grd, blk, 0, st>>>(out, nullptr, nullptr, a);                   \
  else                                                                                 \
    k_noise<T><<<grd, blk, 0, st>>>(out, noise_x, noise_y, a)
  switch (noise_type)
  {
  case 1: HMG_NOISE(1); break;
  case 2: HMG_NOISE(2); break;
  case 3: HMG_NOISE(3); break;
  case 4: HMG_NOISE(4); break;
  case 6: HMG_NOISE(6); break;
  case 7: HMG_NOISE(7); break;
  case 9: HMG_NOISE(9); break;
  case 10: HMG_NOISE(10); break;
  default: k_noise<0><<<grd, blk, 0, st>>>(out, noise_x, noise_y, a); break;
  }
#undef HMG_NOISE
  return check_launch("noise");
}

static int check_noise_args(int nx, int ny, int noise_type, int octaves)
{
  if (nx <= 0 || ny <= 0) return set_error(HMG_ERR_INVALID, "shape must be positive");
  if ((size_t)nx * (size_t)ny > ((size_t)1 << 31))
    return set_error(HMG_ERR_INVALID, "grid larger than 2^31 cells");
  if (octaves < 0) return set_error(HMG_ERR_INVALID, "octaves must be >= 0");
  // hmap::NoiseType has 13 values (PARBERRY 0 .. WORLEY_VALUE 12)
  if (noise_type < 0 || noise_type > 12)
    return set_error(HMG_ERR_INVALID, "noise type %d is not a hmap::NoiseType", noise_type);
  return HMG_OK;
}

// --- remap on the device: min/max reduction then an affine pass --------

__device__ __forceinline__ float wave_min(float v)
{
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ void __launch_bounds__(256) k_minmax(const float *__restrict__ a, size_t n, float *mm)
{
  float  lo = INFINITY, hi = -INFINITY;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
  {
    float x = a[k];
    lo = fminf(lo, x);
    hi = fmaxf(hi, x);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  __shared__ float slo[4], shi[4];
  int              w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
  {
    slo[w] = lo;
    shi[w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0)
  {
    for (int q = 1; q < 4; q++)
    {
      lo = fminf(lo, slo[q]);
      hi = fmaxf(hi, shi[q]);
    }
    // per-block partials; float min/max do not depend on order
    mm[2 + 2 * blockIdx.x] = lo;
    mm[3 + 2 * blockIdx.x] = hi;
  }
}

__global__ void __launch_bounds__(256) k_remap(float *__restrict__ a,
                                               size_t n,
                                               const float *__restrict__ mm,
                                               int          nblk,
                                               float        vmin,
                                               float        vmax)
{
  __shared__ float s_scale, s_offset, s_flat;
  if (threadIdx.x == 0)
  {
    float lo = mm[2], hi = mm[3];
    for (int b = 1; b < nblk; b++)
    {
      lo = fminf(lo, mm[2 + 2 * b]);
      hi = fmaxf(hi, mm[3 + 2 * b]);
    }
    // range.cpp:342-361: scale = (vmax - vmin)/(max - min),
    // offset = vmin - min*scale; x*scale + offset
    s_flat = lo == hi ? 1.f : 0.f;
    float scale = lo == hi ? 0.f : (vmax - vmin) / (hi - lo);
    s_scale = scale;
    s_offset = vmin - lo * scale;
  }
  __syncthreads();
  float  scale = s_scale, offset = s_offset;
  bool   flat = s_flat != 0.f;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    a[k] = flat ? vmin : a[k] * scale + offset;
}

int launch_minmax(const float *a, size_t n, float *mm, int nblk, hipStream_t st)
{
  k_minmax<<<nblk, 256, 0, st>>>(a, n, mm);
  return check_launch("minmax");
}

} // namespace hmg

using namespace hmg;

extern "C" {

int hmg_noise_fbm_dev(hmg_ctx     *ctx,
                      float       *out_dev,
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
                      const float *ctrl_dev,
                      const float *noise_x_dev,
                      const float *noise_y_dev,
                      const float *bbox4,
                      const int   *period2)
{
  clear_error();
  HMG_TRY(check_noise_args(nx, ny, noise_type, octaves));
  if (!out_dev) return set_error(HMG_ERR_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  return launch_noise_fbm(out_dev, nx, ny, noise_type, kx, ky, seed, octaves, weight, persistence,
                          lacunarity, ctrl_dev, noise_x_dev, noise_y_dev, bbox4, period2,
                          ctx->stream);
}

int hmg_noise_fbm(hmg_ctx     *ctx,
                  float       *out_host,
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
                  const float *ctrl_host,
                  const float *noise_x_host,
                  const float *noise_y_host,
                  const float *bbox4,
                  const int   *period2)
{
  clear_error();
  HMG_TRY(check_noise_args(nx, ny, noise_type, octaves));
  if (!out_host) return set_error(HMG_ERR_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  size_t bytes = (size_t)nx * ny * sizeof(float);
  DevBuf out, c, gx, gy;
  HMG_TRY(out.get(ctx->pool, bytes));
  const float *pc = nullptr, *px = nullptr, *py = nullptr;
  if (ctrl_host)
  {
    HMG_TRY(c.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(c.ptr, ctrl_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    pc = c.as<float>();
  }
  if (noise_x_host)
  {
    HMG_TRY(gx.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(gx.ptr, noise_x_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    px = gx.as<float>();
  }
  if (noise_y_host)
  {
    HMG_TRY(gy.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(gy.ptr, noise_y_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    py = gy.as<float>();
  }
  HMG_TRY(launch_noise_fbm(out.as<float>(), nx, ny, noise_type, kx, ky, seed, octaves, weight,
                           persistence, lacunarity, pc, px, py, bbox4, period2,
                           ctx->stream));
  HMG_HIP(hipMemcpyAsync(out_host, out.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

int hmg_noise_dev(hmg_ctx     *ctx,
                  float       *out_dev,
                  int          nx,
                  int          ny,
                  int          noise_type,
                  float        kx,
                  float        ky,
                  uint32_t     seed,
                  const float *noise_x_dev,
                  const float *noise_y_dev,
                  const float *bbox4,
                  const int   *period2)
{
  clear_error();
  HMG_TRY(check_noise_args(nx, ny, noise_type, 0));
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!out_dev) return set_error(HMG_ERR_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  return launch_noise(out_dev, nx, ny, noise_type, kx, ky, seed, noise_x_dev, noise_y_dev, bbox4,
                      period2, ctx->stream);
}

int hmg_noise(hmg_ctx     *ctx,
              float       *out_host,
              int          nx,
              int          ny,
              int          noise_type,
              float        kx,
              float        ky,
              uint32_t     seed,
              const float *noise_x_host,
              const float *noise_y_host,
              const float *bbox4,
              const int   *period2)
{
  clear_error();
  HMG_TRY(check_noise_args(nx, ny, noise_type, 0));
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  if (!out_host) return set_error(HMG_ERR_INVALID, "null output");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  size_t bytes = (size_t)nx * ny * sizeof(float);
  DevBuf out, gx, gy;
  HMG_TRY(out.get(ctx->pool, bytes));
  const float *px = nullptr, *py = nullptr;
  if (noise_x_host)
  {
    HMG_TRY(gx.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(gx.ptr, noise_x_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    px = gx.as<float>();
  }
  if (noise_y_host)
  {
    HMG_TRY(gy.get(ctx->pool, bytes));
    HMG_HIP(hipMemcpyAsync(gy.ptr, noise_y_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    py = gy.as<float>();
  }
  HMG_TRY(launch_noise(out.as<float>(), nx, ny, noise_type, kx, ky, seed, px, py, bbox4,
                       period2, ctx->stream));
  HMG_HIP(hipMemcpyAsync(out_host, out.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

int hmg_remap_dev(hmg_ctx *ctx, float *a_dev, size_t n, float vmin, float vmax)
{
  clear_error();
  if (!a_dev && n) return set_error(HMG_ERR_INVALID, "null array");
  if (n == 0) return HMG_OK;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  HMG_TRY(ctx_enter(ctx));
  int nblk = (int)((n + 255) / 256);
  if (nblk > 1024) nblk = 1024;
  DevBuf mm;
  HMG_TRY(mm.get(ctx->pool, sizeof(float) * (2 + 2 * (size_t)nblk)));
  k_minmax<<<nblk, 256, 0, ctx->stream>>>(a_dev, n, mm.as<float>());
  HMG_TRY(check_launch("minmax"));
  k_remap<<<nblk, 256, 0, ctx->stream>>>(a_dev, n, mm.as<float>(), nblk, vmin, vmax);
  HMG_TRY(check_launch("remap"));
  // the scratch buffer returns to the pool; keep it alive until the stream
  // has consumed it
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

} // extern "C"
