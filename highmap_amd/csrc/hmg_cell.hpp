// hmg_cell.hpp -- per-cell arithmetic of the erosion hot path, shared by
// every kernel (plain and fused) so that they compute bit-identical values.
//
// Build contract: compiled with -ffp-contract=off, correctly rounded FP32
// division and sqrt, denormals kept.  Each source operation below is one
// IEEE-754 FP32 operation; the only fused multiply-adds are explicit fmaf
// calls in the project-written sin/cos.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace hmg
{

// ----------------------------------------------------------------- storage

// Storage codecs for the water, sediment and flux planes.  Values are
// rounded to nearest even when stored; all arithmetic stays FP32.
// st_rounded(x) is st(x) for an x that rnd<S> returned (already a stored
// value): the bit pattern is taken as is, with no second rounding.
struct StF32
{
  using T = float;
  // ZD: heights and water depths in one float2 plane (StF32ZD below)
  static constexpr bool ZD = false;
  static __device__ __forceinline__ float ld(T x) { return x; }
  static __device__ __forceinline__ T     st(float x) { return x; }
  static __device__ __forceinline__ T     st_rounded(float x) { return x; }
};

// FP32 state with the heights and the water depths interleaved in one
// float2 plane (x the height, y the depth): the fused kernels then read and
// write one 8-byte stream per cell for both instead of two 4-byte streams.
// The z planes a kernel receives hold float2; its D planes are unused.
struct StF32ZD : StF32
{
  static constexpr bool ZD = true;
};

struct StBF16
{
  using T = uint16_t;
  static constexpr bool ZD = false;
  static __device__ __forceinline__ float ld(T x)
  {
    return __uint_as_float(((uint32_t)x) << 16);
  }
  // Round to nearest even.  gfx950's conversion instruction
  // (v_cvt_pk_bf16_f32) gives exactly that for every non-NaN float,
  // subnormals and infinities included (compared with the integer form
  // below over all 2^32 patterns); NaNs keep the integer form's rule (the
  // top 16 bits, quiet bit set when low payload bits are dropped), taken in
  // a branch that no lane enters unless it holds a NaN.
  static __device__ __forceinline__ T st(float x)
  {
    const __bf16 h = (__bf16)x;
    T            r = __builtin_bit_cast(T, h);
    if (__builtin_expect(x != x, 0))
    {
      const uint32_t b = __float_as_uint(x);
      r = (T)((b >> 16) | ((b & 0xffffu) ? 0x40u : 0u));
    }
    return r;
  }
  // ld(st(x)) has the 16 stored bits on top and zeros below, NaNs too
  static __device__ __forceinline__ T st_rounded(float x) { return (T)(__float_as_uint(x) >> 16); }
  // the integer form: add half an ulp (ties to even) and truncate
  static __device__ __forceinline__ T st_int(float x)
  {
    uint32_t b = __float_as_uint(x);
    if ((b & 0x7f800000u) == 0x7f800000u) return (T)((b >> 16) | ((b & 0xffffu) ? 0x40u : 0u));
    b += 0x7fffu + ((b >> 16) & 1u);
    return (T)(b >> 16);
  }
};

struct StF16
{
  using T = _Float16;
  static constexpr bool ZD = false;
  static __device__ __forceinline__ float ld(T x) { return (float)x; }
  static __device__ __forceinline__ T     st(float x)
  {
    // The empty asm makes x an FP32 value in a register before the
    // conversion.  Without it the compiler may merge the FP32 multiply or
    // add that produced x with the conversion into one mixed-precision
    // instruction (v_fma_mixlo_f16), which rounds once to FP16 instead of
    // first to FP32 and then to FP16, and so changes results.
    asm volatile("" : "+v"(x));
    return (_Float16)x;
  }
  // an FP16 value widened to FP32 converts back exactly
  static __device__ __forceinline__ T st_rounded(float x) { return st(x); }
};

// Rounds a value to what the storage type would hold.
template <class S>
__device__ __forceinline__ float rnd(float x)
{
  return S::ld(S::st(x));
}

// rnd of four values.  BF16: two packed conversions (v_cvt_pk_bf16_f32,
// round to nearest even as st) and one unordered compare per pair for the
// NaN test, instead of a conversion and a test per value; a wave with a NaN
// in any of them takes rnd of each.  Same results as rnd of each.
template <class S>
__device__ __forceinline__ void rnd4(float &a, float &b, float &c, float &d)
{
  a = rnd<S>(a);
  b = rnd<S>(b);
  c = rnd<S>(c);
  d = rnd<S>(d);
}

template <>
__device__ __forceinline__ void rnd4<StBF16>(float &a, float &b, float &c, float &d)
{
  if (__builtin_expect(__builtin_isunordered(a, b) || __builtin_isunordered(c, d), 0))
  {
    a = rnd<StBF16>(a);
    b = rnd<StBF16>(b);
    c = rnd<StBF16>(c);
    d = rnd<StBF16>(d);
    return;
  }
  typedef float  v2f __attribute__((ext_vector_type(2)));
  typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
  const uint32_t ab = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){a, b}, v2bf));
  const uint32_t cd = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){c, d}, v2bf));
  a = __uint_as_float(ab << 16);
  b = __uint_as_float(ab & 0xffff0000u);
  c = __uint_as_float(cd << 16);
  d = __uint_as_float(cd & 0xffff0000u);
}

// Four flux components stored together (fl, fr, ft, fb).
template <class S>
struct alignas(4 * sizeof(typename S::T)) Flux4
{
  typename S::T l, r, t, b;
};

// Element `idx` of a plane addressed as a uniform base plus an unsigned
// 32-bit byte offset, so the compiler can use the scalar-base + 32-bit
// vector-offset form of global loads and stores instead of per-lane 64-bit
// address arithmetic.  Planes are < 4 GiB (grids are limited to 2^31 cells;
// the largest element is a 16-byte flux quadruple for grids <= 2^28 cells).
template <class T>
__device__ __forceinline__ const T &at_off(const T *base, uint32_t idx)
{
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      idx * (uint32_t)sizeof(T));
}

template <class T>
__device__ __forceinline__ T &at_off(T *base, uint32_t idx)
{
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T));
}

// Loads and stores of streamed planes with the nontemporal hint: the
// streaming kernels read and write each such element once per launch, and
// the hint keeps them from displacing lines that are read again (the
// sediment plane the gathers sample).
template <class T>
__device__ __forceinline__ T ld_stream(const T *base, uint32_t idx)
{
  const char *p = reinterpret_cast<const char *>(base) + idx * (uint32_t)sizeof(T);
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  typedef uint32_t u2 __attribute__((ext_vector_type(2)));
  if constexpr (sizeof(T) == 16)
    return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const u4 *>(p)));
  else if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const u2 *>(p)));
  else if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p)));
  else
    return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(p)));
}

template <class T>
__device__ __forceinline__ void st_stream(T *base, uint32_t idx, T v)
{
  char *p = reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(T);
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  typedef uint32_t u2 __attribute__((ext_vector_type(2)));
  if constexpr (sizeof(T) == 16)
    __builtin_nontemporal_store(__builtin_bit_cast(u4, v), reinterpret_cast<u4 *>(p));
  else if constexpr (sizeof(T) == 8)
    __builtin_nontemporal_store(__builtin_bit_cast(u2, v), reinterpret_cast<u2 *>(p));
  else if constexpr (sizeof(T) == 4)
    __builtin_nontemporal_store(__builtin_bit_cast(uint32_t, v), reinterpret_cast<uint32_t *>(p));
  else
    __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, v), reinterpret_cast<uint16_t *>(p));
}

// ----------------------------------------------------- sin/cos and hashing

// Project-written sin/cos: Cody-Waite reduction by pi/2 with a three-part
// constant (fmaf at each step), then Taylor polynomials of sin (to r^9) and
// cos (to r^10) by Horner with fmaf.  The reference calls OpenCL sin/cos
// under -cl-fast-relaxed-math (HighMap/src/gpu_opencl/kernels/
// _common_rand.cl:15,56), whose results are vendor-defined.
__device__ __forceinline__ float reduce_pio2(float x, int &q)
{
  float k = __builtin_rintf(x * 0x1.45f306p-1f);
  float r = fmaf(-k, 0x1.921fb6p+0f, x);
  r = fmaf(-k, -0x1.777a5cp-25f, r);
  r = fmaf(-k, -0x1.ee59dap-50f, r);
  // q = k mod 4 for every k: floats of magnitude 2^25 and above are
  // multiples of 4, so clamping there keeps the quadrant and keeps the
  // conversion in range (a NaN k gives 0)
  q = ((int)fminf(fmaxf(k, -0x1p25f), 0x1p25f)) & 3;
  return r;
}

__device__ __forceinline__ float sin_poly(float r)
{
  float r2 = r * r;
  float p = fmaf(r2, 1.f / 362880.f, -1.f / 5040.f);
  p = fmaf(r2, p, 1.f / 120.f);
  p = fmaf(r2, p, -1.f / 6.f);
  return fmaf(r * r2, p, r);
}

__device__ __forceinline__ float cos_poly(float r)
{
  float r2 = r * r;
  float p = fmaf(r2, -1.f / 3628800.f, 1.f / 40320.f);
  p = fmaf(r2, p, -1.f / 720.f);
  p = fmaf(r2, p, 1.f / 24.f);
  p = fmaf(r2, p, -0.5f);
  return fmaf(r2, p, 1.f);
}

__device__ __forceinline__ float hm_sin(float x)
{
  int   q;
  float r = reduce_pio2(x, q);
  float s = sin_poly(r), c = cos_poly(r);
  float v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}

// cos and sin of one angle, sharing the reduction.
__device__ __forceinline__ void hm_sincos(float x, float &sn, float &cs)
{
  int   q;
  float r = reduce_pio2(x, q);
  float s = sin_poly(r), c = cos_poly(r);
  // sin: q=0 s, 1 c, 2 -s, 3 -c ; cos: q=0 c, 1 -s, 2 -c, 3 s
  sn = (q & 1) ? c : s;
  sn = (q & 2) ? -sn : sn;
  cs = (q & 1) ? s : c;
  cs = ((q + 1) & 2) ? -cs : cs;
}

// Project-written exp (OpenCL's exp, used by snow_simulation.cl, is not
// correctly rounded): x = k ln2 + r with k = rint(x * RN(1/ln2)), r by two
// fmaf steps (ln2 as its nearest float plus the remainder), e^r by the
// degree-7 Taylor polynomial in Horner form with fmaf, then times 2^k as
// two exact powers of two so that only the last product rounds (into the
// subnormal range for results that small).  About 1 ulp.
__device__ __forceinline__ float hm_exp(float x)
{
  if (x != x) return x;
  if (x > 89.f) return __builtin_inff();
  if (x < -104.f) return 0.f;
  const float k = __builtin_rintf(x * 0x1.715476p+0f);
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
  return (p * __int_as_float((k1 + 127) << 23)) * __int_as_float((k2 + 127) << 23);
}

// p * 2^k for k in [-151, 128]: two exact powers of two, so that only the
// last product rounds (into the subnormal range or to infinity).
__device__ __forceinline__ float hm_scale2(float p, int k)
{
  const int k1 = k / 2, k2 = k - k1;
  return (p * __int_as_float((k1 + 127) << 23)) * __int_as_float((k2 + 127) << 23);
}

// Project-written pow, FP32 only (OpenCL's pow under -cl-fast-relaxed-math,
// which the erosion pass calls for k_discharge_exp != 1,
// hydraulic_vpipes.cl:167, is the vendor's).  The special cases of C99 pow
// first; then |x| = m 2^e with m in [sqrt(1/2), sqrt(2)) (subnormals scaled
// by 2^24 first); ln m = 2 atanh(t) with t = (m - 1) / (m + 1) held as two
// floats (m - 1 exact, m + 1 split by Fast2Sum, the division's fma residual
// corrected for the low part) and 2 atanh(t) = 2t + t^3 P(t^2), P the
// series 2/3 + 2/5 s + ... + 2/15 s^6, summed into two floats; log2 |x| =
// e + ln m / ln 2 and w = y log2 |x| as two floats (products split exactly
// by fma, the integer part added by TwoSum); 2^w = 2^k 2^f, k = rint(w),
// 2^f by the degree-8 Taylor polynomial of exp(f ln 2) with fma, then times
// 2^k as two exact powers of two.  Every step is one IEEE float operation,
// so the oracle's copy gives the same bits.  Within 0.85 ulp of pow for
// |y| <= 10 on a sweep (2.6 ulp at |y| = 128).
__device__ __forceinline__ float hm_powf(float x, float y)
{
  if (y == 0.f) return 1.f;
  if (x == 1.f) return 1.f;
  if (x != x || y != y) return x + y;
  const float ay = fabsf(y);
  const bool  yint = ay >= 0x1p24f || __builtin_rintf(y) == y;
  const bool  yodd = yint && ay < 0x1p24f && (((int)y) & 1);
  const bool  xneg = (__float_as_uint(x) >> 31) != 0;
  float       ax = fabsf(x);
  if (xneg && !yint && ax != 0.f && ax != __builtin_inff()) return __builtin_nanf("");
  if (ay == __builtin_inff())
    return ax == 1.f ? 1.f : ((ax < 1.f) == (y < 0.f) ? __builtin_inff() : 0.f);
  const float sgn = (xneg && yodd) ? -1.f : 1.f;
  if (ax == 0.f) return y > 0.f ? sgn * 0.f : sgn * __builtin_inff();
  if (ax == __builtin_inff()) return y > 0.f ? sgn * __builtin_inff() : sgn * 0.f;

  int e = 0;
  if (ax < 0x1p-126f)
  {
    ax *= 0x1p24f;
    e = -24;
  }
  uint32_t b = __float_as_uint(ax);
  e += (int)(b >> 23) - 127;
  float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (m >= 0x1.6a09e6p+0f)
  {
    m *= 0.5f;
    e += 1;
  }
  // ln m = 2 atanh(t), t = (m - 1) / (m + 1)
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
  // log2 |x| = e + (lh + ll) log2(e)
  const float ph = lh * 0x1.715476p+0f;
  const float pl = fmaf(lh, 0x1.715476p+0f, -ph) + (lh * 0x1.4ae0c0p-26f + ll * 0x1.715476p+0f);
  const float fe = (float)e;
  const float gh = fe + ph;
  const float bv = gh - fe;
  const float gl = ((fe - (gh - bv)) + (ph - bv)) + pl;
  // w = y log2 |x|
  const float wh = y * gh;
  if (wh >= 128.5f) return sgn * __builtin_inff();
  if (wh <= -151.f) return sgn * 0.f;
  const float wl = fmaf(y, gh, -wh) + y * gl;
  const float k = __builtin_rintf(wh);
  const float f = (wh - k) + wl;
  float       q = fmaf(f, 0x1.62c022p-20f, 0x1.ffcbfcp-17f);
  q = fmaf(f, q, 0x1.430912p-13f);
  q = fmaf(f, q, 0x1.5d87fep-10f);
  q = fmaf(f, q, 0x1.3b2ab6p-7f);
  q = fmaf(f, q, 0x1.c6b08ep-5f);
  q = fmaf(f, q, 0x1.ebfbe0p-3f);
  q = fmaf(f, q, 0x1.62e430p-1f);
  q = fmaf(f, q, 1.f);
  return sgn * hm_scale2(q, (int)k);
}

// OpenCL fract(): min(x - floor(x), 0x1.fffffep-1f).
__device__ __forceinline__ float fract_cl(float x)
{
  return fminf(x - floorf(x), 0x1.fffffep-1f);
}

// wang_hash (HighMap/src/gpu_opencl/kernels/_common_rand.cl:66-74).
__host__ __device__ __forceinline__ uint32_t wang_hash(uint32_t s)
{
  s = (s ^ 61u) ^ (s >> 16);
  s *= 9u;
  s = s ^ (s >> 4);
  s *= 0x27d4eb2du;
  s = s ^ (s >> 15);
  return s;
}

// One xorshift32 (13, 17, 5) step (_common_rand.cl:76-82).
__host__ __device__ __forceinline__ uint32_t xorshift32(uint32_t x)
{
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}

// hash12f (_common_rand.cl:12-17); dot product as two products and a sum.
__device__ __forceinline__ float hash12f(float px, float py, float fseed)
{
  float dp = px * 127.1f + py * 311.7f;
  return fract_cl(hm_sin(dp + fseed) * 43758.5453123f);
}

// ----------------------------------------------------------------- thermal

// x / 1.414f, correctly rounded: q0 = x * r with r = RN(1/1.414f), then one
// fma residual and one fma correction.  Equal to the IEEE quotient for every
// float x with 2^-100 <= |x| < inf (checked over all 2^32 inputs by
// tests/cpp/div_const_check.c); other x take the division.
__device__ __forceinline__ float div_1414(float x)
{
  constexpr float B = 1.414f;
  constexpr float R = 0x1.6a17e6p-1f; // RN(1 / 1.414f)
  const float     ax = fabsf(x);
  if (__builtin_expect(ax >= 0x1p-100f && ax < __builtin_inff(), 1))
  {
    const float q0 = x * R;
    return fmaf(fmaf(-q0, B, x), R, q0);
  }
  return x / B;
}

// helper_thermal_exchange (HighMap/src/gpu_opencl/kernels/thermal.cl:5-24),
// without branches.  With d = self - other (other - self is exactly -d):
// the reference returns -0.2 (d - c T) / c when d > c T, 0.2 (-d - c T) / c
// when -d > c T, else 0, i.e. +-(0.2 (|d| - c T)) / c when |d| > c T, with
// the sign of -d (NaN compares false: 0).  The product and the quotient are
// rounded as in the reference; negation is exact.  c is 1 (axial, an exact
// division) or the literal 1.414 (diagonal).
template <bool DIAG>
__device__ __forceinline__ float th_exchange(float self, float other, float talus)
{
  const float max_dif = (DIAG ? 1.414f : 1.f) * talus;
  const float d = self - other;
  const float ad = fabsf(d);
  const bool  over = ad > max_dif;
  const float x = 0.2f * (ad - max_dif);
  // lanes without an exchange divide a harmless 1 (their result is dropped)
  const float m = DIAG ? div_1414(over ? x : 1.f) : x;
  const float e = self > other ? -m : m;
  return over ? e : 0.f;
}

// Interior thermal update (thermal.cl:41-57): neighbours in the reference's
// order (-1,0),(0,1),(0,-1),(1,0),(-1,-1),(-1,1),(1,-1),(1,1); distance
// 1.414 (the literal) on diagonals.  n[k] are the neighbour heights in that
// order.
__device__ __forceinline__ float th_interior(float val, const float n[8], float t)
{
  float amount = 0.f;
  amount += th_exchange<false>(val, n[0], t);
  amount += th_exchange<false>(val, n[1], t);
  amount += th_exchange<false>(val, n[2], t);
  amount += th_exchange<false>(val, n[3], t);
  amount += th_exchange<true>(val, n[4], t);
  amount += th_exchange<true>(val, n[5], t);
  amount += th_exchange<true>(val, n[6], t);
  amount += th_exchange<true>(val, n[7], t);
  return val + amount;
}

// ------------------------------------------------------------------ vpipes

// Constants of one vpipes launch, formed on the host in FP32 exactly as the
// reference forms them.
struct VpConst
{
  int   nx, ny;
  float dt;       // time step
  float dtg;      // dt * gravity (gravity 1 in the reference)
  float c0, c1;   // flux diffusion: strength and 1 - 4*strength
  int   diffuse;  // flux_diffusion
  float dmin;     // 0.001 * water_height / dt (water pass)
  float fnx;      // (float)nx (erosion pass gradient scale)
  float dmax;     // water_height (erosion pass)
  float kc, ke, kd, kexp, ratio;
  float ce;       // 1 - dt * evap_rate
  float rr;       // rain_per_iter (extension)
  int   maintain; // maintain_water_volume
  // p / nx in the sampler is a multiplication by 2^-k when nx = 2^k (the
  // same correctly rounded value); otherwise an IEEE division.
  float inv_nx, inv_ny;
  int   pow2x, pow2y;
};

// hydraulic_vpipes_flow_pass (hydraulic_vpipes.cl:33-68) for one cell.
// Inputs: own z, d1; neighbour z, d1 (left, right, top = j+1, bottom = j-1);
// own fluxes.  Output fluxes in (l, r, t, b).
__device__ __forceinline__ void vp_flow(float        zc,
                                        float        dc,
                                        float        zl,
                                        float        dl,
                                        float        zr,
                                        float        dr,
                                        float        zt,
                                        float        dtp,
                                        float        zb,
                                        float        db,
                                        float       &l,
                                        float       &r,
                                        float       &t,
                                        float       &b,
                                        const VpConst &k)
{
  float h0 = zc + dc;
  float dhl = h0 - zl - dl;
  float dhr = h0 - zr - dr;
  float dht = h0 - zt - dtp;
  float dhb = h0 - zb - db;
  l = fmaxf(0.f, l + k.dtg * dhl);
  r = fmaxf(0.f, r + k.dtg * dhr);
  t = fmaxf(0.f, t + k.dtg * dht);
  b = fmaxf(0.f, b + k.dtg * dhb);
  if (k.diffuse)
  {
    float f_diff = k.c0 * (l + r + t + b);
    l = k.c1 * l + f_diff;
    r = k.c1 * r + f_diff;
    t = k.c1 * t + f_diff;
    b = k.c1 * b + f_diff;
  }
  float sum = l + r + t + b;
  float kk = 0.f;
  if (sum > 1e-5f) kk = dc / (sum * k.dt);
  kk = fminf(fmaxf(kk, 0.f), 1.f);
  l = l * kk;
  r = r * kk;
  t = t * kk;
  b = b * kk;
}

// Net water change of the water pass (hydraulic_vpipes.cl:97-102).
// frl = fr(i-1), ftb = ft(j-1), flr = fl(i+1), fbt = fb(j+1).
__device__ __forceinline__ float vp_depth(float        d1,
                                          float        frl,
                                          float        ftb,
                                          float        flr,
                                          float        fbt,
                                          float        fl,
                                          float        fr,
                                          float        ft,
                                          float        fb,
                                          const VpConst &k)
{
  float dv = k.dt * (frl + ftb + flr + fbt - fl - fr - ft - fb);
  return fmaxf(0.f, d1 + dv);
}

// Velocities of the water pass (hydraulic_vpipes.cl:104-115), given the
// new depth d2.
__device__ __forceinline__ void vp_velocity(float        d2,
                                            float        frl,
                                            float        ftb,
                                            float        flr,
                                            float        fbt,
                                            float        fl,
                                            float        fr,
                                            float        ft,
                                            float        fb,
                                            float       &u,
                                            float       &v,
                                            const VpConst &k)
{
  float un = 0.5f * (frl - fl + fr - flr);
  float vn = 0.5f * (ftb - fb + ft - fbt);
  float dmean = fmaxf(k.dmin, d2);
  u = un / dmean;
  v = vn / dmean;
}

__device__ __forceinline__ float hm_hypot(float x, float y)
{
  return sqrtf(x * x + y * y);
}

// sqrtf(x), correctly rounded, for x >= 2^-96 (and +0, +inf): the hardware
// estimate, then the neighbours one ulp either side are chosen by the signs
// of their fma residuals.  These are the steps the compiler emits for sqrtf,
// less its scaling of inputs below 2^-96 and its select for special classes,
// which leave such x unchanged.  For 0 <= x < 2^-96 the result is not
// correctly rounded but stays below 2^-47.  Both facts are checked over
// every float pattern on the GPU (tests/cpp/arith_check.cpp).
__device__ __forceinline__ float sqrt_big(float x)
{
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  float       r = (fmaf(-sm, s, x) <= 0.f) ? sm : s;
  r = (fmaf(-sp, s, x) > 0.f) ? sp : r;
  return r;
}

// sqrtf(x) for x >= +0 (a sum of squares): sqrt_big, unless a lane of the
// wave holds a value in (0, 2^-96), then the full sequence for the whole
// wave (a uniform branch).
__device__ __forceinline__ float sqrt_sumsq(float x)
{
  if (__builtin_expect(__any(__float_as_uint(x) - 1u < 0x0f7fffffu), 0)) return sqrtf(x);
  return sqrt_big(x);
}

// hydraulic_vpipes_erosion_pass (hydraulic_vpipes.cl:149-183) for one cell.
// zr/zl/zt/zb and sr/sl/st/sb are the neighbour values right (i+1), left
// (i-1), top (j+1), bottom (j-1).  KEXP1: k_discharge_exp is known to be 1,
// so pow(discharge, 1) = discharge is not evaluated at all.
template <bool KEXP1 = false>
__device__ __forceinline__ void vp_erosion(float        zc,
                                           float        sc,
                                           float        zr,
                                           float        zl,
                                           float        zt,
                                           float        zb,
                                           float        sr,
                                           float        sl,
                                           float        st,
                                           float        sb,
                                           float        d2,
                                           float        u,
                                           float        v,
                                           float       &zo,
                                           float       &so,
                                           const VpConst &k)
{
  float dzx = 0.5f * (zr - zl + sr - sl);
  float dzy = 0.5f * (zt - zb + st - sb);
  // The slope only enters through max(1e-3, nx * slope): below 2^-96 the
  // exact root and sqrt_big's (both < 2^-47, nx < 2^31) give 1e-3 alike.
  float talus = sqrt_big(dzx * dzx + dzy * dzy);
  float dzn = fmaxf(1e-3f, k.fnx * talus);
  float salpha = fmaxf(0.01f, dzn / sqrt_big(1.f + dzn * dzn)); // argument >= 1
  float speed = sqrt_sumsq(u * u + v * v);
  float depth = fminf(d2 / k.dmax, k.ratio);
  float discharge = depth * speed;
  float pw = (KEXP1 || k.kexp == 1.f) ? discharge : hm_powf(discharge, k.kexp);
  float capa = k.kc * pw * salpha;
  if (capa > sc)
  {
    float amount = k.ke * (capa - sc);
    zo = zc - amount;
    so = sc + amount;
  }
  else
  {
    float amount = k.kd * (sc - capa);
    zo = zc + amount;
    so = sc - amount;
  }
}

// One axis of the linear mirrored-repeat sampler (OpenCL 1.2 rule for
// CLK_ADDRESS_MIRRORED_REPEAT | CLK_FILTER_LINEAR applied to p/w).
__device__ __forceinline__ void mirror_axis(float p,
                                            int   w,
                                            float inv_w,
                                            int   pow2,
                                            int  &i0,
                                            int  &i1,
                                            float &a)
{
  float fw = (float)w;
  float sg = pow2 ? p * inv_w : p / fw;
  float sp = fabsf(sg - 2.f * __builtin_rintf(0.5f * sg));
  float q = sp * fw - 0.5f;
  float q0 = floorf(q);
  a = q - q0;
  int kq = (int)q0;
  i0 = min(max(kq, 0), w - 1);
  i1 = min(max(kq + 1, 0), w - 1);
}

// The transport sample split in two: `issue` computes the weights and loads
// the four texels, `finish` forms the weighted sum, so that other work can
// sit between the loads and their use.  Same operations as vp_transport.
// With `direct` set, all four texels are read at (i, j) instead (the caller
// then uses s00 alone): selects, not a branch, so that the same loads are in
// flight on both paths.
struct TexelFetch
{
  float w00, w10, w01, w11;
  float s00, s10, s01, s11;
};

template <class S>
__device__ __forceinline__ TexelFetch vp_transport_issue(const typename S::T *sp,
                                                         int                  i,
                                                         int                  j,
                                                         float                u,
                                                         float                v,
                                                         const VpConst       &k,
                                                         bool                 direct = false)
{
  float      px = (float)i - k.dt * u;
  float      py = (float)j - k.dt * v;
  int        i0, i1, j0, j1;
  float      a, b;
  TexelFetch t;
  mirror_axis(px, k.nx, k.inv_nx, k.pow2x, i0, i1, a);
  mirror_axis(py, k.ny, k.inv_ny, k.pow2y, j0, j1, b);
  i0 = direct ? i : i0;
  i1 = direct ? i : i1;
  j0 = direct ? j : j0;
  j1 = direct ? j : j1;
  t.w00 = (1.f - a) * (1.f - b);
  t.w10 = a * (1.f - b);
  t.w01 = (1.f - a) * b;
  t.w11 = a * b;
  const uint32_t r0 = (uint32_t)(j0 * k.nx), r1 = (uint32_t)(j1 * k.nx);
  t.s00 = S::ld(at_off(sp, r0 + i0));
  t.s10 = S::ld(at_off(sp, r0 + i1));
  t.s01 = S::ld(at_off(sp, r1 + i0));
  t.s11 = S::ld(at_off(sp, r1 + i1));
  return t;
}

__device__ __forceinline__ float vp_transport_finish(const TexelFetch &t)
{
  return t.w00 * t.s00 + t.w10 * t.s10 + t.w01 * t.s01 + t.w11 * t.s11;
}

// Sediment transport (hydraulic_vpipes.cl:204-210): bilinear sample of the
// sediment plane at the backtrace (i - dt*u, j - dt*v).
template <class S>
__device__ __forceinline__ float vp_transport(const typename S::T *sp,
                                              int                  i,
                                              int                  j,
                                              float                u,
                                              float                v,
                                              const VpConst       &k)
{
  float px = (float)i - k.dt * u;
  float py = (float)j - k.dt * v;
  int   i0, i1, j0, j1;
  float a, b;
  mirror_axis(px, k.nx, k.inv_nx, k.pow2x, i0, i1, a);
  mirror_axis(py, k.ny, k.inv_ny, k.pow2y, j0, j1, b);
  float  w00 = (1.f - a) * (1.f - b);
  float  w10 = a * (1.f - b);
  float  w01 = (1.f - a) * b;
  float  w11 = a * b;
  size_t r0 = (size_t)j0 * k.nx, r1 = (size_t)j1 * k.nx;
  float  s00 = S::ld(sp[r0 + i0]), s10 = S::ld(sp[r0 + i1]);
  float  s01 = S::ld(sp[r1 + i0]), s11 = S::ld(sp[r1 + i1]);
  return w00 * s00 + w10 * s10 + w01 * s01 + w11 * s11;
}

// Water depth at the start of an iteration from the stored pre-evaporation
// depth d2 of the previous one: evaporation, volume top-up and the optional
// per-iteration rain (hydraulic_vpipes_gpu.cpp:163-169).
__device__ __forceinline__ float vp_d1(float        d2prev,
                                       float        rain,
                                       int          has_prev,
                                       int          topup,
                                       float        wta,
                                       float        rf,
                                       const VpConst &k)
{
  float d = d2prev;
  if (has_prev)
  {
    d = d2prev * k.ce;
    if (topup) d = d + (wta * rain) / rf;
  }
  if (k.rr != 0.f) d = d + k.rr * rain;
  return d;
}

// --------------------------------------------- exact fixed-point volume sum

// A float x maps to the 128-bit integer round(x * 2^64) (exact for
// |x| >= 2^-41); sums of such integers do not depend on order.
struct Fix128
{
  uint64_t lo;
  int64_t  hi;
};

__device__ __forceinline__ Fix128 fix_term(float x)
{
  uint32_t b = __float_as_uint(x);
  uint32_t e = (b >> 23) & 0xffu;
  uint32_t m = (b & 0x7fffffu) | 0x800000u;
  uint64_t lo = 0, hi = 0;
  if (e != 0)
  {
    int sh = (int)e - 86;
    if (sh >= 0)
    {
      sh = sh > 103 ? 103 : sh;
      if (sh < 64)
      {
        lo = (uint64_t)m << sh;
        hi = sh > 40 ? ((uint64_t)m >> (64 - sh)) : 0;
      }
      else
        hi = (uint64_t)m << (sh - 64);
    }
    else if (sh > -25)
    {
      uint32_t rsh = (uint32_t)(-sh);
      uint32_t q = m >> rsh;
      uint32_t rem = m & ((1u << rsh) - 1u);
      uint32_t half = 1u << (rsh - 1);
      if (rem > half || (rem == half && (q & 1u))) q++;
      lo = q;
    }
  }
  if (b >> 31)
  {
    lo = ~lo;
    hi = ~hi;
    lo += 1;
    if (lo == 0) hi += 1;
  }
  return Fix128{lo, (int64_t)hi};
}

// fix_term for x with bit pattern below 1.0f (finite, sign clear, x < 1):
// then round-half-even(x * 2^64) < 2^64 and the high word is 0.  x * 2^64 is
// exact (subnormal x included) and rint rounds it half-even to an integer
// below 2^64, whose two 32-bit halves are exact in float.  Equal to
// fix_term_unit_int for every such pattern (checked on the GPU by
// tests/cpp/arith_check.cpp).
__device__ __forceinline__ uint64_t fix_term_unit(float x)
{
  const float    t = __builtin_rintf(x * 0x1p64f);
  const float    h = t * 0x1p-32f;
  const float    hf = floorf(h);
  const uint32_t hi = (uint32_t)hf;
  const uint32_t lo = (uint32_t)((h - hf) * 0x1p32f);
  return ((uint64_t)hi << 32) | lo;
}

// The same term by integer shifts, equal to fix_term(x).lo for x in [0, 1)
// (same shift, same rounding of the low bits).
__device__ __forceinline__ uint64_t fix_term_unit_int(float x)
{
  const uint32_t b = __float_as_uint(x);
  const uint32_t e = b >> 23;
  const uint32_t m = (b & 0x7fffffu) | 0x800000u;
  const int      sh = (int)e - 86;
  if (sh >= 0) return (uint64_t)m << sh; // e in [86, 126]: sh <= 40
  const uint32_t rsh = (uint32_t)(-sh);
  if (e == 0 || rsh > 24u) return 0; // zero, subnormal or below 2^-65
  const uint32_t q = m >> rsh;
  const uint32_t rem = m & ((1u << rsh) - 1u);
  const uint32_t half = 1u << (rsh - 1u);
  return (uint64_t)(q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u));
}

// a += t for a term below 2^64: one add with carry into the high word.
__device__ __forceinline__ void fix_add_u64(Fix128 &a, uint64_t t)
{
  unsigned long long carry = 0;
  a.lo = __builtin_addcll(a.lo, t, 0ull, &carry);
  a.hi = (int64_t)((uint64_t)a.hi + carry);
}

__device__ __forceinline__ void fix_add(Fix128 &a, const Fix128 &b)
{
  uint64_t lo = a.lo + b.lo;
  uint64_t c = lo < a.lo ? 1u : 0u;
  a.lo = lo;
  a.hi = (int64_t)((uint64_t)a.hi + (uint64_t)b.hi + c);
}

// acc += fix_term(x): the short form when every active lane of the wave holds
// a finite x in [0, 1) (water depths in height units), the general one
// otherwise.  The vote is uniform over the wave, so the branch does not
// diverge.
__device__ __forceinline__ void fix_accumulate(Fix128 &a, float x)
{
  if (__all(__float_as_uint(x) < 0x3f800000u))
    fix_add_u64(a, fix_term_unit(x));
  else
    fix_add(a, fix_term(x));
}

__host__ __device__ __forceinline__ double fix_to_double(uint64_t lo, int64_t hi)
{
  double hid = (double)(int32_t)(hi >> 32) * 4294967296.0 +
               (double)(uint32_t)(uint64_t)hi;
  double lod = (double)(uint32_t)(lo >> 32) * 4294967296.0 + (double)(uint32_t)lo;
  return hid + lod * 0x1p-64;
}

} // namespace hmg
