// hmg_internal.hpp -- runtime internals of libhmg shared by its sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "hmg.h"
#include "hmg_cell.hpp"

namespace hmg
{

// Thread-local error message; returns `code` so call sites can
// `return set_error(...)`.
int set_error(int code, const char *fmt, ...);
void clear_error();

#define HMG_HIP(expr)                                                          \
  do                                                                           \
  {                                                                            \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::hmg::set_error(e_ == hipErrorOutOfMemory ? HMG_ERR_OOM          \
                                                        : HMG_ERR_DEVICE,      \
                              "%s failed: %s (%s:%d)",                         \
                              #expr,                                           \
                              hipGetErrorString(e_),                           \
                              __FILE__,                                        \
                              __LINE__);                                       \
  } while (0)

#define HMG_TRY(expr)                                                          \
  do                                                                           \
  {                                                                            \
    int rc_ = (expr);                                                          \
    if (rc_ != HMG_OK) return rc_;                                             \
  } while (0)

// Caching device allocator of one context: freed blocks are kept by size and
// handed out again, so repeated drop-in calls do not hit hipMalloc.
class Pool
{
public:
  explicit Pool(int device) : device_(device) {}
  ~Pool();
  int  alloc(void **p, size_t bytes);
  void release(void *p);
  void trim();
  size_t held_bytes() const { return held_; }

private:
  int                                  device_;
  std::mutex                           mu_;
  std::multimap<size_t, void *>        free_;
  std::map<void *, size_t>             live_;
  size_t                               held_ = 0;
};

// RAII device buffer from a Pool.
struct DevBuf
{
  Pool  *pool = nullptr;
  void  *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  int get(Pool *pl, size_t n)
  {
    reset();
    pool = pl;
    bytes = n;
    return pl->alloc(&ptr, n);
  }
  void reset()
  {
    if (ptr && pool) pool->release(ptr);
    ptr = nullptr;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

} // namespace hmg

struct hmg_ctx
{
  int             device = 0;
  hipStream_t     stream = nullptr;
  int             num_cus = 0;
  int             wave_size = 64;
  std::recursive_mutex mu;
  hmg::Pool      *pool = nullptr;
};

namespace hmg
{

// Makes ctx->device current for the calling thread.
int ctx_enter(hmg_ctx *ctx);

// Checks the last kernel launch.
inline int check_launch(const char *what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(HMG_ERR_DEVICE, "launch of %s failed: %s", what, hipGetErrorString(e));
  return HMG_OK;
}

// Builds the launch constants of vpipes from the parameters, in FP32 exactly
// as the reference forms them (hydraulic_vpipes.cl:30-31,48-49,109,156-159;
// hydraulic_vpipes_gpu.cpp:163).
VpConst make_vp_const(const hmg_vpipes_params &p, int nx, int ny);

int validate_vp_params(const hmg_vpipes_params &p);

// Kernel-side entry points implemented in the .hip sources.
// Runs `iterations` Jacobi thermal steps on z, ping-ponging with zscratch.
// With result == nullptr the result ends in z (copied back after an odd
// number of launches); otherwise *result receives whichever of the two
// planes holds it and nothing is copied.  With extrapolate, the steps are
// followed by extrapolate_borders (folded into the last step's launch where
// the streaming kernel runs it).
int launch_thermal(hmg_ctx     *ctx,
                   float       *z,
                   float       *zscratch,
                   int          nx,
                   int          ny,
                   const float *talus,
                   float        talus_scalar,
                   int          iterations,
                   const float *bedrock,
                   hipStream_t  st,
                   float      **result = nullptr,
                   bool         extrapolate = false);
int launch_extrapolate_borders(float *z, int nx, int ny, hipStream_t st);

// launch_thermal for heights held beside the water depth in float2 planes
// (x the height, y the depth; the sessions' FP32 fused layout): thermal()
// with a scalar talus and its extrapolate_borders on the x components of zd,
// ping-ponging with zd_scratch (both nx*ny float2); each step carries the
// depths over unchanged.  *result receives the plane that holds the result.
int launch_thermal_zd(hmg_ctx     *ctx,
                      float2      *zd,
                      float2      *zd_scratch,
                      int          nx,
                      int          ny,
                      float        talus_scalar,
                      int          iterations,
                      hipStream_t  st,
                      float2     **result);
// The x (height) or y (depth) components of n float2 into a float plane,
// and back (the other component of zd kept).
int launch_zd_split(const float2 *zd, float *out, int component, size_t n, hipStream_t st);
int launch_zd_merge(float2 *zd, const float *in, int component, size_t n, hipStream_t st);

// HMG_ERR_NONFINITE if the n floats at p (device) hold a NaN or an
// infinity; waits for the context stream (vpipes.hip).
int check_finite_f32(hmg_ctx *ctx, const float *p, size_t n);

// Looks up, once per device, the device address of the static scratch plane
// of hmg_thermal_dev_noalloc (thermal.hip); called when a context is created.
int noalloc_scratch_setup(int device);

// Per-block min and max of a[0..n) into mm[2 + 2b], mm[3 + 2b] for the nblk
// blocks of the launch (nblk <= 1024; mm holds 2 + 2 * nblk floats).
int launch_minmax(const float *a, size_t n, float *mm, int nblk, hipStream_t st);

} // namespace hmg

// One erosion state resident on the device (see vpipes.hip).
struct hmg_session
{
  hmg_ctx          *ctx = nullptr;
  int               nx = 0, ny = 0;
  int               storage = HMG_STORAGE_F32;
  hmg_vpipes_params p{};
  hmg::VpConst      k{};
  size_t            esize = 4; // bytes per stored scalar
  // zd: the FP32 fused layout, heights and water depths as one float2 plane
  // (z[q] holds nx*ny float2: one 8-byte load and store per cell instead of
  // two 4-byte streams; D[q] is not allocated)
  bool              zd = false;
  hmg::DevBuf       z[2], D[2], Sp[2], F[2], rain, slots;
  hmg::DevBuf       d1, s, u, v; // plain-form intermediates
  hmg::DevBuf       zscratch;    // thermal ping-pong
  int               cur = 0;
  long              iter = 0;
  double            W0 = 0.0, R = 0.0;
  float             rf = 0.f;
  int               chunk_rows = 64;
  int               prefetch = 1;
  // Online launch-shape tuning of the fused kernel (vpipes.hip): the first
  // hydraulic iterations of a large session each run one candidate shape
  // between two events; the fastest shape is then kept (and cached for the
  // process).  Results do not depend on the shape.  tune_pos < 0: done.
  int                     tune_pos = -1;
  std::vector<hipEvent_t> tune_ev;
  std::vector<std::pair<int, int>> tune_cand; // (rows, variant) this session probes
  ~hmg_session()
  {
    for (hipEvent_t e : tune_ev) (void)hipEventDestroy(e);
  }
};

namespace hmg
{

// Advances sessions[i] by iters[i] hydraulic iterations with batched
// launches (vpipes.hip).  *ms (may be null) receives the device time.
// HMG_ERR_NONFINITE if the session's state holds a NaN or an infinity;
// waits for the context stream (vpipes.hip).
int session_check_finite(hmg_session *s);

// hmg_session_create(_dev) with the heights and the water depths always in
// separate planes (for sessions that only run thermal steps on z directly).
int session_create_separate(hmg_ctx                 *ctx,
                            const float             *z,
                            bool                     z_on_dev,
                            int                      nx,
                            int                      ny,
                            const hmg_vpipes_params *p,
                            int                      storage,
                            hmg_session            **out);

int batch_advance(hmg_ctx           *ctx,
                  hmg_session *const *ss,
                  const int          *iters,
                  int                 n,
                  int                 use_graph,
                  int                 chunk_rows,
                  int                 prefetch,
                  float              *ms);
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
                     hipStream_t  st);

} // namespace hmg
