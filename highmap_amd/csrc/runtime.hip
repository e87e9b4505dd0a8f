// runtime.hip -- context, errors, device memory pool, parameter checks.
//
// Replaces the OpenCL runtime of the reference: hmap::gpu::init_opencl
// (HighMap/src/gpu_opencl/gpu_opencl.cpp:32-134) and the CLWrapper
// DeviceManager/KernelManager singletons it drives.  There is no run-time
// kernel compilation: the kernels are gfx950 code objects in this library.
#include <cmath>
#include <cstring>

#include "hmg_internal.hpp"

namespace hmg
{

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error()
{
  g_err[0] = 0;
}

Pool::~Pool()
{
  trim();
  for (auto &kv : live_) (void)hipFree(kv.first);
}

static size_t round_bytes(size_t n)
{
  const size_t g = (size_t)1 << 16; // 64 KiB granularity
  return (n + g - 1) / g * g;
}

int Pool::alloc(void **p, size_t bytes)
{
  size_t                      n = round_bytes(bytes ? bytes : 1);
  std::lock_guard<std::mutex> lk(mu_);
  auto                        it = free_.lower_bound(n);
  if (it != free_.end() && it->first <= n + n / 4)
  {
    *p = it->second;
    live_[*p] = it->first;
    free_.erase(it);
    return HMG_OK;
  }
  void      *q = nullptr;
  hipError_t e = hipMalloc(&q, n);
  if (e != hipSuccess)
  {
    // release cached blocks and retry once
    for (auto &kv : free_)
    {
      (void)hipFree(kv.second);
      held_ -= kv.first;
    }
    free_.clear();
    (void)hipGetLastError();
    e = hipMalloc(&q, n);
  }
  if (e != hipSuccess)
  {
    (void)hipGetLastError();
    return set_error(HMG_ERR_OOM, "device allocation of %zu bytes failed: %s", n,
                     hipGetErrorString(e));
  }
  held_ += n;
  live_[q] = n;
  *p = q;
  return HMG_OK;
}

void Pool::release(void *p)
{
  std::lock_guard<std::mutex> lk(mu_);
  auto                        it = live_.find(p);
  if (it == live_.end()) return;
  free_.emplace(it->second, p);
  live_.erase(it);
}

void Pool::trim()
{
  std::lock_guard<std::mutex> lk(mu_);
  for (auto &kv : free_)
  {
    (void)hipFree(kv.second);
    held_ -= kv.first;
  }
  free_.clear();
}

int ctx_enter(hmg_ctx *ctx)
{
  if (!ctx) return set_error(HMG_ERR_INVALID, "null context");
  HMG_HIP(hipSetDevice(ctx->device));
  return HMG_OK;
}

VpConst make_vp_const(const hmg_vpipes_params &p, int nx, int ny)
{
  VpConst k;
  k.nx = nx;
  k.ny = ny;
  k.dt = p.dt;
  k.dtg = p.dt * p.gravity;
  k.c0 = p.flux_diffusion_strength;
  k.c1 = 1.f - 4.f * p.flux_diffusion_strength;
  k.diffuse = p.flux_diffusion ? 1 : 0;
  k.dmin = 0.001f * p.water_height / p.dt;
  k.fnx = (float)nx;
  k.dmax = p.water_height;
  k.kc = p.k_capacity;
  k.ke = p.k_erode;
  k.kd = p.k_depose;
  k.kexp = p.k_discharge_exp;
  k.ratio = p.downcutting_max_depth_ratio;
  k.ce = 1.f - p.dt * p.evap_rate;
  k.rr = p.rain_per_iter;
  k.maintain = p.maintain_water_volume ? 1 : 0;
  k.pow2x = (nx & (nx - 1)) == 0;
  k.pow2y = (ny & (ny - 1)) == 0;
  k.inv_nx = k.pow2x ? 1.f / (float)nx : 0.f;
  k.inv_ny = k.pow2y ? 1.f / (float)ny : 0.f;
  return k;
}

int validate_vp_params(const hmg_vpipes_params &p)
{
  const float f[] = {p.water_height, p.evap_rate, p.dt, p.k_capacity, p.k_erode,
                     p.k_depose, p.k_discharge_exp, p.downcutting_max_depth_ratio,
                     p.flux_diffusion_strength, p.gravity, p.rain_per_iter};
  for (float x : f)
    if (!std::isfinite(x)) return set_error(HMG_ERR_INVALID, "non-finite vpipes parameter");
  if (!(p.dt > 0.f)) return set_error(HMG_ERR_INVALID, "dt must be > 0");
  if (!(p.water_height > 0.f))
    return set_error(HMG_ERR_INVALID, "water_height must be > 0");
  if (p.iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  if (p.mode != HMG_VP_FUSED && p.mode != HMG_VP_PLAIN)
    return set_error(HMG_ERR_INVALID, "unknown vpipes mode %d", p.mode);
  return HMG_OK;
}

} // namespace hmg

using namespace hmg;

extern "C" {

const char *hmg_last_error(void)
{
  return g_err;
}

int hmg_abi_version(void)
{
  return HMG_ABI_VERSION;
}

void hmg_vpipes_defaults(hmg_vpipes_params *p)
{
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  // hmap::gpu::hydraulic_vpipes defaults (erosion.hpp:1206-1223)
  p->water_height = 1e-2f;
  p->maintain_water_volume = 1;
  p->evap_rate = 0.1f;
  p->iterations = 50;
  p->dt = 0.5f;
  p->k_capacity = 0.5f;
  p->k_erode = 1e-3f;
  p->k_depose = 1e-2f;
  p->k_discharge_exp = 1.f;
  p->downcutting_max_depth_ratio = 10.f;
  p->flux_diffusion = 1;
  p->flux_diffusion_strength = 0.01f;
  // extensions
  p->gravity = 1.f;
  p->rain_per_iter = 0.f;
  p->mode = HMG_VP_FUSED;
}

int hmg_init(int device, hmg_ctx **out)
{
  clear_error();
  if (!out) return set_error(HMG_ERR_INVALID, "null output pointer");
  *out = nullptr;
  int        n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
  {
    (void)hipGetLastError();
    return set_error(HMG_ERR_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= n)
    return set_error(HMG_ERR_INVALID, "device %d out of range [0, %d)", device, n);
  HMG_HIP(hipSetDevice(device));
  HMG_TRY(noalloc_scratch_setup(device));
  hipDeviceProp_t prop;
  HMG_HIP(hipGetDeviceProperties(&prop, device));
  hmg_ctx *c = new hmg_ctx();
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  c->wave_size = prop.warpSize;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess)
  {
    delete c;
    return set_error(HMG_ERR_DEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
  }
  c->pool = new Pool(device);
  *out = c;
  return HMG_OK;
}

int hmg_destroy(hmg_ctx *ctx)
{
  if (!ctx) return HMG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx->pool;
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HMG_OK;
}

int hmg_sync(hmg_ctx *ctx)
{
  HMG_TRY(ctx_enter(ctx));
  HMG_HIP(hipStreamSynchronize(ctx->stream));
  return HMG_OK;
}

void *hmg_stream(hmg_ctx *ctx)
{
  return ctx ? (void *)ctx->stream : nullptr;
}

} // extern "C"
