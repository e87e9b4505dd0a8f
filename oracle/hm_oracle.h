/*
 * hm_oracle.h -- CPU oracle for the HighMap grid-erosion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (highmap_amd/, include/)
 * links, loads or calls this code.  It is used by tests/, by the smoke check
 * and by the bench's cpu_baseline leg, always as the checker, never as the
 * thing measured or shipped.
 *
 * Two restatements of the reference algorithm live here:
 *   *_f32 : the FP32 parity contract.  Same operation order as the HIP
 *           kernels, no floating-point contraction (built with
 *           -ffp-contract=off), correctly rounded division and sqrt,
 *           project-written sin/cos.  The GPU must match it bit for bit.
 *   *_f64 : the FP64 error oracle.  The same algorithm with every float
 *           constant of the reference promoted to double.  Used to state the
 *           error of every FP32/BF16/FP16 variant.
 *
 * Reference files followed (read as text, HighMap source tree):
 *   HighMap/src/gpu_opencl/kernels/noise.cl:12-63,194-216,456-504
 *   HighMap/src/gpu_opencl/kernels/_common_rand.cl:12-17,53-57,66-88
 *   HighMap/src/gpu_opencl/kernels/_common_index.cl:8-24,88-132
 *   HighMap/src/gpu_opencl/kernels/_common_math.cl:47-50,126-129
 *   HighMap/src/gpu_opencl/kernels/thermal.cl:5-58
 *   HighMap/src/gpu_opencl/kernels/hydraulic_vpipes.cl:5-211
 *   HighMap/src/erosion/hydraulic_vpipes_gpu.cpp:34-177
 *   HighMap/src/erosion/thermal_gpu.cpp:17-96
 *   HighMap/src/boundary/boundary.cpp:19-41
 *   HighMap/src/range/range.cpp:342-361
 *
 * Parity status of each function is stated in its comment in hm_oracle.c.
 */
#ifndef HM_ORACLE_H
#define HM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of one vpipes run.  Field meaning follows
 * hmap::gpu::hydraulic_vpipes (HighMap/include/highmap/erosion.hpp:1206-1223);
 * gravity and rain_per_iter are extensions (reference values: 1 and 0). */
typedef struct orc_vp_params
{
  float water_height;
  float evap_rate;
  float dt;
  float k_capacity;
  float k_erode;
  float k_depose;
  float k_discharge_exp;
  float downcutting_max_depth_ratio;
  float flux_diffusion_strength;
  float gravity;
  float rain_per_iter;
  int   maintain_water_volume;
  int   flux_diffusion;
  /* Storage of the water, sediment and flux state (FP32 instance only):
   * 0 FP32, 1 BF16, 2 FP16.  Each value of those planes is rounded to the
   * storage type as soon as it is produced (initial water, flow-pass
   * fluxes, water-pass depth, erosion-pass sediment) and every consumer uses
   * the rounded value; arithmetic stays FP32. */
  int   storage;
} orc_vp_params;

float orc_round_bf16(float x);
float orc_round_f16(float x);

/* --- integer seed path and project-written transcendental functions --- */
uint32_t orc_wang_hash(uint32_t seed);
uint32_t orc_xorshift32(uint32_t state);
float    orc_seed_to_fseed(uint32_t seed);
float    orc_sinf(float x);
float    orc_cosf(float x);
float    orc_fractf(float x);
float    orc_hash12f(float px, float py, float fseed);
void     orc_grad22f(float px, float py, float fseed, float *gx, float *gy);
float    orc_base_perlin(float px, float py, float fseed, int per_x, int per_y);
/* Single-octave value of hmap::NoiseType `type` (0 for types the kernels do
 * not implement), and whether the kernels implement it. */
float    orc_base_noise(int type, float px, float py, float fseed, int per_x, int per_y);
int      orc_noise_type_supported(int type);
/* gpu::noise (single octave); returns 1 (zeros written) for types the
 * kernels do not implement. */
int orc_noise_f32(float *out, int nx, int ny, int noise_type, float kx, float ky, uint32_t seed,
                  const float *noise_x, const float *noise_y, const float *bbox4,
                  const int *period2);

/* --- noise_fbm, every implemented type; returns 0, or 1 (zeros written)
 * for types the kernels do not implement --- */
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
                      const int   *period2);

void orc_remap_f32(float *a, size_t n, float vmin, float vmax);

/* --- thermal --- */
void orc_thermal_steps_f32(float       *z,
                           int          nx,
                           int          ny,
                           const float *talus,
                           float        talus_scalar,
                           int          iterations);
void orc_thermal_steps_f64(double       *z,
                           int           nx,
                           int           ny,
                           const double *talus,
                           double        talus_scalar,
                           int           iterations);
void orc_extrapolate_borders_f32(float *z, int nx, int ny);
void orc_extrapolate_borders_f64(double *z, int nx, int ny);

/* thermal_auto_bedrock Jacobi steps (no border extrapolation); talus may be
 * NULL (talus_scalar everywhere). */
void orc_thermal_auto_bedrock_steps_f32(float *z, int nx, int ny, const float *talus,
                                        float talus_scalar, int iterations);
void orc_thermal_auto_bedrock_steps_f64(double *z, int nx, int ny, const double *talus,
                                        double talus_scalar, int iterations);

/* hmap::gpu::sediment_deposition without a mask; z updated in place. */
void orc_sediment_deposition_f32(float *z, int nx, int ny, const float *talus,
                                 float talus_scalar, float *deposition, float max_deposition,
                                 int iterations, int thermal_subiterations);
void orc_sediment_deposition_f64(double *z, int nx, int ny, const double *talus,
                                 double talus_scalar, double *deposition,
                                 float max_deposition, int iterations,
                                 int thermal_subiterations);

/* --- fixed-point water volume (order-independent, exact per term) --- */
double orc_fixsum_f32(const float *a, size_t n);
void   orc_fixsum_parts(const float *a, size_t n, uint64_t *lo, int64_t *hi);
uint64_t orc_xorshift_period(uint32_t start);

/* --- vpipes --- */
/* Initial state: d = water_height * rain, s = fl = fr = ft = fb = u = v = 0.
 * Writes W0 = sum(d) and R = sum(rain) (rain==NULL means all ones). */
void orc_vpipes_init_f32(float         *d,
                         float         *s,
                         float         *fl,
                         float         *fr,
                         float         *ft,
                         float         *fb,
                         float         *u,
                         float         *v,
                         int            nx,
                         int            ny,
                         const orc_vp_params *p,
                         const float   *rain,
                         double        *W0,
                         double        *R);
void orc_vpipes_init_f64(double        *d,
                         double        *s,
                         double        *fl,
                         double        *fr,
                         double        *ft,
                         double        *fb,
                         double        *u,
                         double        *v,
                         int            nx,
                         int            ny,
                         const orc_vp_params *p,
                         const double  *rain,
                         double        *W0,
                         double        *R);
/* Runs `iterations` full reference iterations (flow, water, erosion,
 * transport, evaporation, volume top-up) on the caller-owned state. */
void orc_vpipes_run_f32(float         *z,
                        float         *d,
                        float         *s,
                        float         *fl,
                        float         *fr,
                        float         *ft,
                        float         *fb,
                        float         *u,
                        float         *v,
                        int            nx,
                        int            ny,
                        const orc_vp_params *p,
                        const float   *rain,
                        double         W0,
                        double         R,
                        int            iterations);
void orc_vpipes_run_f64(double        *z,
                        double        *d,
                        double        *s,
                        double        *fl,
                        double        *fr,
                        double        *ft,
                        double        *fb,
                        double        *u,
                        double        *v,
                        int            nx,
                        int            ny,
                        const orc_vp_params *p,
                        const double  *rain,
                        double         W0,
                        double         R,
                        int            iterations);

/* hmap::gpu::flow_simulation: the water depth after `iterations` flow and
 * water passes (and the optional dry-out) into d (nx*ny). */
void orc_flow_simulation_f32(const float *z, int nx, int ny, float water_height,
                             const float *depth_map, int iterations, float dt,
                             int flux_diffusion, float flux_diffusion_strength,
                             float dry_out_ratio, float *d);
void orc_flow_simulation_f64(const double *z, int nx, int ny, float water_height,
                             const double *depth_map, int iterations, float dt,
                             int flux_diffusion, float flux_diffusion_strength,
                             float dry_out_ratio, double *d);

/* exp of the FP32 contract (shared with the kernels). */
float orc_expf(float x);
/* the contract's pow (FP32 only, shared with the kernels), see hm_oracle.c */
float orc_powf(float x, float y);
void  orc_powf_many(const float *x, const float *y, float *out, long n);
/* orc_sinf and orc_cosf elementwise */
void  orc_sincos_many(const float *x, float *s, float *c, long n);

/* hmap::gpu::snow_simulation: the final snow depth into s (nx*ny). */
void orc_snow_simulation_f32(const float *z, int nx, int ny, float snow_depth,
                             const float *fall_map, const float *melting_map,
                             const float *talus, int iterations, float dt,
                             float fall_iterations_ratio, float k_snow, float k_visc,
                             float k_melt_factor, float k_depth_ratio,
                             float k_depth_slope_ratio, int post_filter,
                             float thermal_talus_ratio, float *s);
void orc_snow_simulation_f64(const double *z, int nx, int ny, float snow_depth,
                             const double *fall_map, const double *melting_map,
                             const double *talus, int iterations, float dt,
                             float fall_iterations_ratio, float k_snow, float k_visc,
                             float k_melt_factor, float k_depth_ratio,
                             float k_depth_slope_ratio, int post_filter,
                             float thermal_talus_ratio, double *s);
/* One snow_simulation kernel launch (s -> so, Jacobi), k_melt as passed. */
void orc_snow_step_f32(const float *z, const float *s, const float *talus,
                       const float *melt, float *so, int nx, int ny, float dt,
                       float snow_depth, float k_snow, float k_melt, float k_visc,
                       float k_depth_ratio, float k_depth_slope_ratio);
void orc_snow_step_f64(const double *z, const double *s, const double *talus,
                       const double *melt, double *so, int nx, int ny, double dt,
                       double snow_depth, double k_snow, double k_melt, double k_visc,
                       double k_depth_ratio, double k_depth_slope_ratio);

/* Single-pass entry points, for pass-level tests. */
void orc_vp_flow_f32(const float *z, const float *d1, const float *fl,
                     const float *fr, const float *ft, const float *fb,
                     float *flo, float *fro, float *fto, float *fbo, int nx,
                     int ny, const orc_vp_params *p);
void orc_vp_water_f32(const float *fl, const float *fr, const float *ft,
                      const float *fb, const float *d1, float *d2, float *u,
                      float *v, int nx, int ny, const orc_vp_params *p);
void orc_vp_erosion_f32(const float *z, const float *s, const float *d2,
                        const float *u, const float *v, float *zo, float *so,
                        int nx, int ny, const orc_vp_params *p);
void orc_vp_transport_f32(const float *u, const float *v, const float *s,
                          float *so, int nx, int ny, float dt);
void orc_vp_erosion_f64(const double *z, const double *s, const double *d2,
                        const double *u, const double *v, double *zo, double *so,
                        int nx, int ny, const orc_vp_params *p);

#ifdef __cplusplus
}
#endif

#endif
