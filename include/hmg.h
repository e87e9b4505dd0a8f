/*
 * hmg.h -- C ABI of libhmg, the MI355X (gfx950) implementation of HighMap's
 * grid-erosion hot path: gpu::noise_fbm / gpu::noise, gpu::thermal (and
 * thermal_auto_bedrock, sediment_deposition, snow_simulation),
 * gpu::hydraulic_vpipes and gpu::flow_simulation.
 *
 * This header replaces the OpenCL runtime layer the reference binds through
 * (HighMap/src/gpu_opencl/gpu_opencl.cpp and the CLWrapper `Run` calls in
 * HighMap/src/erosion/hydraulic_vpipes_gpu.cpp:67-159,
 * HighMap/src/erosion/thermal_gpu.cpp:26-67 and
 * HighMap/src/primitives/coherent_noise/primitives_gpu.cpp:558-588).
 * The C++ drop-in (include/hmap_gpu.hpp) keeps the reference's hmap::gpu::
 * signatures on top of it.
 *
 * Conventions
 *  - Arrays are float32, i-fastest, linear index j*nx + i (hmap::Array,
 *    HighMap/include/highmap/array.hpp:236-239).
 *  - Pointers named *_host are host memory, borrowed for the call; the call
 *    is synchronous.  Pointers named *_dev are device memory on the context's
 *    device; those calls are asynchronous on the context stream (or on the
 *    stream given) and return once queued.
 *  - Every entry point returns 0 (HMG_OK) or a negative hmg_status;
 *    hmg_last_error() then holds a message for the calling thread.
 *  - One context may be used from several threads; calls on it serialise.
 *    Separate contexts on one device run concurrently on their own streams.
 */
#ifndef HMG_H
#define HMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HMG_ABI_VERSION 1

typedef enum hmg_status
{
  HMG_OK = 0,
  HMG_ERR_INVALID = -1,     /* bad argument (size, null, range)          */
  HMG_ERR_OOM = -2,         /* device memory exhausted / budget exceeded */
  HMG_ERR_DEVICE = -3,      /* HIP runtime or kernel error               */
  HMG_ERR_NONFINITE = -4,   /* NaN/Inf detected in the state             */
  HMG_ERR_UNSUPPORTED = -5, /* option outside the implemented path       */
  HMG_ERR_NOT_READY = -6    /* asynchronous result not available yet     */
} hmg_status;

/* Storage type of the water, sediment and flux planes in a session.
 * Arithmetic is FP32 in every case; z is always FP32. */
typedef enum hmg_storage
{
  HMG_STORAGE_F32 = 0,
  HMG_STORAGE_BF16 = 1,
  HMG_STORAGE_F16 = 2
} hmg_storage;

/* How vpipes iterations are launched. */
typedef enum hmg_vp_mode
{
  HMG_VP_FUSED = 0, /* one launch per iteration (default)                  */
  HMG_VP_PLAIN = 1  /* the reference's pass structure: flow, water,
                       erosion, transport, evaporation, top-up launches    */
} hmg_vp_mode;

/* Parameters of hydraulic_vpipes.  The first 13 fields and their defaults
 * are those of hmap::gpu::hydraulic_vpipes
 * (HighMap/include/highmap/erosion.hpp:1206-1223):
 *   water_height 1e-2, maintain_water_volume 1, evap_rate 0.1,
 *   iterations 50, dt 0.5, k_capacity 0.5, k_erode 1e-3, k_depose 1e-2,
 *   k_discharge_exp 1, downcutting_max_depth_ratio 10, flux_diffusion 1,
 *   flux_diffusion_strength 0.01.
 * gravity (reference: hard-coded 1, hydraulic_vpipes.cl:31) and
 * rain_per_iter (reference: none; water added before each flow pass as
 * rain_per_iter * rain_map) are extensions.  Use hmg_vpipes_defaults(). */
typedef struct hmg_vpipes_params
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
  int   iterations;
  int   mode; /* hmg_vp_mode */
} hmg_vpipes_params;

typedef struct hmg_ctx     hmg_ctx;
typedef struct hmg_session hmg_session;

/* ------------------------------------------------------------ runtime -- */

/* Creates a context on HIP device `device`.  Replaces
 * hmap::gpu::init_opencl() (HighMap/src/gpu_opencl/gpu_opencl.cpp:32-134,
 * decl. HighMap/include/highmap/opencl/gpu_opencl.hpp:30), which returns
 * false when no device is ready; here that is HMG_ERR_DEVICE. */
int hmg_init(int device, hmg_ctx **out);
int hmg_destroy(hmg_ctx *ctx);
/* Thread-local message of the last failing call ("" if none). */
const char *hmg_last_error(void);
int         hmg_abi_version(void);
/* Fills defaults of hmap::gpu::hydraulic_vpipes plus gravity 1,
 * rain_per_iter 0, mode HMG_VP_FUSED. */
void hmg_vpipes_defaults(hmg_vpipes_params *p);
/* Waits for all work queued on the context stream. */
int hmg_sync(hmg_ctx *ctx);
/* The HIP stream (hipStream_t) of the context, for callers that enqueue
 * their own work around the *_dev entry points. */
void *hmg_stream(hmg_ctx *ctx);

/* -------------------------------------------------------------- noise -- */

/* hmap::gpu::noise_fbm (HighMap/src/primitives/coherent_noise/
 * primitives_gpu.cpp:543-591; defaults coherent_noise.hpp:821-834:
 * octaves 8, weight 0.7, persistence 0.5, lacunarity 2, bbox {0,1,0,1},
 * period {0,0}).  noise_type follows hmap::NoiseType
 * (HighMap/include/highmap/functions.hpp:66-81): PERLIN 1, PERLIN_BILLOW 2,
 * PERLIN_HALF 3, SIMPLEX2 4, VALUE 6, VALUE_CUBIC 7, VALUE_LINEAR 9 and
 * WORLEY 10 are computed (noise.cl:28-383); the reference kernel writes 0 for
 * PARBERRY 0, SIMPLEX2S 5, VALUE_DELAUNAY 8, WORLEY_DOUBLE 11 and
 * WORLEY_VALUE 12, and so does this call; other values are
 * HMG_ERR_INVALID.  ctrl, noise_x, noise_y may be NULL; bbox and period may
 * be NULL for the defaults. */
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
                  const int   *period2);
/* Same on device buffers, asynchronous on the context stream. */
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
                      const int   *period2);

/* hmap::gpu::noise, one octave (HighMap/src/primitives/coherent_noise/
 * primitives_gpu.cpp:504-541, kernel noise.cl:385-454; defaults
 * coherent_noise.hpp:804-812: bbox {0,1,0,1}, period {0,0}); types as for
 * hmg_noise_fbm. */
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
              const int   *period2);
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
                  const int   *period2);

/* hmap::remap(array, vmin, vmax) (HighMap/src/range/range.cpp:342-361) on a
 * device buffer; min and max are reduced on the device. */
int hmg_remap_dev(hmg_ctx *ctx, float *a_dev, size_t n, float vmin, float vmax);

/* ------------------------------------------------------------ thermal -- */

/* hmap::gpu::thermal (HighMap/src/erosion/thermal_gpu.cpp:17-96; defaults
 * HighMap/include/highmap/erosion.hpp:1784-1801, iterations 10).
 * talus_host == NULL means the scalar overload (talus_scalar everywhere).
 * bedrock_host selects the thermal_with_bedrock kernel
 * (thermal.cl:60-100).  deposition_host, if not NULL, receives
 * max(z - z_before, 0).  mask_host, if not NULL, blends the result as
 * apply_with_mask does (HighMap/include/highmap/operator.hpp:405-418).
 * Semantics: Jacobi steps (each step reads step t, writes t+1), then
 * extrapolate_borders (HighMap/src/boundary/boundary.cpp:19-41).
 * nx and ny must be >= 3. */
int hmg_thermal(hmg_ctx     *ctx,
                float       *z_host,
                int          nx,
                int          ny,
                const float *talus_host,
                float        talus_scalar,
                int          iterations,
                const float *bedrock_host,
                float       *deposition_host,
                const float *mask_host);
/* Device form: `iterations` Jacobi steps then extrapolate_borders, in place
 * on z_dev.  The second Jacobi plane comes from the context's buffer pool
 * and goes back to it when the call returns, so the call waits for its work
 * on the context stream.  talus_dev / bedrock_dev may be NULL. */
int hmg_thermal_dev(hmg_ctx     *ctx,
                    float       *z_dev,
                    int          nx,
                    int          ny,
                    const float *talus_dev,
                    float        talus_scalar,
                    int          iterations,
                    const float *bedrock_dev);
/* Kernel choice of the uniform-talus Jacobi steps: 8 (the default) runs
 * grids below 2^18 cells 8 steps per launch on LDS tiles and larger grids
 * with the streaming kernels, two steps per launch while two are left (the
 * last odd step alone); 2 uses the streaming kernels that way at every
 * size; 0 the one-step streaming kernel at every size.  Results do not
 * depend on it.  Process-wide; HMG_ERR_INVALID for other values. */
int hmg_thermal_set_block_steps(int steps);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same
 * computation as hmg_thermal_dev, with the second Jacobi plane in a static
 * device array of 4096 x 4096 floats (64 MiB) that each GPU holds for the
 * life of the process; it calls no hipMalloc/hipFree and allocates nothing.
 * Grids above 4096 x 4096 cells are HMG_ERR_INVALID.  Asynchronous on the
 * context stream.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (from different contexts or streams) race and give
 * wrong results; calls on one context are ordered by its stream. */
int hmg_thermal_dev_noalloc(hmg_ctx     *ctx,
                            float       *z_dev,
                            int          nx,
                            int          ny,
                            const float *talus_dev,
                            float        talus_scalar,
                            int          iterations,
                            const float *bedrock_dev);

/* hmap::gpu::thermal_auto_bedrock, the three overloads
 * (HighMap/src/erosion/thermal_gpu.cpp:98-146, kernel thermal.cl:102-145;
 * defaults iterations 10, include/highmap/erosion.hpp:1824-1838): Jacobi
 * steps in which each interior cell keeps a bedrock that starts at
 * -FLT_MAX and becomes max(z0, z) after each update, z = max(max(z + exchange,
 * z0), bedrock); then extrapolate_borders.  talus_host NULL means
 * talus_scalar everywhere; deposition (max(z - z0, 0)) and mask (the
 * p_mask overload) as in hmg_thermal. */
int hmg_thermal_auto_bedrock(hmg_ctx     *ctx,
                             float       *z_host,
                             int          nx,
                             int          ny,
                             const float *talus_host,
                             float        talus_scalar,
                             int          iterations,
                             float       *deposition_host,
                             const float *mask_host);
/* Device form, in place on z_dev (no deposition or mask). */
int hmg_thermal_auto_bedrock_dev(hmg_ctx     *ctx,
                                 float       *z_dev,
                                 int          nx,
                                 int          ny,
                                 const float *talus_dev,
                                 float        talus_scalar,
                                 int          iterations);

/* hmap::gpu::sediment_deposition, both overloads
 * (HighMap/src/erosion/deposition.cpp:19-63; defaults max_deposition 0.01,
 * iterations 5, thermal_subiterations 10, include/highmap/erosion.hpp:
 * 1366-1379): `iterations` rounds of raising a sediment layer by
 * max_deposition / iterations and letting it settle with thermal(z + layer,
 * talus, thermal_subiterations); z += layer.  deposition_host (may be NULL)
 * receives the layer; mask_host (may be NULL) blends z = lerp(z, result,
 * mask). */
int hmg_sediment_deposition(hmg_ctx     *ctx,
                            float       *z_host,
                            int          nx,
                            int          ny,
                            const float *talus_host,
                            float        talus_scalar,
                            float       *deposition_host,
                            float        max_deposition,
                            int          iterations,
                            int          thermal_subiterations,
                            const float *mask_host);

/* ------------------------------------------------------------- vpipes -- */

/* hmap::gpu::hydraulic_vpipes (HighMap/src/erosion/
 * hydraulic_vpipes_gpu.cpp:15-178).  z is updated in place; rain_host may be
 * NULL (uniform 1); d/s/u/v outputs may be NULL.  p == NULL means the
 * reference defaults. */
int hmg_vpipes(hmg_ctx                 *ctx,
               float                   *z_host,
               int                      nx,
               int                      ny,
               const hmg_vpipes_params *p,
               const float             *rain_host,
               float                   *d_out_host,
               float                   *s_out_host,
               float                   *u_out_host,
               float                   *v_out_host);

/* ---------------------------------------------------- flow simulation -- */

/* hmap::gpu::flow_simulation (HighMap/src/hydrology/flow_simulation.cpp:
 * 26-110; defaults in include/highmap/hydrology/hydrology.hpp:875-882:
 * dt 0.5, flux_diffusion 1, flux_diffusion_strength 0.01, dry_out_ratio 0).
 * d = water_height * depth_map, then `iterations` flow + water passes of
 * hydraulic_vpipes (no erosion, evaporation or top-up), then
 * water_depth_dry_out(d, dry_out_ratio) when dry_out_ratio != 0.  Writes the
 * water depth (nx*ny floats) to d_out_host; z and depth_map are not
 * changed.  At most 2^28 cells. */
int hmg_flow_simulation(hmg_ctx     *ctx,
                        const float *z_host,
                        int          nx,
                        int          ny,
                        float        water_height,
                        const float *depth_map_host,
                        int          iterations,
                        float        dt,
                        int          flux_diffusion,
                        float        flux_diffusion_strength,
                        float        dry_out_ratio,
                        float       *d_out_host);
/* The same on device pointers (d_out_dev may not alias the inputs). */
int hmg_flow_simulation_dev(hmg_ctx     *ctx,
                            const float *z_dev,
                            int          nx,
                            int          ny,
                            float        water_height,
                            const float *depth_map_dev,
                            int          iterations,
                            float        dt,
                            int          flux_diffusion,
                            float        flux_diffusion_strength,
                            float        dry_out_ratio,
                            float       *d_out_dev);

/* ---------------------------------------------------- snow simulation -- */

/* hmap::gpu::snow_simulation (HighMap/src/hydrology/snow_simulation.cpp:
 * 81-154, kernel HighMap/src/gpu_opencl/kernels/snow_simulation.cl; defaults
 * in include/highmap/hydrology/hydrology.hpp:942-956: dt 0.5,
 * fall_iterations_ratio 1, k_snow 0.5, k_visc 0.1, k_melt_factor 0.8,
 * k_depth_ratio 1, k_depth_slope_ratio 1, post_filter 1,
 * thermal_talus_ratio 0.2).  fall_map, melting_map and talus are nx*ny
 * planes.  Writes the snow depth (nx*ny floats) to s_out_host; the inputs
 * are not changed.  nx, ny >= 3, iterations >= 0. */
int hmg_snow_simulation(hmg_ctx     *ctx,
                        const float *z_host,
                        int          nx,
                        int          ny,
                        float        snow_depth,
                        const float *fall_map_host,
                        const float *melting_map_host,
                        const float *talus_host,
                        int          iterations,
                        float        dt,
                        float        fall_iterations_ratio,
                        float        k_snow,
                        float        k_visc,
                        float        k_melt_factor,
                        float        k_depth_ratio,
                        float        k_depth_slope_ratio,
                        int          post_filter,
                        float        thermal_talus_ratio,
                        float       *s_out_host);
/* The same on device pointers (s_out_dev may not alias the inputs). */
int hmg_snow_simulation_dev(hmg_ctx     *ctx,
                            const float *z_dev,
                            int          nx,
                            int          ny,
                            float        snow_depth,
                            const float *fall_map_dev,
                            const float *melting_map_dev,
                            const float *talus_dev,
                            int          iterations,
                            float        dt,
                            float        fall_iterations_ratio,
                            float        k_snow,
                            float        k_visc,
                            float        k_melt_factor,
                            float        k_depth_ratio,
                            float        k_depth_slope_ratio,
                            int          post_filter,
                            float        thermal_talus_ratio,
                            float       *s_out_dev);

/* ------------------------------------------------------------ session -- */

/* A session keeps one erosion state (z, water, sediment, fluxes) resident
 * in device memory between calls, so that hydraulic iterations and thermal
 * steps can be interleaved on one continuous state without host transfers.
 * The water, sediment and flux planes are stored as `storage`; all
 * arithmetic is FP32; the water-volume reduction is exact fixed point.
 * Grids of more than 2^28 cells run the plain passes whatever mode the
 * parameters ask for (the fused kernel addresses its planes with 32-bit
 * offsets); results do not depend on the mode. */
int hmg_session_create(hmg_ctx                 *ctx,
                       const float             *z_host,
                       int                      nx,
                       int                      ny,
                       const hmg_vpipes_params *p,
                       const float             *rain_host,
                       int                      storage,
                       hmg_session            **out);
/* Same, with z and rain already on the device (copied into the session). */
int hmg_session_create_dev(hmg_ctx                 *ctx,
                           const float             *z_dev,
                           int                      nx,
                           int                      ny,
                           const hmg_vpipes_params *p,
                           const float             *rain_dev,
                           int                      storage,
                           hmg_session            **out);
/* Queues `iterations` hydraulic iterations (asynchronous). */
int hmg_session_vpipes(hmg_session *s, int iterations);
/* Queues one thermal() call on the session's z: `iterations` Jacobi steps
 * with uniform talus, then extrapolate_borders (asynchronous). */
int hmg_session_thermal(hmg_session *s, float talus, int iterations);
/* Waits, then copies the state out, in the form hydraulic_vpipes returns it
 * (d after evaporation and top-up, s after transport, u and v of the last
 * iteration).  Any pointer may be NULL. */
int hmg_session_export(hmg_session *s,
                       float       *z_host,
                       float       *d_host,
                       float       *s_host,
                       float       *u_host,
                       float       *v_host);
/* Waits, then returns sum(z), sum(s), sum(d) in FP64 (mass check). */
int hmg_session_sums(hmg_session *s, double *sum_z, double *sum_s, double *sum_d);
/* Waits, then returns max over the cells of dt * max(|u|, |v|) for the
 * state's velocities (those hmg_session_export returns), in FP64: how far, in
 * cells, the next iteration's sediment backtrace (the semi-Lagrangian
 * transport, hydraulic_vpipes.cl:204-210) reaches; the reference computes no
 * such bound (its dt is fixed).  NaN if a velocity is NaN. */
int hmg_session_max_backtrace(hmg_session *s, double *max_backtrace);
/* Launch shape of the fused kernel: rows each wavefront streams through
 * and the variant, `prefetch`: 0 = each row's loads issued when the row is
 * needed, 1 = one row ahead, 2 = two rows ahead with each row's sediment
 * gathers issued one row ahead.  Results do not depend on either.  By
 * default a fused session of at least 2^20 cells tunes the shape online:
 * from its 20th hydraulic iteration on, a warm-up and two rounds over the
 * candidate shapes (those that give every SIMD at least one wavefront, and
 * per variant the chunk length that fills the kernel's wavefront slots)
 * each run one candidate between two events, the next one keeps the
 * fastest (a process-wide cache per storage, shape, rain plane and exponent
 * skips the probes for later sessions; a session that ends after a full
 * round caches its best so far).  Setting the shape ends the tuning. */
int hmg_session_set_tuning(hmg_session *s, int chunk_rows, int prefetch);
/* The current shape; *tuning is 1 while the online probes still run. */
int hmg_session_get_tuning(hmg_session *s, int *chunk_rows, int *prefetch, int *tuning);
/* Device pointer of the session's current heights (FP32, nx*ny).  Sessions
 * of FP32 state in the fused form keep the heights beside the water depths
 * in one float2 plane; for them this is a copy of the heights, made on the
 * context stream, valid until the session's next call (null if the copy
 * cannot be made). */
float *hmg_session_z_dev(hmg_session *s);
int    hmg_session_sync(hmg_session *s);
int    hmg_session_destroy(hmg_session *s);
/* Device bytes a session of this shape and storage holds. */
size_t hmg_session_bytes(int nx, int ny, int storage);

/* -------------------------------------------------------------- batch -- */

/* A batch runs many independent maps (any shapes, each with its own
 * iteration count) together: every launch advances all maps still running
 * by one iteration, so small maps share the GPU instead of each paying a
 * launch per pass.  Each map's result is identical to running it alone.
 * No reference counterpart (the reference processes one Array per call);
 * the per-map semantics are hmg_vpipes'. */
typedef struct hmg_batch hmg_batch;
int hmg_batch_create(hmg_ctx *ctx, int storage, const hmg_vpipes_params *p, hmg_batch **out);
/* Adds a map (host arrays; rain may be NULL, at most 2^28 cells) to run
 * `iterations` iterations; *map_id receives its index. */
int hmg_batch_add(hmg_batch   *b,
                  const float *z_host,
                  int          nx,
                  int          ny,
                  const float *rain_host,
                  int          iterations,
                  int         *map_id);
/* Launch shape of the batched kernel: rows per wavefront and variant as in
 * hmg_session_set_tuning (0, 1, or 2 = gathers ahead). */
int hmg_batch_set_tuning(hmg_batch *b, int chunk_rows, int prefetch);
/* Picks the launch shape by timing `probe` iterations of each candidate on
 * the live maps (they count toward each map's iterations; results do not
 * depend on the shape), then the form (graph or persistent, `probe`
 * iterations each with that shape); caches the choice per storage type and
 * batch size.  Candidates: chunk lengths 16-96 for each variant, plus the
 * lengths whose work items fill one, two or three rounds of the variant's
 * wavefront slots.  When every map has 20 iterations to spare beyond the
 * probes, 20 ordinary iterations run first. */
int hmg_batch_autotune(hmg_batch *b, int probe, int *chunk_rows_out, int *prefetch_out);
/* Forms of hmg_batch_run. */
#define HMG_BATCH_STREAM 0     /* one launch per iteration on the stream */
#define HMG_BATCH_GRAPH 1      /* the same launches captured in one hipGraph */
#define HMG_BATCH_PERSISTENT 2 /* one cooperative launch for all iterations: a
                                  grid no larger than the device holds at once,
                                  every wavefront looping over the items of
                                  each iteration, a grid-wide barrier between
                                  iterations (the sediment gathers reach any
                                  distance, so an iteration needs all of the
                                  previous one); no per-row workgroup barrier */
#define HMG_BATCH_AUTO 3       /* the form hmg_batch_autotune picked (graph
                                  until it has run) */
/* Runs every map to its iteration count in the given form (above; for
 * compatibility any other nonzero value means HMG_BATCH_GRAPH).
 * *device_ms (may be NULL) receives the elapsed device time of the
 * launches.  A persistent run the runtime refuses (observed when another
 * process on the device has made cooperative launches) returns
 * HMG_ERR_UNSUPPORTED with the maps unchanged; one whose barrier wait gave
 * up (seconds without progress) returns HMG_ERR_DEVICE and leaves the maps'
 * state invalid. */
int hmg_batch_run(hmg_batch *b, int use_graph, float *device_ms);
/* The form hmg_batch_autotune picked (HMG_BATCH_GRAPH or
 * HMG_BATCH_PERSISTENT) into *form. */
int hmg_batch_tuned_form(hmg_batch *b, int *form);
/* Copies one map's state out as hmg_session_export does. */
int hmg_batch_export(hmg_batch *b, int map_id, float *z, float *d, float *s, float *u, float *v);
int hmg_batch_destroy(hmg_batch *b);

/* ------------------------------------------------------------- server -- */

/* A continuous-batching erosion server on one device: jobs are admitted in
 * arrival order while their state fits in an HBM budget, keep their state
 * resident on the device, and advance together (batched vpipes launches,
 * then due thermal steps) tick by tick; each finished job's z is copied to
 * the caller's buffer.  A job's result equals the same job run alone.  The
 * batched launches' shape is tuned online on the first ticks (device time
 * per cell-iteration of six candidate shapes, two ticks each).
 * No reference counterpart (HighMap runs one call at a time). */
typedef struct hmg_server hmg_server;

enum
{
  HMG_JOB_VPIPES_THERMAL = 0, /* hydraulic iterations, a thermal step (1 Jacobi
                                 step + borders) after every thermal_every-th */
  HMG_JOB_THERMAL = 1         /* thermal(z, talus, iterations) only           */
};
enum
{
  HMG_JOB_QUEUED = 1,
  HMG_JOB_RUNNING = 2,
  HMG_JOB_DONE = 3 /* a negative state is the hmg_status of a failed job */
};

typedef struct hmg_job_desc
{
  int      nx, ny;
  int      kind;          /* HMG_JOB_*                                      */
  int      iterations;    /* hydraulic iterations, or thermal steps          */
  int      thermal_every; /* HMG_JOB_VPIPES_THERMAL: 10 in the served mix    */
  float    talus;         /* thermal talus                                   */
  int      has_input;     /* 1: z given at submit; 0: generated on the device
                             as remap(noise_fbm(PERLIN, kw 4, seed)) in [0,1] */
  uint32_t seed;
} hmg_job_desc;

/* hbm_budget 0 means 80% of the device memory free at creation. */
int hmg_server_create(hmg_ctx                 *ctx,
                      size_t                   hbm_budget,
                      int                      storage,
                      const hmg_vpipes_params *p,
                      hmg_server             **out);
/* Queues a job (at least 3 x 3, at most 2^28 cells); z_out_host (nx*ny
 * floats) must stay valid until the job is done.  Returns immediately. */
int hmg_server_submit(hmg_server         *sv,
                      const hmg_job_desc *d,
                      const float        *z_host,
                      float              *z_out_host,
                      long long          *job_id);
int hmg_server_poll(hmg_server *sv, long long job_id, int *state);
/* steady-clock seconds of submission, admission and result delivery */
int hmg_server_job_times(hmg_server *sv,
                         long long   job_id,
                         double     *t_submit,
                         double     *t_admit,
                         double     *t_done);
/* Caps the cells (sum of nx*ny) of the jobs that run at once; admission
 * waits while the cap is reached (a job larger than the cap still runs when
 * nothing else does).  0, the default, means no cap.  Under overload a cap
 * of a few full-device tiles keeps the device saturated while later jobs
 * wait in arrival order instead of all sharing it. */
int hmg_server_set_max_active_cells(hmg_server *sv, size_t cells);
int hmg_server_wait_all(hmg_server *sv);
int hmg_server_stats(hmg_server *sv,
                     double     *cell_iters,
                     long long  *ticks,
                     double     *busy_s,
                     size_t     *budget,
                     size_t     *used);
/* Finishes every queued job, then stops the server. */
int hmg_server_destroy(hmg_server *sv);

#ifdef __cplusplus
}
#endif

#endif
