/*
 * cpu_bench.c -- times the FP32 CPU restatement (hm_oracle.c) as the host
 * CPU baseline.  TEST/BENCH INFRASTRUCTURE ONLY: built with the reference's
 * host flags (-O3 -Ofast -ffast-math) and OpenMP over rows, and run as a
 * child process by the bench's cpu_baseline leg.
 *
 * usage: cpu_bench <workload> <nx> <ny> <iterations> <threads> <reps>
 *   workload: vpipes_thermal  vpipes with one thermal step after every 10th
 *                             hydraulic iteration (talus 1/nx)
 *             vpipes          vpipes only
 *             thermal         thermal only (talus 2/nx)
 *             flow            flow_simulation (water_height 1e-2, uniform
 *                             depth map, dt 0.5, flux diffusion 0.01)
 *             snow            snow_simulation as ex_snow_simulation (snow
 *                             0.1, uniform fall, no melting, talus 2.5/nx,
 *                             defaults, post-filter on)
 * The water-volume sum of each iteration is a serial FP32 accumulate, as the
 * reference's host loop does it (methods.cpp:428-431).
 * Input: remap(noise_fbm(PERLIN, kw 4, seed 2)) to [0,1] (not timed).
 * Output: one JSON line with the median wall time of `reps` timed runs after
 * one warm-up run.
 */
#include "hm_oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int cmp_d(const void *a, const void *b)
{
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
  if (argc < 7)
  {
    fprintf(stderr, "usage: %s workload nx ny iterations threads reps\n", argv[0]);
    return 2;
  }
  const char *wl = argv[1];
  int         nx = atoi(argv[2]), ny = atoi(argv[3]);
  int         iters = atoi(argv[4]), threads = atoi(argv[5]), reps = atoi(argv[6]);
  if (nx < 3 || ny < 3 || iters < 1 || threads < 1 || reps < 1) return 2;
  omp_set_num_threads(threads);

  size_t n = (size_t)nx * ny;
  float *z0 = malloc(n * sizeof(float));
  float  bbox[4] = {0.f, 1.f, 0.f, 1.f};
  int    per[2] = {0, 0};
  orc_noise_fbm_f32(z0, nx, ny, 1, 4.f, 4.f, 2u, 8, 0.7f, 0.5f, 2.f, NULL,
                    NULL, NULL, bbox, per);
  orc_remap_f32(z0, n, 0.f, 1.f);

  orc_vp_params p;
  memset(&p, 0, sizeof(p));
  p.water_height = 1e-3f;
  p.evap_rate = 0.01f;
  p.dt = 0.5f;
  p.k_capacity = 0.1f;
  p.k_erode = 0.05f;
  p.k_depose = 0.05f;
  p.k_discharge_exp = 1.f;
  p.downcutting_max_depth_ratio = 10.f;
  p.flux_diffusion_strength = 0.01f;
  p.gravity = 1.f;
  p.maintain_water_volume = 1;
  p.flux_diffusion = 1;

  float *pl = malloc(9 * n * sizeof(float));
  float *z = pl, *d = pl + n, *s = pl + 2 * n, *fl = pl + 3 * n, *fr = pl + 4 * n;
  float *ft = pl + 5 * n, *fb = pl + 6 * n, *u = pl + 7 * n, *v = pl + 8 * n;

  double *times = malloc(sizeof(double) * (size_t)reps);
  for (int r = -1; r < reps; r++)
  {
    memcpy(z, z0, n * sizeof(float));
    double W0, R;
    orc_vpipes_init_f32(d, s, fl, fr, ft, fb, u, v, nx, ny, &p, NULL, &W0, &R);
    double t0 = now_s();
    if (!strcmp(wl, "flow"))
    {
      for (size_t k = 0; k < n; k++) u[k] = 1.f; /* depth map */
      orc_flow_simulation_f32(z, nx, ny, 1e-2f, u, iters, 0.5f, 1, 0.01f, 0.f, d);
      z[n / 2] = d[n / 2]; /* checksum below */
    }
    else if (!strcmp(wl, "snow"))
    {
      for (size_t k = 0; k < n; k++)
      {
        u[k] = 1.f;              /* fall map */
        v[k] = 0.f;              /* melting map */
        fl[k] = 2.5f / (float)nx; /* talus */
      }
      orc_snow_simulation_f32(z, nx, ny, 0.1f, u, v, fl, iters, 0.5f, 1.f, 0.5f, 0.1f, 0.8f,
                              1.f, 1.f, 1, 0.2f, d);
      z[n / 2] = d[n / 2]; /* checksum below */
    }
    else if (!strcmp(wl, "thermal"))
    {
      orc_thermal_steps_f32(z, nx, ny, NULL, 2.f / (float)nx, iters);
      orc_extrapolate_borders_f32(z, nx, ny);
    }
    else
    {
      int with_thermal = !strcmp(wl, "vpipes_thermal");
      for (int done = 0; done < iters;)
      {
        int chunk = with_thermal ? (iters - done < 10 ? iters - done : 10)
                                 : iters - done;
        orc_vpipes_run_f32(z, d, s, fl, fr, ft, fb, u, v, nx, ny, &p, NULL,
                           W0, R, chunk);
        done += chunk;
        if (with_thermal && chunk == 10)
        {
          orc_thermal_steps_f32(z, nx, ny, NULL, 1.f / (float)nx, 1);
          orc_extrapolate_borders_f32(z, nx, ny);
        }
      }
    }
    double t1 = now_s();
    if (r >= 0) times[r] = t1 - t0;
  }
  qsort(times, (size_t)reps, sizeof(double), cmp_d);
  double med = times[reps / 2];
  double gci = (double)n * (double)iters / med / 1e9;
  printf("{\"workload\": \"%s\", \"nx\": %d, \"ny\": %d, \"iterations\": %d, "
         "\"threads\": %d, \"reps\": %d, \"median_s\": %.6f, "
         "\"gcell_iter_s\": %.6f, \"checksum\": %.9g}\n",
         wl, nx, ny, iters, threads, reps, med, gci, (double)z[n / 2]);
  free(times);
  free(pl);
  free(z0);
  return 0;
}
