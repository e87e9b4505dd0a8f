// server.hip -- continuous-batching erosion server on one device.
//
// Requests (erosion jobs of any grid size and iteration count) are queued on
// submit and admitted in arrival order while their device state fits in an
// HBM budget.  Admitted jobs keep their state resident in device memory and
// advance together: each tick runs up to 10 hydraulic iterations of every
// vpipes job as batched launches (one launch per iteration for all jobs),
// the thermal step those jobs are due, and up to 8 Jacobi steps of every
// thermal-only job.  Jobs join at the next tick after they are admitted and
// leave as soon as they finish, when their result is copied to the caller's
// host buffer and their memory returns to the budget.
//
// A job's result depends only on its own state: it is bit-identical to the
// same job run alone (hmg_session_* / hmg_thermal_dev), whatever else was
// running.  No reference counterpart: HighMap runs one call at a time.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>

#include "hmg_internal.hpp"

namespace
{

double now_s()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Job
{
  long long          id = 0;
  hmg_job_desc       desc{};
  std::vector<float> z_in; // host copy when the caller passed z
  float             *out = nullptr;
  double             t_submit = 0, t_admit = 0, t_done = 0;
  int                state = HMG_JOB_QUEUED;
  hmg_session       *s = nullptr;
  int                done = 0; // iterations (vpipes) or steps (thermal) done
  size_t             bytes = 0;
};

} // namespace

struct hmg_server
{
  hmg_ctx                      *ctx = nullptr;
  int                           storage = HMG_STORAGE_F32;
  hmg_vpipes_params             p{};
  size_t                        budget = 0, used = 0;
  size_t                        max_cells = 0, active_cells = 0; // 0: no cap
  std::mutex                    mu;
  std::condition_variable       cv, cv_done;
  std::deque<Job *>             queue;
  std::vector<Job *>            active;
  // jobs taken from the queue whose sessions are being created (in neither
  // the queue nor the active list until they join it)
  size_t                        admitting = 0;
  std::map<long long, Job *>    jobs;
  std::thread                   worker;
  bool                          stop = false;
  long long                     next_id = 1;
  double                        cell_iters = 0, busy_s = 0;
  long long                     ticks = 0;
  // launch shape of the batched hydraulic launches, tuned online per size
  // class of a tick's hydraulic work (tick())
  struct Shape
  {
    int    probe = 0, rows = 48, variant = 1;
    double cost[6] = {0, 0, 0, 0, 0, 0};
  };
  std::map<int, Shape>          shapes;
};

namespace
{

using namespace hmg;

size_t job_bytes(const hmg_server *sv, const hmg_job_desc &d)
{
  const size_t n = (size_t)d.nx * (size_t)d.ny;
  const size_t g = (size_t)1 << 16; // pool granularity
  size_t       b = hmg_session_bytes(d.nx, d.ny, sv->storage) + 12 * g;
  if (!d.has_input) b += n * 4 + g; // the generated input terrain
  return b;
}

int admit(hmg_server *sv, Job *j)
{
  hmg_ctx           *ctx = sv->ctx;
  const hmg_job_desc &d = j->desc;
  hmg_vpipes_params  p = sv->p;
  p.mode = HMG_VP_FUSED;
  // thermal-only jobs step their z plane directly (tick): separate planes
  const bool thermal_only = d.kind == HMG_JOB_THERMAL;
  if (d.has_input)
  {
    if (thermal_only)
      return session_create_separate(ctx, j->z_in.data(), false, d.nx, d.ny, &p, sv->storage,
                                     &j->s);
    return hmg_session_create(ctx, j->z_in.data(), d.nx, d.ny, &p, nullptr, sv->storage, &j->s);
  }
  // generate remap(noise_fbm(PERLIN, kw 4, seed)) on the device
  DevBuf t;
  HMG_TRY(t.get(ctx->pool, (size_t)d.nx * d.ny * sizeof(float)));
  HMG_TRY(hmg_noise_fbm_dev(ctx, t.as<float>(), d.nx, d.ny, 1, 4.f, 4.f, d.seed, 8, 0.7f, 0.5f,
                            2.f, nullptr, nullptr, nullptr, nullptr, nullptr));
  HMG_TRY(hmg_remap_dev(ctx, t.as<float>(), (size_t)d.nx * d.ny, 0.f, 1.f));
  if (thermal_only)
    return session_create_separate(ctx, t.as<float>(), true, d.nx, d.ny, &p, sv->storage, &j->s);
  return hmg_session_create_dev(ctx, t.as<float>(), d.nx, d.ny, &p, nullptr, sv->storage, &j->s);
}

// Publishes a job's state to pollers: state, t_admit and t_done are read
// under sv->mu by other threads, and the lock orders the result copy into
// the caller's buffer before a poller can see HMG_JOB_DONE.
void set_state(hmg_server *sv, Job *j, int state)
{
  std::lock_guard<std::mutex> lk(sv->mu);
  j->state = state;
}

// Candidate shapes of the batched hydraulic launches (rows per wavefront,
// variant: 1 one row ahead, 2 gathers ahead).  Ticks are grouped in size
// classes, floor(log2(cells of the tick's hydraulic jobs)); the first 2 x 6
// ticks of a class run the candidates in turn, each timed on the device; the
// cost of a tick is its device time per cell-iteration, the cheaper of a
// shape's two ticks counts, and the cheapest shape is kept for the class.
// Only ticks whose jobs are all past their first TUNE_AFTER_IT hydraulic
// iterations probe: while the water spreads from rest the sediment
// backtraces are short and hit the cache, which ranks the shapes
// differently from the rest of a run.  Results do not depend on the shape.
constexpr int SERVE_CAND[6][2] = {{32, 1}, {48, 1}, {64, 1}, {40, 2}, {56, 2}, {72, 2}};
constexpr int N_SERVE_CAND = 6;
constexpr int TUNE_AFTER_IT = 20;

// Rows per wavefront for a tick: at most `rows`, shortened (not below 8)
// while the tick's maps would give fewer than one wavefront per SIMD, so
// that a few small maps still spread over the device.
int capped_rows(const hmg_server *sv, const std::vector<Job *> &vj, int rows)
{
  const long target = 4L * (sv->ctx->num_cus > 0 ? sv->ctx->num_cus : 256);
  for (;;)
  {
    long items = 0;
    for (const Job *j : vj)
    {
      const long strips = (j->desc.nx + 59) / 60; // 60 columns per wavefront
      items += strips * ((j->desc.ny + rows - 1) / rows);
    }
    if (items >= target || rows <= 8) return rows;
    rows = rows / 2 < 8 ? 8 : rows / 2;
  }
}

// One tick over the active jobs; returns the cell-iterations done.
double tick(hmg_server *sv, std::vector<Job *> &act)
{
  hmg_ctx                    *ctx = sv->ctx;
  std::vector<hmg_session *> vs;
  std::vector<int>           vi;
  std::vector<Job *>         vj;
  double                     work = 0;
  for (Job *j : act)
    if (j->desc.kind == HMG_JOB_VPIPES_THERMAL)
    {
      const int every = j->desc.thermal_every > 0 ? j->desc.thermal_every : (1 << 30);
      int       k = every - j->done % every;
      if (k > j->desc.iterations - j->done) k = j->desc.iterations - j->done;
      if (k > 10) k = 10;
      vs.push_back(j->s);
      vi.push_back(k);
      vj.push_back(j);
    }
  int rc = HMG_OK;
  if (!vs.empty())
  {
    double cells = 0, tick_cells = 0;
    bool   young = false;
    for (size_t q = 0; q < vj.size(); q++)
    {
      const double c = (double)vj[q]->desc.nx * vj[q]->desc.ny;
      tick_cells += c;
      cells += c * vi[q];
      young = young || vj[q]->done < TUNE_AFTER_IT;
    }
    hmg_server::Shape &sh = sv->shapes[(int)std::floor(std::log2(tick_cells))];
    const bool         probing = !young && sh.probe < 2 * N_SERVE_CAND;
    const int          ci = sh.probe % N_SERVE_CAND;
    const int          rows = capped_rows(sv, vj, probing ? SERVE_CAND[ci][0] : sh.rows);
    const int          variant = probing ? SERVE_CAND[ci][1] : sh.variant;
    float              ms = 0.f;
    rc = batch_advance(ctx, vs.data(), vi.data(), (int)vs.size(), 0, rows, variant,
                       probing ? &ms : nullptr);
    if (probing && rc == HMG_OK)
    {
      const double c = cells > 0 ? ms / cells : 0.0;
      if (sh.probe < N_SERVE_CAND || c < sh.cost[ci]) sh.cost[ci] = c;
      if (++sh.probe == 2 * N_SERVE_CAND)
      {
        int b = 0;
        for (int q = 1; q < N_SERVE_CAND; q++)
          if (sh.cost[q] < sh.cost[b]) b = q;
        sh.rows = SERVE_CAND[b][0];
        sh.variant = SERVE_CAND[b][1];
      }
    }
    for (size_t q = 0; q < vj.size(); q++)
    {
      Job *j = vj[q];
      if (rc != HMG_OK)
      {
        set_state(sv, j, rc);
        continue;
      }
      j->done += vi[q];
      work += (double)j->desc.nx * j->desc.ny * vi[q];
      const int every = j->desc.thermal_every;
      if (every > 0 && j->done % every == 0)
      {
        int r2 = hmg_session_thermal(j->s, j->desc.talus, 1);
        if (r2 != HMG_OK) set_state(sv, j, r2);
      }
    }
  }
  for (Job *j : act)
    if (j->desc.kind == HMG_JOB_THERMAL)
    {
      int k = j->desc.iterations - j->done;
      if (k > 8) k = 8;
      hmg_session *s = j->s;
      float       *zc = s->z[s->cur].as<float>(), *res = zc;
      // the job's last steps end with thermal()'s extrapolate_borders
      const bool   last = j->done + k == j->desc.iterations;
      int          r2 = launch_thermal(ctx, zc, s->z[s->cur ^ 1].as<float>(), s->nx, s->ny,
                                       nullptr, j->desc.talus, k, nullptr, ctx->stream, &res,
                                       last);
      if (r2 != HMG_OK)
      {
        set_state(sv, j, r2);
        continue;
      }
      // an odd number of launches leaves the result in the other plane
      if (res != zc) std::swap(s->z[s->cur].ptr, s->z[s->cur ^ 1].ptr);
      j->done += k;
      work += (double)j->desc.nx * j->desc.ny * k;
    }
  (void)hipStreamSynchronize(ctx->stream);
  return work;
}

void worker_main(hmg_server *sv)
{
  for (;;)
  {
    std::vector<Job *> admitted;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv.wait(lk, [&] { return sv->stop || !sv->queue.empty() || !sv->active.empty(); });
      if (sv->stop && sv->queue.empty() && sv->active.empty()) return;
      // admission in arrival order while the state fits in the budget
      while (!sv->queue.empty())
      {
        Job         *j = sv->queue.front();
        const bool   alone = sv->active.empty() && admitted.empty();
        const size_t cells = (size_t)j->desc.nx * j->desc.ny;
        if (sv->used + j->bytes > sv->budget && !alone) break;
        if (sv->max_cells && sv->active_cells + cells > sv->max_cells && !alone) break;
        sv->queue.pop_front();
        sv->used += j->bytes;
        sv->active_cells += cells;
        admitted.push_back(j);
      }
      sv->admitting = admitted.size();
    }
    const double t0 = now_s();
    std::lock_guard<std::recursive_mutex> ctx_lock(sv->ctx->mu);
    (void)ctx_enter(sv->ctx);
    for (Job *j : admitted)
    {
      if (j->bytes > sv->budget)
      {
        set_state(sv, j, HMG_ERR_OOM);
        continue;
      }
      int                         rc = admit(sv, j);
      std::lock_guard<std::mutex> lk(sv->mu);
      j->t_admit = now_s();
      j->state = rc == HMG_OK ? HMG_JOB_RUNNING : rc;
    }
    std::vector<Job *> act;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (Job *j : admitted) sv->active.push_back(j);
      sv->admitting = 0;
      for (Job *j : sv->active)
        if (j->state == HMG_JOB_RUNNING && j->done < j->desc.iterations) act.push_back(j);
    }
    const double work = act.empty() ? 0.0 : tick(sv, act);
    // results of finished (or failed) jobs to the callers' host buffers
    std::vector<Job *> finished;
    std::vector<int>   final_state;
    for (Job *j : sv->active)
      if (j->state != HMG_JOB_RUNNING || j->done >= j->desc.iterations) finished.push_back(j);
    for (Job *j : finished)
    {
      int st = j->state;
      if (st == HMG_JOB_RUNNING && j->s)
      {
        int rc = HMG_OK;
        // thermal() extrapolates the borders even after zero steps
        if (j->desc.kind == HMG_JOB_THERMAL && j->desc.iterations == 0)
          rc = launch_extrapolate_borders(j->s->z[j->s->cur].as<float>(), j->s->nx, j->s->ny,
                                          sv->ctx->stream);
        // the export checks the state: a job whose state holds a NaN or an
        // infinity fails with HMG_ERR_NONFINITE, alone
        if (rc == HMG_OK)
          rc = hmg_session_export(j->s, j->out, nullptr, nullptr, nullptr, nullptr);
        st = rc == HMG_OK ? HMG_JOB_DONE : rc;
      }
      if (j->s) hmg_session_destroy(j->s);
      j->s = nullptr;
      final_state.push_back(st);
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      const double                t_done = now_s();
      for (size_t q = 0; q < finished.size(); q++)
      {
        Job *j = finished[q];
        j->state = final_state[q];
        j->t_done = t_done;
        sv->used -= j->bytes;
        sv->active_cells -= (size_t)j->desc.nx * j->desc.ny;
        sv->active.erase(std::find(sv->active.begin(), sv->active.end(), j));
      }
      sv->cell_iters += work;
      sv->busy_s += now_s() - t0;
      sv->ticks++;
    }
    if (!finished.empty()) sv->cv_done.notify_all();
  }
}

} // namespace

extern "C" {

int hmg_server_create(hmg_ctx                 *ctx,
                      size_t                   hbm_budget,
                      int                      storage,
                      const hmg_vpipes_params *p,
                      hmg_server             **out)
{
  clear_error();
  if (!ctx || !out) return set_error(HMG_ERR_INVALID, "null argument");
  if (storage != HMG_STORAGE_F32 && storage != HMG_STORAGE_BF16 && storage != HMG_STORAGE_F16)
    return set_error(HMG_ERR_INVALID, "unknown storage %d", storage);
  hmg_server *sv = new hmg_server();
  sv->ctx = ctx;
  sv->storage = storage;
  if (p)
    sv->p = *p;
  else
    hmg_vpipes_defaults(&sv->p);
  int rc = validate_vp_params(sv->p);
  if (rc != HMG_OK)
  {
    delete sv;
    return rc;
  }
  if (hbm_budget == 0)
  {
    size_t free_b = 0, total_b = 0;
    (void)ctx_enter(ctx);
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    hbm_budget = free_b / 10 * 8;
  }
  sv->budget = hbm_budget;
  sv->worker = std::thread(worker_main, sv);
  *out = sv;
  return HMG_OK;
}

int hmg_server_submit(hmg_server         *sv,
                      const hmg_job_desc *d,
                      const float        *z_host,
                      float              *z_out_host,
                      long long          *job_id)
{
  clear_error();
  if (!sv || !d || !z_out_host) return set_error(HMG_ERR_INVALID, "null argument");
  // jobs advance in batched streaming launches, which address planes with
  // 32-bit offsets: at most 2^28 cells
  if (d->nx < 3 || d->ny < 3 || (size_t)d->nx * (size_t)d->ny > ((size_t)1 << 28))
    return set_error(HMG_ERR_INVALID, "job grid must be at least 3x3 and at most 2^28 cells");
  if (d->iterations < 0) return set_error(HMG_ERR_INVALID, "iterations must be >= 0");
  if (d->kind != HMG_JOB_VPIPES_THERMAL && d->kind != HMG_JOB_THERMAL)
    return set_error(HMG_ERR_INVALID, "unknown job kind %d", d->kind);
  if ((d->has_input != 0) != (z_host != nullptr))
    return set_error(HMG_ERR_INVALID, "has_input must match z_host");
  Job *j = new Job();
  j->desc = *d;
  j->out = z_out_host;
  if (z_host) j->z_in.assign(z_host, z_host + (size_t)d->nx * d->ny);
  j->bytes = job_bytes(sv, *d);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop)
    {
      delete j;
      return set_error(HMG_ERR_INVALID, "server is stopping");
    }
    j->id = sv->next_id++;
    j->t_submit = now_s();
    sv->jobs[j->id] = j;
    sv->queue.push_back(j);
  }
  sv->cv.notify_one();
  if (job_id) *job_id = j->id;
  return HMG_OK;
}

int hmg_server_poll(hmg_server *sv, long long job_id, int *state)
{
  if (!sv || !state) return set_error(HMG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto                        it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return set_error(HMG_ERR_INVALID, "unknown job %lld", job_id);
  *state = it->second->state;
  return HMG_OK;
}

int hmg_server_job_times(hmg_server *sv,
                         long long   job_id,
                         double     *t_submit,
                         double     *t_admit,
                         double     *t_done)
{
  if (!sv) return set_error(HMG_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto                        it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return set_error(HMG_ERR_INVALID, "unknown job %lld", job_id);
  if (t_submit) *t_submit = it->second->t_submit;
  if (t_admit) *t_admit = it->second->t_admit;
  if (t_done) *t_done = it->second->t_done;
  return HMG_OK;
}

int hmg_server_set_max_active_cells(hmg_server *sv, size_t cells)
{
  if (!sv) return set_error(HMG_ERR_INVALID, "null server");
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->max_cells = cells;
  }
  sv->cv.notify_all();
  return HMG_OK;
}

int hmg_server_wait_all(hmg_server *sv)
{
  if (!sv) return set_error(HMG_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  sv->cv_done.wait(lk,
                   [&] { return sv->queue.empty() && sv->active.empty() && sv->admitting == 0; });
  return HMG_OK;
}

int hmg_server_stats(hmg_server *sv,
                     double     *cell_iters,
                     long long  *ticks,
                     double     *busy_s,
                     size_t     *budget,
                     size_t     *used)
{
  if (!sv) return set_error(HMG_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (cell_iters) *cell_iters = sv->cell_iters;
  if (ticks) *ticks = sv->ticks;
  if (busy_s) *busy_s = sv->busy_s;
  if (budget) *budget = sv->budget;
  if (used) *used = sv->used;
  return HMG_OK;
}

int hmg_server_destroy(hmg_server *sv)
{
  if (!sv) return HMG_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (auto &kv : sv->jobs)
  {
    if (kv.second->s) hmg_session_destroy(kv.second->s);
    delete kv.second;
  }
  delete sv;
  return HMG_OK;
}

} // extern "C"
