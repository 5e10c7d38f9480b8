// letkf_serve.hip -- continuous batching of LETKF analysis jobs on one GPU.
//
// FIFO admission by HBM budget and an active-job limit; every admitted job
// runs on its own host thread, engine handle and HIP stream, so kernels of
// different jobs overlap on the device.  The budget is allocated once, when
// the server is created: an admitted job gets a contiguous region of it
// (first fit), and every device buffer of the job -- its engine handle's,
// its state and operator buffers -- is carved from that region, so no
// hipMalloc or hipFree (which synchronises the device) runs while other
// jobs' kernels are in flight.  A job's ensemble stays in device
// memory between its cycles: the analysis of cycle c (mean-zero perturbations
// and mean) is the background of cycle c+1, ping-ponging between the caller's
// output arrays and one server-owned state buffer.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "dida_letkf.h"
#include "letkf_kernels.h"
#include "letkf_pool.h"

using dida::DevicePool;
using dida::FitAlloc;
using dida::set_error;

namespace {

double now_s() {
  using clk = std::chrono::steady_clock;
  static const clk::time_point t0 = clk::now();
  return std::chrono::duration<double>(clk::now() - t0).count();
}

struct Job {
  long long id = 0;
  dida_serve_job spec{};
  std::vector<float> glat, glon;
  std::vector<double> hyai, hybi;      // the config's pointers are redirected here
  bool op_mode = false;                // some cycle computes Yb, d on the device
  std::vector<dida_serve_obs> obs;
  dida_serve_status st{};
  std::atomic<bool> cancel{false};
  int max_sites = 0, max_rows = 0;
  size_t region = FitAlloc::npos;      // offset of the job's region in the arena
  DevicePool pool;                     // the region and its allocator
};

// A block of a job's region, given back when it goes out of scope.
struct PoolBlock {
  DevicePool* pool = nullptr;
  size_t off = FitAlloc::npos, n = 0;
  double* get(DevicePool* pl, size_t bytes) {
    pool = pl;
    n = bytes;
    off = pl->fit.alloc(bytes);
    return off == FitAlloc::npos ? nullptr : reinterpret_cast<double*>(pl->base + off);
  }
  void drop() {
    if (pool && off != FitAlloc::npos) (void)pool->fit.release(off, n);
    pool = nullptr;
    off = FitAlloc::npos;
  }
  ~PoolBlock() { drop(); }
};

}  // namespace

struct dida_serve {
  dida_serve_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv;          // job state changes
  std::deque<std::shared_ptr<Job>> queue;
  std::map<long long, std::shared_ptr<Job>> jobs;
  std::map<long long, std::thread> workers;   // running (or just finished) jobs
  std::vector<long long> finished;            // workers to join
  size_t used = 0;
  char* arena = nullptr;               // the whole budget, allocated at create
  FitAlloc fit;                        // regions of admitted jobs
  int active = 0;
  long long next_id = 1;
  bool stopping = false;
  std::thread dispatcher;
};

static void run_job(dida_serve* s, std::shared_ptr<Job> j) {
  (void)hipSetDevice(s->device);
  // every device buffer below comes from the job's region
  DevicePool* prev_pool = dida::bind_thread_pool(&j->pool);
  dida_serve_job& sp = j->spec;
  dida_letkf_handle* h = nullptr;
  int rc = dida_letkf_create(&sp.cfg, &h);
  PoolBlock b_xt, b_at, b_yb, b_dep;
  double* xt = nullptr;
  double* at = nullptr;
  double* yb_op = nullptr;
  double* dep_op = nullptr;
  long long n_an = 0;
  const size_t n_state = (size_t)(sp.cfg.nlat + 2 * sp.cfg.halo) * (sp.cfg.nlon + 2 * sp.cfg.halo) *
                         (sp.cfg.n2d + sp.cfg.n3d * sp.cfg.nz);
  const size_t k = sp.cfg.k;
  if (rc == DIDA_OK) rc = dida_letkf_set_grid(h, j->glat.data(), j->glon.data());
  if (rc == DIDA_OK && sp.n_cycles > 1) {
    xt = b_xt.get(&j->pool, n_state * k * 8);
    at = b_at.get(&j->pool, n_state * 8);
    if (!xt || !at) rc = set_error(DIDA_ERR_NOMEM, "job region exhausted (state buffers)");
  }
  if (rc == DIDA_OK && j->op_mode) {
    const size_t nr = (size_t)std::max(j->max_rows, 1);
    yb_op = b_yb.get(&j->pool, nr * k * 8);
    dep_op = b_dep.get(&j->pool, nr * 8);
    if (!yb_op || !dep_op) rc = set_error(DIDA_ERR_NOMEM, "job region exhausted (operator)");
  }
  // ping-pong: cycle c reads (bx, ba) and writes (ox, oa); the last cycle
  // writes the caller's (xa, axa)
  const double* bx = sp.xb;
  const double* ba = sp.axb;
  for (int c = 0; c < sp.n_cycles && rc >= 0; ++c) {
    if (j->cancel.load()) {
      rc = DIDA_ERR_STATE;
      break;
    }
    const int left = sp.n_cycles - 1 - c;     // cycles after this one
    double* ox = (left % 2 == 0) ? sp.xa : xt;
    double* oa = (left % 2 == 0) ? sp.axa : at;
    const dida_serve_obs& o = j->obs[c];
    const double* yb = o.yb;
    const double* dep = o.dep;
    if (!yb) {
      // H(x) and centring from this cycle's background, on the job's stream
      int r = dida_letkf_obs_operator_device(h, bx, ba, o.n_sites, o.site_lat, o.site_lon,
                                             o.site_row_ptr, o.n_rows, o.row_var, o.row_ph, o.yo,
                                             yb_op, dep_op);
      if (r < 0) { rc = r; break; }
      yb = yb_op;
      dep = dep_op;
    }
    int r = dida_letkf_set_obs_device(h, o.n_sites, o.site_lat, o.site_lon, o.site_row_ptr,
                                      o.n_rows, yb, dep, o.rinv, o.vcoord);
    if (r < 0) { rc = r; break; }
    dida_letkf_stats stt;
    r = dida_letkf_analyze_device(h, bx, ba, ox, oa, &stt);
    if (r < 0) { rc = r; break; }
    if (j->op_mode && sp.cfg.halo > 0) {
      // the next cycle's observation operator reads the halo
      const int rh = dida_letkf_fill_halo_device(h, ox, oa);
      if (rh < 0) { rc = rh; break; }
    }
    rc = std::max(rc, r);
    n_an += stt.n_analyses;
    bx = ox;
    ba = oa;
  }
  if (h) dida_letkf_destroy(h);
  for (PoolBlock* b : {&b_xt, &b_at, &b_yb, &b_dep}) b->drop();
  dida::bind_thread_pool(prev_pool);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // every block is back in the job's region; the region goes back to the arena
    s->fit.release(j->region, j->st.bytes);
    j->st.t_end = now_s();
    j->st.rc = rc;
    j->st.n_analyses = n_an;
    j->st.state = j->cancel.load() && rc == DIDA_ERR_STATE ? DIDA_JOB_CANCELLED
                  : (rc < 0 ? DIDA_JOB_FAILED : DIDA_JOB_DONE);
    s->used -= j->st.bytes;
    s->active -= 1;
    s->finished.push_back(j->id);
  }
  s->cv.notify_all();
}

// Joins the worker threads of finished jobs (called with s->mu held; a
// finished worker only has its return left to run).
static void reap_workers(dida_serve* s) {
  for (long long id : s->finished) {
    auto it = s->workers.find(id);
    if (it == s->workers.end()) continue;
    if (it->second.joinable()) it->second.join();
    s->workers.erase(it);
  }
  s->finished.clear();
}

static void dispatch_loop(dida_serve* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // the head job fits when the arena has a free block of its size
    auto fits = [&] {
      return FitAlloc::round(s->queue.front()->st.bytes) <= s->fit.largest();
    };
    s->cv.wait(lk, [&] {
      if (s->stopping || !s->finished.empty()) return true;
      if (s->queue.empty() || s->active >= s->cfg.max_active) return false;
      return fits();
    });
    reap_workers(s);
    if (!s->stopping && (s->queue.empty() || s->active >= s->cfg.max_active || !fits()))
      continue;
    if (s->stopping && s->queue.empty()) break;
    if (s->stopping) {
      // drain: cancel what was never admitted
      for (auto& j : s->queue) {
        j->st.state = DIDA_JOB_CANCELLED;
        j->st.t_end = now_s();
      }
      s->queue.clear();
      s->cv.notify_all();
      break;
    }
    std::shared_ptr<Job> j = s->queue.front();
    s->queue.pop_front();
    if (j->cancel.load()) {
      j->st.state = DIDA_JOB_CANCELLED;
      j->st.t_end = now_s();
      s->cv.notify_all();
      continue;
    }
    j->region = s->fit.alloc(j->st.bytes);
    j->pool.base = s->arena + j->region;
    j->pool.fit.reset(FitAlloc::round(j->st.bytes));
    s->used += j->st.bytes;
    s->active += 1;
    j->st.state = DIDA_JOB_RUNNING;
    j->st.t_start = now_s();
    s->workers.emplace(j->id, std::thread(run_job, s, j));
  }
}

extern "C" {

double dida_serve_now(void) { return now_s(); }

int dida_serve_create(const dida_serve_config* cfg, dida_serve** out) {
  if (!cfg || !out || cfg->max_active < 1 || cfg->hbm_budget_bytes == 0)
    return set_error(DIDA_ERR_ARG, "serve config: max_active >= 1 and a budget are needed");
  dida_serve* s = new (std::nothrow) dida_serve();
  if (!s) return DIDA_ERR_NOMEM;
  s->cfg = *cfg;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return DIDA_ERR_HIP;
  }
  s->fit.reset(cfg->hbm_budget_bytes);
  if (hipMalloc(&s->arena, s->fit.size()) != hipSuccess) {
    delete s;
    return set_error(DIDA_ERR_NOMEM, "serve: the HBM budget of %zu bytes cannot be allocated",
                     (size_t)cfg->hbm_budget_bytes);
  }
  now_s();
  s->dispatcher = std::thread(dispatch_loop, s);
  *out = s;
  return DIDA_OK;
}

int dida_serve_submit(dida_serve* s, const dida_serve_job* job, long long* job_id) {
  if (!s || !job || !job_id || job->n_cycles < 1 || !job->obs || !job->grid_lat ||
      !job->grid_lon || !job->xb || !job->axb || !job->xa || !job->axa)
    return set_error(DIDA_ERR_ARG, "submit: null argument or n_cycles < 1");
  auto j = std::make_shared<Job>();
  j->spec = *job;
  if (job->cfg.hyai && job->cfg.hybi) {
    j->hyai.assign(job->cfg.hyai, job->cfg.hyai + job->cfg.nz + 1);
    j->hybi.assign(job->cfg.hybi, job->cfg.hybi + job->cfg.nz + 1);
    j->spec.cfg.hyai = j->hyai.data();
    j->spec.cfg.hybi = j->hybi.data();
  } else {
    j->spec.cfg.hyai = nullptr;
    j->spec.cfg.hybi = nullptr;
  }
  j->glat.assign(job->grid_lat, job->grid_lat + job->cfg.nlat);
  j->glon.assign(job->grid_lon, job->grid_lon + job->cfg.nlon);
  j->obs.assign(job->obs, job->obs + job->n_cycles);
  for (const auto& o : j->obs) {
    if (o.n_sites < 0 || o.n_rows < 0) return set_error(DIDA_ERR_ARG, "submit: negative obs sizes");
    if (!o.yb) {
      if (!o.row_var || !o.row_ph || !o.yo)
        return set_error(DIDA_ERR_ARG, "submit: a cycle without yb needs row_var, row_ph and yo");
      if (job->cfg.halo < 1 || j->hyai.empty())
        return set_error(DIDA_ERR_ARG, "submit: the observation operator needs halo >= 1 and hyai/hybi");
      j->op_mode = true;
    } else if (!o.dep) {
      return set_error(DIDA_ERR_ARG, "submit: yb without dep");
    }
    j->max_sites = std::max(j->max_sites, o.n_sites);
    j->max_rows = std::max(j->max_rows, o.n_rows);
  }
  // a serving job's batches are bounded to 4 GiB of scratch, row lists
  // included, unless set
  if (j->spec.cfg.max_batch_columns <= 0)
    j->spec.cfg.max_batch_columns = (int)std::min<long long>(
        dida::batch_columns_for_bytes(job->cfg, 4ull << 30, j->max_rows), 1 << 30);
  j->st.bytes = dida_letkf_job_bytes(&j->spec.cfg, j->max_sites, j->max_rows);
  if (j->st.bytes == 0) return set_error(DIDA_ERR_ARG, "submit: bad job configuration");
  if (FitAlloc::round(j->st.bytes) > s->fit.size())
    return set_error(DIDA_ERR_ADMISSION, "submit: job needs %zu bytes, budget %zu", j->st.bytes,
                     (size_t)s->cfg.hbm_budget_bytes);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return set_error(DIDA_ERR_STATE, "submit: server is stopping");
    j->id = s->next_id++;
    j->st.state = DIDA_JOB_QUEUED;
    j->st.t_submit = now_s();
    s->jobs[j->id] = j;
    s->queue.push_back(j);
    *job_id = j->id;
  }
  s->cv.notify_all();
  return DIDA_OK;
}

int dida_serve_poll(dida_serve* s, long long id, dida_serve_status* st) {
  if (!s || !st) return DIDA_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return DIDA_ERR_ARG;
  *st = it->second->st;
  return DIDA_OK;
}

int dida_serve_wait(dida_serve* s, long long id, dida_serve_status* st) {
  if (!s) return DIDA_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return DIDA_ERR_ARG;
  auto j = it->second;
  s->cv.wait(lk, [&] { return j->st.state >= DIDA_JOB_DONE; });
  if (st) *st = j->st;
  return DIDA_OK;
}

int dida_serve_cancel(dida_serve* s, long long id) {
  if (!s) return DIDA_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return DIDA_ERR_ARG;
  it->second->cancel.store(true);
  if (it->second->st.state == DIDA_JOB_QUEUED) {
    for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
      if ((*q)->id == id) {
        s->queue.erase(q);
        break;
      }
    it->second->st.state = DIDA_JOB_CANCELLED;
    it->second->st.t_end = now_s();
    s->cv.notify_all();
  }
  return DIDA_OK;
}

int dida_serve_destroy(dida_serve* s) {
  if (!s) return DIDA_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  // running jobs finish on their own threads
  for (auto& t : s->workers)
    if (t.second.joinable()) t.second.join();
  if (s->arena) (void)hipFree(s->arena);
  delete s;
  return DIDA_OK;
}

}  // extern "C"
