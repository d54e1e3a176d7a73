// nb_serve.hip — the serving loop of nbody.h (nb_server_*): continuous
// batching of simulation jobs on one GPU.
//
// A job moves queued -> preparing -> running -> done. Each tick:
//   1. select  scan the first `lookahead` queued jobs in arrival order; a job
//              whose device memory fits what is left of the HBM budget (Barnes-Hut
//              contexts) or of the shared direct-sum pool, and whose estimated
//              step time fits what is left of the tick budget, is selected: its
//              memory and tick time are reserved and a host worker generates its
//              initial conditions. A job estimated above the whole tick budget is
//              selected only onto an idle device. Later, smaller jobs may pass a
//              job that does not fit yet (bounded by the look-ahead);
//   2. start   selected jobs whose initial conditions are ready are uploaded:
//              direct-sum jobs become clusters of the shared nb_batch pool,
//              Barnes-Hut jobs get their own nb_ctx (own stream);
//   3. step    one nb_batch_step for every direct-sum job at once (one persistent
//              force launch) and one nb_step per Barnes-Hut job, all enqueued
//              before any wait, so Barnes-Hut kernels overlap with each other and
//              with the batch;
//   4. retire  jobs that reached their step count copy their FP64 state to the
//              host and release their reservations.
// Only selected jobs hold host memory for initial conditions, so a long queue
// costs a few bytes per job.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "nb_internal.h"
#include "nbody_ic.h"

namespace {

struct Job {
  nb_job desc{};
  int32_t state = NB_JOB_QUEUED;  // preparing jobs report NB_JOB_QUEUED
  int32_t solver = NB_SOLVER_DIRECT;
  int32_t steps_done = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  // reservations held from selection to retirement
  int64_t bytes = 0;        // HBM budget (Barnes-Hut contexts)
  int64_t pool_bodies = 0;  // direct-sum pool bodies (padded)
  double cost_ms = 0;       // estimated device time of one step
  int32_t slot = -1;        // batch cluster id
  nb_ctx* ctx = nullptr;
  // initial conditions (made by a worker) and, later, the final state
  std::atomic<int> prepared{0};  // 0 pending, 1 ready, -1 failed
  std::vector<double> pos, vel;
  std::vector<float> mass;
};

int64_t tree_bound(int64_t n) {
  int64_t total = 0, w = 1;
  for (int l = 0; l <= kMaxLevel; ++l) {
    total += w < n ? w : n;
    if (w < n) w *= 8;
  }
  return total;
}

// Device bytes of a Barnes-Hut context of n bodies (nb_create + nb_bh_alloc).
int64_t bh_bytes(int64_t n) {
  const int64_t np = (n + 255) / 256 * 256;
  return np * (16 + 4) + n * (4 * 24 + 4 * 5 + 16 + 32 + 8) + 32 * tree_bound(n) + (4 << 20);
}

int64_t padded(int64_t n) { return (n + kPadBodies - 1) / kPadBodies * kPadBodies; }

// Estimated device time of one step, from rates measured on MI355X: the
// compensated direct sum runs about 3.2e12 pair interactions/s once the device
// is full, and a Barnes-Hut step (theta = 0.5, build + traversal) costs about
// 2 ns per body plus some 60 us of launches (8.5 ms at 4,194,304 bodies).
double step_cost_ms(int64_t n, int32_t solver) {
  if (solver == NB_SOLVER_DIRECT) return 2e-3 + (double)n * (double)n / 3.2e9;
  return 0.06 + (double)n * 2.0e-6;
}

}  // namespace

struct nb_server {
  nb_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  nb_batch* batch = nullptr;
  int64_t budget_total = 0;  // HBM budget beyond the direct-sum pool
  int64_t bytes_left = 0;    // of budget_total, not reserved
  int64_t pool_left = 0;     // direct-sum pool bodies not reserved
  double load_ms = 0;        // reserved tick time (preparing and running jobs)
  int64_t next_id = 0;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::deque<int64_t> queue;       // not yet selected, arrival order
  std::vector<int64_t> preparing;  // selected, initial conditions pending or ready
  std::vector<int64_t> running;    // on the device
  // initial-condition workers
  std::mutex mu;
  std::condition_variable cv;        // workers wait for jobs to prepare
  std::condition_variable ready_cv;  // an idle tick waits for a prepared job
  std::deque<Job*> prep_queue;
  bool stop = false;
  std::vector<std::thread> workers;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  void worker() {
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !prep_queue.empty(); });
        if (stop) return;
        j = prep_queue.front();
        prep_queue.pop_front();
      }
      const int64_t n = j->desc.n;
      int rc = -1;
      try {
        j->pos.resize(3 * n);
        j->vel.resize(3 * n);
        j->mass.resize(n);
        rc = j->desc.ic == NB_IC_BALL
                 ? nb_ic_ball(j->desc.seed, n, 100.0, 1.0, j->pos.data(), j->vel.data(),
                              j->mass.data())
                 : nb_ic_plummer(j->desc.seed, n, 50.0, 1e4, 10.0,
                                 j->desc.G == 0.f ? 1.0 : j->desc.G, j->pos.data(),
                                 j->vel.data(), j->mass.data());
      } catch (const std::bad_alloc&) {
        rc = -1;
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        j->prepared.store(rc == 0 ? 1 : -1);
      }
      ready_cv.notify_all();
    }
  }

  void release(Job& j) {
    bytes_left += j.bytes;
    pool_left += j.pool_bodies;
    load_ms -= j.cost_ms;
    if (load_ms < 1e-9) load_ms = 0;
    j.bytes = j.pool_bodies = 0;
    j.cost_ms = 0;
  }
};

namespace {

// Upload a prepared job. NB_E_BUDGET: the pool has no free segment of that
// size yet (fragmentation); the job stays prepared and is retried next tick.
int start(nb_server* s, Job& j) {
  const int64_t n = j.desc.n;
  if (j.solver == NB_SOLVER_DIRECT) {
    int32_t id = -1;
    const int rc = nb_batch_add(s->batch, j.pos.data(), j.vel.data(), nullptr, j.mass.data(), n,
                                j.desc.eps2, j.desc.dt, j.desc.G, &id);
    if (rc != NB_OK) return rc;
    j.slot = id;
  } else {
    nb_config c{};
    c.n = n;
    c.eps2 = j.desc.eps2;
    c.dt = j.desc.dt;
    c.G = j.desc.G;
    c.theta = s->cfg.theta;
    c.leaf_size = s->cfg.leaf_size;
    c.solver = NB_SOLVER_BH;
    int rc = nb_create(&j.ctx, &c);
    if (rc != NB_OK) return rc == NB_E_OOM ? NB_E_BUDGET : rc;
    rc = nb_upload_f64(j.ctx, j.pos.data(), j.vel.data(), nullptr, j.mass.data(), n);
    if (rc != NB_OK) {
      nb_destroy(j.ctx);
      j.ctx = nullptr;
      return rc;
    }
  }
  // the host copy of the initial conditions is no longer needed
  std::vector<double>().swap(j.pos);
  std::vector<double>().swap(j.vel);
  return NB_OK;
}

int retire(nb_server* s, Job& j) {
  const int64_t n = j.desc.n;
  int rc = NB_OK;
  if (j.solver == NB_SOLVER_DIRECT) {
    if (j.slot >= 0) {
      j.pos.resize(3 * n);
      j.vel.resize(3 * n);
      rc = nb_batch_download(s->batch, j.slot, j.pos.data(), j.vel.data(), nullptr);
      const int rr = nb_batch_remove(s->batch, j.slot);
      if (rc == NB_OK) rc = rr;
      j.slot = -1;
    }
  } else if (j.ctx) {
    j.pos.resize(3 * n);
    j.vel.resize(3 * n);
    rc = nb_download_f64(j.ctx, j.pos.data(), j.vel.data(), nullptr);
    nb_destroy(j.ctx);
    j.ctx = nullptr;
  }
  std::vector<float>().swap(j.mass);
  s->release(j);
  if (j.state != NB_JOB_FAILED) j.state = rc == NB_OK ? NB_JOB_DONE : NB_JOB_FAILED;
  if (j.state == NB_JOB_FAILED) {
    std::vector<double>().swap(j.pos);
    std::vector<double>().swap(j.vel);
  }
  j.t_done = s->now();
  return rc;
}

void free_server(nb_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->ready_cv.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->jobs)
    if (kv.second->ctx) nb_destroy(kv.second->ctx);
  if (s->batch) nb_batch_destroy(s->batch);
  delete s;
}

}  // namespace

extern "C" {

int nb_server_create(nb_server** out, const nb_server_config* cfg) {
  if (!out || !cfg) return nb_fail(NB_E_ARG, "nb_server_create: null argument");
  *out = nullptr;
  nb_server* s = new (std::nothrow) nb_server();
  if (!s) return nb_fail(NB_E_OOM, "nb_server_create: host allocation failed");
  s->cfg = *cfg;
  if (s->cfg.direct_pool <= 0) s->cfg.direct_pool = int64_t(1) << 24;
  if (s->cfg.max_jobs <= 0) s->cfg.max_jobs = 1024;
  if (s->cfg.bh_threshold <= 0) s->cfg.bh_threshold = 65536;
  if (s->cfg.theta <= 0.f) s->cfg.theta = 0.5f;
  if (s->cfg.ic_threads <= 0) s->cfg.ic_threads = 4;
  if (s->cfg.lookahead <= 0) s->cfg.lookahead = 16;
  if (!(s->cfg.tick_budget_ms > 0.f)) s->cfg.tick_budget_ms = 10.f;
  s->t0 = std::chrono::steady_clock::now();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete s;
    return nb_fail(NB_E_NOGPU, "nb_server_create: no HIP device visible");
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    delete s;
    return nb_fail(NB_E_HIP, "nb_server_create: hipMemGetInfo failed");
  }
  int64_t budget = s->cfg.hbm_budget > 0 ? s->cfg.hbm_budget : (int64_t)(0.6 * (double)free_b);
  const int64_t pool_bytes = s->cfg.direct_pool * 116;  // pm, mass, state, force per body
  if (pool_bytes >= budget) {
    delete s;
    return nb_fail(NB_E_BUDGET, "nb_server_create: the direct-sum pool exceeds the HBM budget");
  }
  nb_batch_config bc{};
  bc.max_bodies = s->cfg.direct_pool;
  bc.max_clusters = s->cfg.max_jobs;
  int rc = nb_batch_create(&s->batch, &bc);
  if (rc != NB_OK) {
    delete s;
    return rc;
  }
  s->budget_total = s->bytes_left = budget - pool_bytes;
  s->pool_left = s->cfg.direct_pool;
  for (int t = 0; t < s->cfg.ic_threads; ++t) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return NB_OK;
}

void nb_server_destroy(nb_server* s) { free_server(s); }

double nb_server_now(nb_server* s) { return s ? s->now() : 0.0; }

int nb_server_submit(nb_server* s, const nb_job* job, int64_t* id) {
  if (!s || !job || !id) return nb_fail(NB_E_ARG, "nb_server_submit: null argument");
  if (job->n < 1 || job->n >= (int64_t(1) << 27) || job->steps < 0 ||
      (job->ic != NB_IC_PLUMMER && job->ic != NB_IC_BALL) || !(job->eps2 >= 0.f))
    return nb_fail(NB_E_ARG, "nb_server_submit: bad job description");
  const int32_t solver = job->n >= s->cfg.bh_threshold ? NB_SOLVER_BH : NB_SOLVER_DIRECT;
  if (solver == NB_SOLVER_BH ? bh_bytes(job->n) > s->budget_total
                             : padded(job->n) > s->cfg.direct_pool)
    return nb_fail(NB_E_BUDGET, "nb_server_submit: the job can never fit the HBM budget");
  auto j = std::make_unique<Job>();
  j->desc = *job;
  j->solver = solver;
  j->cost_ms = step_cost_ms(job->n, solver);
  j->t_submit = s->now();
  const int64_t jid = s->next_id++;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->jobs[jid] = std::move(j);
  }
  s->queue.push_back(jid);
  *id = jid;
  return NB_OK;
}

int nb_server_tick(nb_server* s, int32_t* running_out, int32_t* finished_out) {
  if (!s) return nb_fail(NB_E_ARG, "nb_server_tick: null server");
  int finished = 0;
  // 1. select (arrival order, bounded look-ahead)
  int scanned = 0;
  for (auto it = s->queue.begin(); it != s->queue.end() && scanned < s->cfg.lookahead;) {
    if ((int)(s->running.size() + s->preparing.size()) >= s->cfg.max_jobs) break;
    ++scanned;
    Job& j = *s->jobs.at(*it);
    const bool idle = s->running.empty() && s->preparing.empty();
    const bool fits_tick = idle || s->load_ms + j.cost_ms <= s->cfg.tick_budget_ms;
    const bool fits_mem = j.solver == NB_SOLVER_BH ? bh_bytes(j.desc.n) <= s->bytes_left
                                                   : padded(j.desc.n) <= s->pool_left;
    if (!fits_tick || !fits_mem) {
      ++it;  // a later, smaller job may fit
      continue;
    }
    if (j.solver == NB_SOLVER_BH) {
      j.bytes = bh_bytes(j.desc.n);
      s->bytes_left -= j.bytes;
    } else {
      j.pool_bodies = padded(j.desc.n);
      s->pool_left -= j.pool_bodies;
    }
    s->load_ms += j.cost_ms;
    s->preparing.push_back(*it);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->prep_queue.push_back(&j);
    }
    s->cv.notify_one();
    it = s->queue.erase(it);
  }
  // Idle with work being prepared: wait (up to 50 ms) for the first prepared
  // job instead of returning an empty tick.
  if (s->running.empty() && !s->preparing.empty()) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->ready_cv.wait_for(lk, std::chrono::milliseconds(50), [&] {
      for (int64_t id : s->preparing)
        if (s->jobs.at(id)->prepared.load() != 0) return true;
      return false;
    });
  }
  // 2. start the prepared jobs (selection order)
  for (auto it = s->preparing.begin(); it != s->preparing.end();) {
    Job& j = *s->jobs.at(*it);
    const int p = j.prepared.load();
    if (p == 0) {
      ++it;
      continue;
    }
    const int rc = p < 0 ? NB_E_OOM : start(s, j);
    if (rc == NB_E_BUDGET) {
      ++it;
      continue;
    }
    if (rc != NB_OK) {
      j.state = NB_JOB_FAILED;
      retire(s, j);
      ++finished;
    } else {
      j.state = NB_JOB_RUNNING;
      j.t_admit = s->now();
      s->running.push_back(*it);
    }
    it = s->preparing.erase(it);
  }
  // 3. one step for every running job, everything enqueued before any wait
  bool any_direct = false;
  for (int64_t id : s->running) {
    Job& j = *s->jobs.at(id);
    if (j.steps_done >= j.desc.steps) continue;
    if (j.solver == NB_SOLVER_DIRECT) {
      any_direct = true;
    } else {
      const int rc = nb_step(j.ctx, 1);
      if (rc != NB_OK) return rc;
    }
  }
  if (any_direct) {
    const int rc = nb_batch_step(s->batch, 1);
    if (rc != NB_OK) return rc;
  }
  for (int64_t id : s->running) {
    Job& j = *s->jobs.at(id);
    if (j.steps_done >= j.desc.steps) continue;
    if (j.solver == NB_SOLVER_BH) {
      const int rc = nb_sync(j.ctx);
      if (rc != NB_OK) {
        j.state = NB_JOB_FAILED;
        continue;
      }
    }
    ++j.steps_done;
  }
  if (any_direct) {
    const int rc = nb_batch_sync(s->batch);
    if (rc != NB_OK) return rc;
  }
  // 4. retire
  for (auto it = s->running.begin(); it != s->running.end();) {
    Job& j = *s->jobs.at(*it);
    if (j.steps_done >= j.desc.steps || j.state == NB_JOB_FAILED) {
      retire(s, j);
      ++finished;
      it = s->running.erase(it);
    } else {
      ++it;
    }
  }
  if (running_out) *running_out = (int32_t)s->running.size();
  if (finished_out) *finished_out = finished;
  return NB_OK;
}

int nb_server_status(nb_server* s, int64_t id, nb_job_status* out) {
  if (!s || !out) return nb_fail(NB_E_ARG, "nb_server_status: null argument");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return nb_fail(NB_E_ARG, "nb_server_status: unknown job");
  const Job& j = *it->second;
  out->state = j.state;
  out->solver = j.solver;
  out->steps_done = j.steps_done;
  out->steps = j.desc.steps;
  out->t_submit = j.t_submit;
  out->t_admit = j.t_admit;
  out->t_done = j.t_done;
  return NB_OK;
}

int nb_server_result(nb_server* s, int64_t id, double* pos3, double* vel3) {
  if (!s) return nb_fail(NB_E_ARG, "nb_server_result: null server");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return nb_fail(NB_E_ARG, "nb_server_result: unknown job");
  const Job& j = *it->second;
  if (j.state != NB_JOB_DONE) return nb_fail(NB_E_ARG, "nb_server_result: job not finished");
  if (pos3) std::copy(j.pos.begin(), j.pos.end(), pos3);
  if (vel3) std::copy(j.vel.begin(), j.vel.end(), vel3);
  return NB_OK;
}

int nb_server_forget(nb_server* s, int64_t id) {
  if (!s) return nb_fail(NB_E_ARG, "nb_server_forget: null server");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return nb_fail(NB_E_ARG, "nb_server_forget: unknown job");
  if (it->second->state != NB_JOB_DONE && it->second->state != NB_JOB_FAILED)
    return nb_fail(NB_E_ARG, "nb_server_forget: job still queued or running");
  std::lock_guard<std::mutex> lk(s->mu);
  s->jobs.erase(it);
  return NB_OK;
}

}  // extern "C"
