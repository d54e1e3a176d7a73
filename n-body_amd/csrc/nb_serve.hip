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
//   2. start   selected jobs whose initial conditions are ready start running:
//              direct-sum jobs become clusters of the shared nb_batch pool;
//              Barnes-Hut jobs have their own nb_ctx (own stream), which the
//              worker that made their initial conditions already took from the
//              context pool (or created), uploaded, and captured the step of;
//   3. step    every running job gets one step, and the tick budget left over
//              goes to extra steps for the jobs with the least remaining work
//              (at most kMaxQuota per tick); direct-sum jobs step together in
//              one nb_batch_step (per-cluster quotas), Barnes-Hut jobs in one
//              nb_step each, all enqueued before any wait, so Barnes-Hut kernels
//              overlap with each other and with the batch;
//   4. retire  jobs that reached their step count copy their FP64 state to the
//              host and release their reservations; a Barnes-Hut job's copy runs
//              on a worker thread, so the loop never waits on a large transfer,
//              and its context goes back to the pool.
// Barnes-Hut contexts are sized in power-of-two capacity classes and rebound
// to each new job of their class (nb_ctx_rebind): a steady stream allocates
// and frees no device memory; idle pooled contexts are freed only when a new
// class does not fit the HBM budget.
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
  int32_t quota = 1;        // steps in the current tick
  double t_submit = 0, t_admit = 0, t_done = 0;
  // reservations held from selection to retirement
  int64_t bytes = 0;        // HBM budget (Barnes-Hut contexts)
  int64_t pool_bodies = 0;  // direct-sum pool bodies (padded)
  double cost_ms = 0;       // estimated device time of one step
  int32_t slot = -1;        // batch cluster id
  nb_ctx* ctx = nullptr;    // Barnes-Hut context (from the pool when `reused`)
  int64_t cap = 0;          // its capacity class
  bool reused = false;
  bool in_pool = false;     // its context went back to the pool (keeps its bytes)
  // initial conditions (made by a worker) and, later, the final state
  std::atomic<int> prepared{0};  // 0 pending, 1 ready, -1 failed
  std::atomic<int> finished{0};  // Barnes-Hut retirement on a worker: 1 done, -1 failed
  double t_fin = 0;              // written by that worker before `finished`
  std::vector<double> pos, vel;
  std::vector<float> mass;
};

// Capacity class of a Barnes-Hut job: the power of two at or above n (at least
// 4096), so pooled contexts fit the jobs of their class.
int64_t capacity(int64_t n) {
  int64_t c = int64_t(1) << 12;
  while (c < n) c <<= 1;
  return c;
}

// Device bytes of a Barnes-Hut context sized for n bodies (nb_create +
// nb_bh_alloc): FP64 state and forces, FP32 view and masses, reductions, and the
// single octree block.
int64_t bh_bytes(int64_t n) {
  const int64_t np = (n + 255) / 256 * 256;
  return 4 * 24 * n + np * (16 + 4) + 8 * (4 * 256 + np / 256 + 16) +
         nb_bh_layout_for(n).off[kBhBufCount] + (1 << 20);
}

int64_t padded(int64_t n) { return (n + kPadBodies - 1) / kPadBodies * kPadBodies; }

// Estimated device time of one step, from rates measured on MI355X: the
// compensated direct sum runs about 3.2e12 pair interactions/s once the device
// is full, and a Barnes-Hut step (theta = 0.5, build + traversal) costs about
// 1.7 ns per body plus some 0.35 ms of latency-bound launches (7.3 ms at
// 4,194,304 bodies, about 0.5 ms at 65,536).
double step_cost_ms(int64_t n, int32_t solver) {
  if (solver == NB_SOLVER_DIRECT) return 2e-3 + (double)n * (double)n / 3.2e9;
  return 0.35 + (double)n * 1.7e-6;
}

// Most steps a job takes in one tick (bounds the batch's replays per tick).
constexpr int kMaxQuota = 32;
// Shortest-first admits jobs up to this many ticks' budget of single steps.
constexpr double kOversubscribe = 4.0;

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
  std::vector<int64_t> retiring;   // Barnes-Hut jobs a worker is copying out
  std::vector<int64_t> retired;    // finished or failed since the last nb_server_poll
  // initial-condition workers
  std::mutex mu;
  std::condition_variable cv;        // workers wait for jobs to prepare
  std::condition_variable ready_cv;  // an idle tick waits for a prepared job
  std::deque<Job*> prep_queue;
  std::deque<Job*> retire_queue;
  bool stop = false;
  std::vector<std::thread> workers;
  // idle Barnes-Hut contexts by capacity class; workers return them (under mu),
  // the loop takes and frees them. Their bytes stay out of bytes_left.
  std::multimap<int64_t, nb_ctx*> ctx_pool;
  nb_server_stats stats{};

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Barnes-Hut context for a prepared job: the pooled context the loop handed
  // over, rebound to this job, or a new one of the job's capacity class; then
  // the upload and the step-graph capture, all on the job's own stream, so the
  // serving loop does not wait for them.
  int make_context(Job& j) {
    nb_config c{};
    c.n = j.desc.n;
    c.eps2 = j.desc.eps2;
    c.dt = j.desc.dt;
    c.G = j.desc.G;
    c.theta = cfg.theta;
    c.leaf_size = cfg.leaf_size;
    c.solver = NB_SOLVER_BH;
    nb_ctx* ctx = j.ctx;
    int rc = NB_OK;
    if (!ctx) {
      nb_config cc = c;
      cc.n = j.cap;
      rc = nb_create(&ctx, &cc);
      if (rc != NB_OK) return rc;
      j.ctx = ctx;
    }
    rc = nb_ctx_rebind(ctx, &c);
    if (rc == NB_OK)
      rc = nb_upload_f64(ctx, j.pos.data(), j.vel.data(), nullptr, j.mass.data(), j.desc.n);
    if (rc == NB_OK) rc = nb_prepare_step_graph(ctx);
    if (rc != NB_OK) return rc;  // the loop destroys j.ctx when it retires the failed job
    std::vector<double>().swap(j.pos);
    std::vector<double>().swap(j.vel);
    return NB_OK;
  }

  void worker() {
    for (;;) {
      Job* j = nullptr;
      bool out = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !retire_queue.empty() || !prep_queue.empty(); });
        if (stop) return;
        out = !retire_queue.empty();  // copies out first: they free device memory
        std::deque<Job*>& q = out ? retire_queue : prep_queue;
        j = q.front();
        q.pop_front();
      }
      const int64_t n = j->desc.n;
      if (out) {
        int rc = NB_OK;
        try {
          j->pos.resize(3 * n);
          j->vel.resize(3 * n);
          rc = nb_download_f64(j->ctx, j->pos.data(), j->vel.data(), nullptr);
        } catch (const std::bad_alloc&) {
          rc = NB_E_OOM;
        }
        j->t_fin = now();
        {
          std::lock_guard<std::mutex> lk(mu);
          if (rc == NB_OK) {  // the context serves the next job of its class
            ctx_pool.emplace(j->cap, j->ctx);
            j->ctx = nullptr;
            j->in_pool = true;
          }
          j->finished.store(rc == NB_OK ? 1 : -1);
        }
        ready_cv.notify_all();
        continue;
      }
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
        if (rc == 0 && j->solver == NB_SOLVER_BH) rc = make_context(*j) == NB_OK ? 0 : -1;
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

  void release_time(Job& j) {
    load_ms -= j.cost_ms;
    if (load_ms < 1e-9) load_ms = 0;
    j.cost_ms = 0;
  }
  void release(Job& j) {
    release_time(j);
    if (!j.in_pool) bytes_left += j.bytes;  // a pooled context keeps its bytes reserved
    pool_left += j.pool_bodies;
    j.bytes = j.pool_bodies = 0;
  }

  // Reserve the device memory of a Barnes-Hut job: an idle pooled context of
  // its class if there is one, else the bytes of a new one, freeing idle pooled
  // contexts of other classes when the budget needs them. False: no room yet.
  bool reserve_bh(Job& j) {
    j.cap = capacity(j.desc.n);
    j.bytes = bh_bytes(j.cap);
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = ctx_pool.find(j.cap);
      if (it != ctx_pool.end()) {
        j.ctx = it->second;
        j.reused = true;
        ctx_pool.erase(it);
        return true;
      }
    }
    while (j.bytes > bytes_left) {
      nb_ctx* victim = nullptr;
      int64_t vcap = 0;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (ctx_pool.empty()) break;
        auto it = std::prev(ctx_pool.end());  // the largest idle context
        vcap = it->first;
        victim = it->second;
        ctx_pool.erase(it);
      }
      nb_destroy(victim);
      bytes_left += bh_bytes(vcap);
    }
    if (j.bytes > bytes_left) {
      j.bytes = 0;
      return false;
    }
    bytes_left -= j.bytes;
    return true;
  }
};

namespace {

// Start a prepared job. A direct-sum job is uploaded into the shared pool here;
// NB_E_BUDGET means the pool has no free segment of that size yet
// (fragmentation), and the job stays prepared and is retried next tick. A
// Barnes-Hut job's context was created and uploaded by its worker.
int start(nb_server* s, Job& j) {
  if (j.solver == NB_SOLVER_DIRECT) {
    int32_t id = -1;
    const int rc = nb_batch_add(s->batch, j.pos.data(), j.vel.data(), nullptr, j.mass.data(),
                                j.desc.n, j.desc.eps2, j.desc.dt, j.desc.G, &id);
    if (rc != NB_OK) return rc;
    j.slot = id;
    std::vector<double>().swap(j.pos);
    std::vector<double>().swap(j.vel);
  }
  return j.ctx || j.solver == NB_SOLVER_DIRECT ? NB_OK : NB_E_HIP;
}

// Finish a job on the loop's thread (direct-sum jobs, failed jobs, and
// Barnes-Hut jobs whose worker has copied them out: `rc` is its result).
int retire(nb_server* s, Job& j, int rc = NB_OK) {
  const int64_t n = j.desc.n;
  if (j.solver == NB_SOLVER_DIRECT) {
    if (j.slot >= 0) {
      j.pos.resize(3 * n);
      j.vel.resize(3 * n);
      rc = nb_batch_download(s->batch, j.slot, j.pos.data(), j.vel.data(), nullptr);
      const int rr = nb_batch_remove(s->batch, j.slot);
      if (rc == NB_OK) rc = rr;
      j.slot = -1;
    }
  } else if (j.ctx) {  // a failed Barnes-Hut job: no copy, and its context is not reused
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
  j.t_done = j.t_fin > 0 ? j.t_fin : s->now();
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
  for (auto& kv : s->ctx_pool) nb_destroy(kv.second);
  if (s->batch) nb_batch_destroy(s->batch);
  delete s;
}

}  // namespace

extern "C" {

int nb_server_create(nb_server** out, const nb_server_config* cfg) {
  NB_TRY
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
  if (s->cfg.policy != NB_SERVE_ARRIVAL && s->cfg.policy != NB_SERVE_SHORTEST) {
    delete s;
    return nb_fail(NB_E_ARG, "nb_server_create: unknown policy");
  }
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
  try {
    for (int t = 0; t < s->cfg.ic_threads; ++t) s->workers.emplace_back([s] { s->worker(); });
  } catch (...) {  // stop and join the workers that did start
    free_server(s);
    return nb_fail(NB_E_HIP, "nb_server_create: cannot start the worker threads");
  }
  *out = s;
  return NB_OK;
  NB_CATCH("nb_server_create")
}

void nb_server_destroy(nb_server* s) { free_server(s); }

double nb_server_now(nb_server* s) { return s ? s->now() : 0.0; }

int nb_server_submit(nb_server* s, const nb_job* job, int64_t* id) {
  NB_TRY
  if (!s || !job || !id) return nb_fail(NB_E_ARG, "nb_server_submit: null argument");
  if (job->n < 1 || job->n >= (int64_t(1) << 27) || job->steps < 0 ||
      (job->ic != NB_IC_PLUMMER && job->ic != NB_IC_BALL) || !(job->eps2 >= 0.f))
    return nb_fail(NB_E_ARG, "nb_server_submit: bad job description");
  const int32_t solver = job->n >= s->cfg.bh_threshold ? NB_SOLVER_BH : NB_SOLVER_DIRECT;
  if (solver == NB_SOLVER_BH ? bh_bytes(capacity(job->n)) > s->budget_total
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
  NB_CATCH("nb_server_submit")
}

int nb_server_tick(nb_server* s, int32_t* running_out, int32_t* finished_out) {
  NB_TRY
  NB_RANGE("nb_server_tick");
  if (!s) return nb_fail(NB_E_ARG, "nb_server_tick: null server");
  int finished = 0;
  using clk = std::chrono::steady_clock;
  auto t_mark = clk::now();
  auto lap = [&](double& acc) {
    const auto t = clk::now();
    acc += std::chrono::duration<double, std::milli>(t - t_mark).count();
    t_mark = t;
  };
  s->stats.ticks += 1;
  // 0. Barnes-Hut jobs whose worker has copied them out are done
  auto collect = [&] {
    for (auto it = s->retiring.begin(); it != s->retiring.end();) {
      Job& j = *s->jobs.at(*it);
      const int f = j.finished.load();
      if (f == 0) {
        ++it;
        continue;
      }
      retire(s, j, f > 0 ? NB_OK : NB_E_HIP);
      s->retired.push_back(*it);
      ++finished;
      it = s->retiring.erase(it);
    }
  };
  collect();
  // 1. select among the first `lookahead` queued jobs, in arrival order or
  //    smallest estimated total work first (policy)
  {
    // (sort key, job): the key is the estimated total work under shortest-first
    // and 0 in arrival order; ids rise with arrival, so equal keys keep it. One
    // map lookup per candidate, not one per comparison (4096 candidates a tick).
    std::vector<std::pair<std::pair<double, int64_t>, Job*>> cand;
    const bool shortest = s->cfg.policy == NB_SERVE_SHORTEST;
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)cand.size() < s->cfg.lookahead; ++it) {
      Job* j = s->jobs.at(*it).get();
      cand.push_back({{shortest ? j->cost_ms * j->desc.steps : 0.0, *it}, j});
    }
    if (shortest)
      std::sort(cand.begin(), cand.end(),
                [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int64_t> taken;
    for (const auto& c : cand) {
      if ((int)(s->running.size() + s->preparing.size()) >= s->cfg.max_jobs) break;
      const int64_t id = c.first.second;
      Job& j = *c.second;
      const bool idle = s->running.empty() && s->preparing.empty();
      // shortest-first oversubscribes: admitted jobs beyond one tick's budget
      // wait resident (no steps) while jobs with less remaining work run
      const double gate = s->cfg.policy == NB_SERVE_SHORTEST ? kOversubscribe * s->cfg.tick_budget_ms
                                                             : s->cfg.tick_budget_ms;
      const bool fits_tick = idle || s->load_ms + j.cost_ms <= gate;
      if (!fits_tick) continue;  // a later, smaller job may fit
      if (j.solver == NB_SOLVER_BH) {
        if (!s->reserve_bh(j)) continue;
      } else {
        if (padded(j.desc.n) > s->pool_left) continue;
        j.pool_bodies = padded(j.desc.n);
        s->pool_left -= j.pool_bodies;
      }
      s->load_ms += j.cost_ms;
      s->preparing.push_back(id);
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->prep_queue.push_back(&j);
      }
      s->cv.notify_one();
      taken.push_back(id);
    }
    if (!taken.empty()) {
      std::sort(taken.begin(), taken.end());
      std::deque<int64_t> rest;
      for (int64_t id : s->queue)
        if (!std::binary_search(taken.begin(), taken.end(), id)) rest.push_back(id);
      s->queue.swap(rest);
    }
  }
  lap(s->stats.select_ms);
  // Idle with work on the workers: wait (up to 50 ms) for the first prepared
  // or copied-out job instead of returning an empty tick.
  if (s->running.empty() && (!s->preparing.empty() || !s->retiring.empty())) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->ready_cv.wait_for(lk, std::chrono::milliseconds(50), [&] {
        for (int64_t id : s->preparing)
          if (s->jobs.at(id)->prepared.load() != 0) return true;
        for (int64_t id : s->retiring)
          if (s->jobs.at(id)->finished.load() != 0) return true;
        return false;
      });
    }
    collect();
  }
  lap(s->stats.idle_wait_ms);
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
      s->retired.push_back(*it);
      ++finished;
    } else {
      j.state = NB_JOB_RUNNING;
      j.t_admit = s->now();
      s->running.push_back(*it);
    }
    it = s->preparing.erase(it);
  }
  lap(s->stats.start_ms);
  // 3. steps, least remaining work first. ARRIVAL: one for every running job,
  //    and the tick budget their single steps leave goes to extra steps, so
  //    short jobs finish in a few ticks. SHORTEST (preemptive shortest remaining
  //    work first): the tick budget goes to the jobs in that order, as many
  //    steps each as fit, and a job it does not reach waits this tick (the
  //    first always steps). Everything is enqueued before any wait.
  std::vector<Job*> act;
  double base_ms = 0;
  const bool srpt = s->cfg.policy == NB_SERVE_SHORTEST;
  for (int64_t id : s->running) {
    Job& j = *s->jobs.at(id);
    if (j.steps_done >= j.desc.steps) continue;
    j.quota = srpt ? 0 : 1;
    base_ms += srpt ? 0.0 : j.cost_ms;
    act.push_back(&j);
  }
  std::stable_sort(act.begin(), act.end(), [](const Job* a, const Job* b) {
    return a->cost_ms * (a->desc.steps - a->steps_done) < b->cost_ms * (b->desc.steps - b->steps_done);
  });
  double spare_ms = s->cfg.tick_budget_ms - base_ms;
  for (Job* j : act) {
    if (spare_ms <= 0 && !(srpt && j == act.front())) break;
    const int left = j->desc.steps - j->steps_done;
    const int have = j->quota;
    int add = (int)std::min<double>({(double)(left - have), (double)(kMaxQuota - have),
                                     std::floor(spare_ms / j->cost_ms)});
    if (srpt && j == act.front()) add = std::max(add, 1);
    if (add <= 0) continue;
    j->quota += add;
    spare_ms -= add * j->cost_ms;
  }
  std::vector<int32_t> qid, qsteps;
  int batch_steps = 0;
  for (Job* j : act) {
    if (j->solver == NB_SOLVER_DIRECT) {
      qid.push_back(j->slot);
      qsteps.push_back(j->quota);
      batch_steps = std::max(batch_steps, j->quota);
    } else if (j->quota > 0 && nb_step(j->ctx, j->quota) != NB_OK) {
      j->state = NB_JOB_FAILED;  // this job alone; it is retired below
    }
  }
  if (batch_steps > 0) {
    int rc = nb_batch_set_quotas(s->batch, qid.data(), qsteps.data(), (int32_t)qid.size());
    if (rc == NB_OK) rc = nb_batch_step(s->batch, batch_steps);
    if (rc != NB_OK) return rc;
  }
  lap(s->stats.enqueue_ms);
  for (Job* j : act) {
    s->stats.steps += j->quota;
    if (j->state == NB_JOB_FAILED || j->quota == 0) continue;
    if (j->solver == NB_SOLVER_BH) {
      const int rc = nb_sync(j->ctx);
      if (rc != NB_OK) {
        j->state = NB_JOB_FAILED;
        continue;
      }
    }
    j->steps_done += j->quota;
  }
  const bool any_direct = batch_steps > 0;
  if (any_direct) {
    const int rc = nb_batch_sync(s->batch);
    if (rc != NB_OK) return rc;
  }
  lap(s->stats.device_wait_ms);
  // 4. retire: direct-sum and failed jobs here, Barnes-Hut jobs on a worker
  for (auto it = s->running.begin(); it != s->running.end();) {
    Job& j = *s->jobs.at(*it);
    if (j.steps_done >= j.desc.steps || j.state == NB_JOB_FAILED) {
      if (j.solver == NB_SOLVER_BH && j.state != NB_JOB_FAILED) {
        s->release_time(j);  // off the device's schedule; its memory until the copy ends
        s->retiring.push_back(*it);
        {
          std::lock_guard<std::mutex> lk(s->mu);
          s->retire_queue.push_back(&j);
        }
        s->cv.notify_one();
      } else {
        retire(s, j);
        s->retired.push_back(*it);
        ++finished;
      }
      it = s->running.erase(it);
    } else {
      ++it;
    }
  }
  lap(s->stats.retire_ms);
  if (running_out) *running_out = (int32_t)s->running.size();
  if (finished_out) *finished_out = finished;
  return NB_OK;
  NB_CATCH("nb_server_tick")
}

int nb_server_get_stats(nb_server* s, nb_server_stats* out) {
  NB_TRY
  if (!s || !out) return nb_fail(NB_E_ARG, "nb_server_get_stats: null argument");
  *out = s->stats;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    out->pooled_contexts = (int64_t)s->ctx_pool.size();
  }
  return NB_OK;
  NB_CATCH("nb_server_get_stats")
}

int nb_server_poll(nb_server* s, int64_t* ids, int32_t cap, int32_t* count) {
  NB_TRY
  if (!s || !count || (cap > 0 && !ids)) return nb_fail(NB_E_ARG, "nb_server_poll: bad argument");
  const int32_t k = (int32_t)std::min<size_t>(cap > 0 ? (size_t)cap : 0, s->retired.size());
  std::copy(s->retired.begin(), s->retired.begin() + k, ids);
  s->retired.erase(s->retired.begin(), s->retired.begin() + k);
  *count = k;
  return NB_OK;
  NB_CATCH("nb_server_poll")
}

int nb_server_status(nb_server* s, int64_t id, nb_job_status* out) {
  NB_TRY
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
  NB_CATCH("nb_server_status")
}

int nb_server_result(nb_server* s, int64_t id, double* pos3, double* vel3) {
  NB_TRY
  if (!s) return nb_fail(NB_E_ARG, "nb_server_result: null server");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return nb_fail(NB_E_ARG, "nb_server_result: unknown job");
  const Job& j = *it->second;
  if (j.state != NB_JOB_DONE) return nb_fail(NB_E_ARG, "nb_server_result: job not finished");
  if (pos3) std::copy(j.pos.begin(), j.pos.end(), pos3);
  if (vel3) std::copy(j.vel.begin(), j.vel.end(), vel3);
  return NB_OK;
  NB_CATCH("nb_server_result")
}

int nb_server_forget(nb_server* s, int64_t id) {
  NB_TRY
  if (!s) return nb_fail(NB_E_ARG, "nb_server_forget: null server");
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return nb_fail(NB_E_ARG, "nb_server_forget: unknown job");
  if (it->second->state != NB_JOB_DONE && it->second->state != NB_JOB_FAILED)
    return nb_fail(NB_E_ARG, "nb_server_forget: job still queued or running");
  std::lock_guard<std::mutex> lk(s->mu);
  s->jobs.erase(it);
  return NB_OK;
  NB_CATCH("nb_server_forget")
}

}  // extern "C"
