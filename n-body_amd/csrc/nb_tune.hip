// nb_tune.hip — the autotuner of nbody.h (nb_autotune / nb_get_tuning).
//
// Candidates are timed with HIP events on a context's own stream; the winners
// go into a process-wide table indexed by the power-of-two size class of n,
// which nb_create consults when a config leaves the shape fields at 0.
#include <cfloat>
#include <mutex>
#include <vector>

#include "nb_internal.h"
#include "nbody_ic.h"

namespace {

constexpr int kClasses = 31;
std::mutex g_mu;
nb_tuning g_table[kClasses] = {};

int size_class(int64_t n) {
  int c = 0;
  while (c < kClasses - 1 && (int64_t(1) << c) < n) ++c;
  return c;
}

struct Input {
  std::vector<double> pos, vel;
  std::vector<float> mass;
};

Input plummer(int64_t n) {
  Input in;
  in.pos.resize(3 * n);
  in.vel.resize(3 * n);
  in.mass.resize(n);
  nb_ic_plummer(20240, n, 50.0, 1e4, 10.0, 1.0, in.pos.data(), in.vel.data(), in.mass.data());
  return in;
}

// Mean device time of `reps` force passes (after one warm-up pass).
int time_forces(const nb_config& cfg, const Input& in, int reps, double* ms) {
  nb_ctx* c = nullptr;
  int rc = nb_create(&c, &cfg);
  if (rc) return rc;
  rc = nb_upload_f64(c, in.pos.data(), in.vel.data(), nullptr, in.mass.data(), cfg.n);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = nb_fail(NB_E_HIP, "nb_autotune: event creation failed");
  for (int r = -1; !rc && r < reps; ++r) {
    if (r == 0) (void)hipEventRecord(e0, c->stream);
    hipError_t e;
    if (cfg.solver == NB_SOLVER_BH) {
      e = nb_bh_build(c);
      if (e == hipSuccess) e = nb_bh_traverse(c, false);
    } else {
      e = nb_launch_direct(c);
    }
    if (e != hipSuccess) rc = nb_fail(NB_E_HIP, std::string("nb_autotune: ") + hipGetErrorString(e));
  }
  if (!rc) {
    (void)hipEventRecord(e1, c->stream);
    float t = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      rc = nb_fail(NB_E_HIP, "nb_autotune: timing failed");
    *ms = t / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  nb_destroy(c);
  return rc;
}

}  // namespace

// Consulted by nb_create / nb_batch_create: fills zero shape fields of the
// class of n from the table; returns false if nothing was tuned for it.
bool nb_tuned_shape(int64_t n, int* bpl, int* jsplit, int* leaf) {
  std::lock_guard<std::mutex> lk(g_mu);
  const nb_tuning& t = g_table[size_class(n)];
  if (t.n == 0) return false;
  if (bpl && *bpl == 0) *bpl = t.bodies_per_lane;
  if (jsplit && *jsplit == 0) *jsplit = t.j_split;
  if (leaf && *leaf == 0) *leaf = t.leaf_size;
  return true;
}

extern "C" int nb_autotune(int64_t n, int32_t what, int32_t reps, nb_tuning* out) {
  NB_TRY
  NB_RANGE("nb_autotune");
  if (n < 1 || n >= (int64_t(1) << 27)) return nb_fail(NB_E_ARG, "nb_autotune: n out of range");
  if (reps < 1) reps = 3;
  const Input in = plummer(n);
  nb_tuning best{};
  best.n = n;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    const nb_tuning& t = g_table[size_class(n)];
    best.leaf_size = t.leaf_size;
    best.bodies_per_lane = t.bodies_per_lane;
    best.j_split = t.j_split;
  }
  if (what & NB_TUNE_DIRECT) {
    double best_ms = DBL_MAX;
    for (int b : {1, 2, 4})
      for (int k : {1, 2, 4, 8, 16}) {
        if (!nb_valid_shape(b, k)) continue;
        nb_config cfg{};
        cfg.n = n;
        cfg.eps2 = 40.f;
        cfg.dt = 0.0035f;
        cfg.solver = NB_SOLVER_DIRECT;
        cfg.bodies_per_lane = b;
        cfg.j_split = k;
        double ms = 0;
        const int rc = time_forces(cfg, in, reps, &ms);
        if (rc) return rc;
        if (ms < best_ms) {
          best_ms = ms;
          best.bodies_per_lane = b;
          best.j_split = k;
          best.direct_ms = ms;
        }
      }
  }
  if (what & NB_TUNE_BH) {
    double best_ms = DBL_MAX;
    for (int leaf : {1, 2, 4, 8, 16, 32}) {
      nb_config cfg{};
      cfg.n = n;
      cfg.eps2 = 40.f;
      cfg.dt = 0.0035f;
      cfg.solver = NB_SOLVER_BH;
      cfg.leaf_size = leaf;
      double ms = 0;
      const int rc = time_forces(cfg, in, reps, &ms);
      if (rc) return rc;
      if (ms < best_ms) {
        best_ms = ms;
        best.leaf_size = leaf;
        best.bh_ms = ms;
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[size_class(n)] = best;
  }
  if (out) *out = best;
  return NB_OK;
  NB_CATCH("nb_autotune")
}

extern "C" int nb_tune_leaf(nb_ctx* c, int32_t reps, nb_tuning* out) {
  NB_TRY
  NB_RANGE("nb_tune_leaf");
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_tune_leaf: null context");
  if (c->cfg.solver != NB_SOLVER_BH) return nb_fail(NB_E_ARG, "nb_tune_leaf: not a Barnes-Hut context");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_tune_leaf: nothing uploaded");
  if (reps < 1) reps = 3;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return nb_fail(NB_E_HIP, "nb_tune_leaf: event creation failed");
  }
  // Candidates are timed in interleaved rounds (one pass of each per round, the
  // first round untimed) and each keeps its fastest pass, so clock and
  // neighbour drift during the sweep affects every candidate alike.
  // Each leaf size is timed with sibling ranges walked in pairs
  // (NB_F_BH_PAIRS) and alone (NB_F_BH_SINGLE_RANGES).
  constexpr int kLeaves[] = {1, 2, 4, 8, 16, 32};
  constexpr int kLeafCount = sizeof(kLeaves) / sizeof(kLeaves[0]);
  constexpr int kCount = 2 * kLeafCount;
  const int flags0 = c->cfg.flags & ~(NB_F_BH_SINGLE_RANGES | NB_F_BH_PAIRS);
  double ms[kCount];
  for (double& m : ms) m = DBL_MAX;
  int rc = NB_OK, best_leaf = c->cfg.leaf_size, best_single = 0;
  double best_ms = DBL_MAX;
  for (int r = -1; !rc && r < reps; ++r) {
    for (int k = 0; !rc && k < kCount; ++k) {
      c->cfg.leaf_size = kLeaves[k % kLeafCount];
      c->cfg.flags = flags0 | (k >= kLeafCount ? NB_F_BH_SINGLE_RANGES : NB_F_BH_PAIRS);
      (void)hipEventRecord(e0, c->stream);
      hipError_t e = nb_bh_build(c);
      if (e == hipSuccess) e = nb_bh_traverse(c, false);
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) rc = nb_fail(NB_E_HIP, std::string("nb_tune_leaf: ") + hipGetErrorString(e));
      else if (r >= 0 && t < ms[k]) ms[k] = t;
    }
  }
  for (int k = 0; !rc && k < kCount; ++k)
    if (ms[k] < best_ms) {
      best_ms = ms[k];
      best_leaf = kLeaves[k % kLeafCount];
      best_single = k >= kLeafCount ? 1 : 0;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->cfg.leaf_size = best_leaf;
  c->cfg.flags = flags0 | (best_single ? NB_F_BH_SINGLE_RANGES : NB_F_BH_PAIRS);
  // the captured step graph holds the old leaf size; the tree and forces of the
  // last timed pass are the last candidate's, so the chosen one runs once more
  nb_drop_graph(c->stream, c->graph_exec, c->graph);
  if (rc) return rc;
  NB_HIP(nb_bh_build(c));
  NB_HIP(nb_bh_traverse(c, false));
  NB_HIP(hipStreamSynchronize(c->stream));
  if (out) {
    *out = nb_tuning{};
    out->n = c->n;
    out->leaf_size = best_leaf;
    out->bh_single_ranges = best_single;
    out->bh_ms = best_ms;
  }
  return NB_OK;
  NB_CATCH("nb_tune_leaf")
}

extern "C" int nb_get_tuning(int64_t n, nb_tuning* out) {
  NB_TRY
  if (!out || n < 1) return nb_fail(NB_E_ARG, "nb_get_tuning: bad argument");
  nb_config cfg{};
  cfg.n = n;
  *out = nb_tuning{};
  out->n = n;
  int bpl = 0, js = 0, leaf = 0;
  nb_tuned_shape(n, &bpl, &js, &leaf);
  nb_default_shape(n, &bpl, &js, &leaf);
  out->bodies_per_lane = bpl;
  out->j_split = js;
  out->leaf_size = leaf;
  return NB_OK;
  NB_CATCH("nb_get_tuning")
}
