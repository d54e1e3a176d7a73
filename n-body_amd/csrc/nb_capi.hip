// nb_capi.hip — the C ABI of nbody.h: context lifetime, device memory,
// host<->device copies, the step sequence and its hipGraph, diagnostics.
//
// Replaces the reference's GL plumbing: SSBO creation at bindings 0-4
// (ParticleSystem.cpp:73-83, OpenGLBuffer.cpp:12-60), the per-step dispatch
// sequences of ParticleSolverGPU.cpp:31-41 and ParticleSolverBHutGPU.cpp:25-61,
// and buffer read-back (OpenGLBuffer.cpp:135-165).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "nb_internal.h"

static thread_local std::string g_last_error;

void nb_set_error(const std::string& msg) { g_last_error = msg; }
int nb_fail(int code, const std::string& msg) {
  try {
    g_last_error = msg;
  } catch (...) {  // keep the code; the message is lost only when the copy cannot be made
  }
  return code;
}

extern "C" const char* nb_last_error(void) { return g_last_error.c_str(); }

std::shared_mutex& nb_capture_mutex() {
  static std::shared_mutex m;
  return m;
}

namespace {

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

void free_ctx(nb_ctx* c) {
  if (!c) return;
  nb_drop_graph(c->stream, c->graph_exec, c->graph);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  nb_bh_free(c);
  if (c->tmp) (void)hipFree(c->tmp);
  (void)hipFree(c->pos);
  (void)hipFree(c->vel);
  (void)hipFree(c->acc);
  (void)hipFree(c->force);
  (void)hipFree(c->mass);
  (void)hipFree(c->pm);
  (void)hipFree(c->flags);
  (void)hipFree(c->red);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int check_flags(nb_ctx* c) {
  int h = 0;
  NB_HIP(hipMemcpyAsync(&h, c->flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  if (h & 1) return nb_fail(NB_E_NONFINITE, "non-finite position or velocity in the particle state");
  if (h & 2) return nb_fail(NB_E_TREE_OVERFLOW, "octree node pool exhausted");
  if (h & 4) return nb_fail(NB_E_TREE_OVERFLOW, "octree traversal left the node range");
  if (h & 8) return nb_fail(NB_E_TREE_OVERFLOW, "octree traversal stack overflow");
  return NB_OK;
}


// Structure-of-arrays host copy [k][slot] -> [k][caller index] after nb_reorder.
template <typename T>
void to_caller_order(const nb_ctx* c, std::vector<T>& a, int comps) {
  if (c->order.empty()) return;
  const size_t n = c->n;
  std::vector<T> out(a.size());
  for (int k = 0; k < comps; ++k)
    for (size_t s = 0; s < n; ++s) out[k * n + (size_t)c->order[s]] = a[k * n + s];
  a.swap(out);
}

}  // namespace

// Built-in launch shapes, from a sweep on MI355X (Plummer spheres, force pass):
// one body per lane; 16 waves per i-block up to 16384 bodies (4096 bodies:
// 30 us against 83 us at 4 waves), 8 above (2^15: 0.41 ms against 0.61 ms;
// 2^18: 19.8 ms against 20.2 ms).
void nb_default_shape(int64_t n, int* bpl, int* jsplit, int* leaf) {
  if (bpl && *bpl == 0) *bpl = 1;
  if (jsplit && *jsplit == 0) {
    int k = n <= 16384 ? 16 : 8;
    if (bpl && *bpl == 2) k = 8;
    if (bpl && *bpl == 4) k = 4;
    *jsplit = k;
  }
  if (leaf && *leaf == 0) *leaf = 8;
}

// Valid direct-sum launch shapes (see launch_b in nb_direct.hip).
bool nb_valid_shape(int bpl, int jsplit) {
  switch (bpl) {
    case 1: return jsplit == 1 || jsplit == 2 || jsplit == 4 || jsplit == 8 || jsplit == 16;
    case 2: return jsplit == 1 || jsplit == 2 || jsplit == 4 || jsplit == 8;
    case 4: return jsplit == 1 || jsplit == 2 || jsplit == 4;
    default: return false;
  }
}

namespace {

void default_direct_shape(nb_ctx* c) {
  int bpl = c->cfg.bodies_per_lane, js = c->cfg.j_split, leaf = c->cfg.leaf_size;
  nb_tuned_shape(c->cfg.n, &bpl, &js, &leaf);
  nb_default_shape(c->cfg.n, &bpl, &js, &leaf);
  if (!nb_valid_shape(bpl, js)) js = std::min(js, bpl == 1 ? 16 : bpl == 2 ? 8 : 4);
  c->bpl = bpl;
  c->jsplit = js;
  c->cfg.leaf_size = leaf;
}

}  // namespace

hipError_t nb_enqueue_step(nb_ctx* c) {
  hipError_t e;
  const bool t = c->timing;
  if (t) (void)hipEventRecord(c->ev[0], c->stream);
  if (c->cfg.solver == NB_SOLVER_BH) {
    e = nb_bh_build(c);
    if (e != hipSuccess) return e;
    if (t) (void)hipEventRecord(c->ev[1], c->stream);
    e = nb_bh_traverse(c, false);
  } else {
    if (t) (void)hipEventRecord(c->ev[1], c->stream);
    e = nb_launch_direct(c);
  }
  if (e != hipSuccess) return e;
  if (t) (void)hipEventRecord(c->ev[2], c->stream);
  e = nb_launch_integrate(c);
  if (t) (void)hipEventRecord(c->ev[3], c->stream);
  return e;
}

static int accumulate_timing(nb_ctx* c) {
  NB_HIP(hipEventSynchronize(c->ev[3]));
  float build = 0.f, force = 0.f, step = 0.f;
  NB_HIP(hipEventElapsedTime(&build, c->ev[0], c->ev[1]));
  NB_HIP(hipEventElapsedTime(&force, c->ev[1], c->ev[2]));
  NB_HIP(hipEventElapsedTime(&step, c->ev[0], c->ev[3]));
  c->tim.force_ms += force;
  c->tim.force_calls += 1;
  c->tim.step_ms += step;
  c->tim.step_calls += 1;
  if (c->cfg.solver == NB_SOLVER_BH) {
    c->tim.build_ms += build;
    c->tim.build_calls += 1;
  }
  return NB_OK;
}

// Direct-sum systems up to this size capture kChunkSteps steps per graph: at
// N = 4096 a step replayed from a one-step graph took 0.0333 ms and from an
// eight-step graph 0.0285 ms (N = 16384: 4% less; Barnes-Hut steps, bound by
// their dependent build kernels, gained under 1%).
constexpr int64_t kChunkMaxBodies = 32768;
constexpr int kChunkSteps = 8;

int nb_prepare_step_graph(nb_ctx* c) {
  if (c->graph_exec || (c->cfg.flags & NB_F_NO_GRAPH)) return NB_OK;
  NB_API_CAPTURE;
  const int steps =
      c->cfg.solver == NB_SOLVER_DIRECT && c->n <= kChunkMaxBodies ? kChunkSteps : 1;
  NB_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int s = 0; s < steps && e == hipSuccess; ++s) e = nb_enqueue_step(c);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return nb_fail(NB_E_HIP, std::string("nb_step: graph capture failed: ") +
                                 hipGetErrorString(e != hipSuccess ? e : e2));
  }
  const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    c->graph_exec = nullptr;
    return nb_fail(NB_E_HIP, std::string("nb_step: graph instantiation failed: ") +
                                 hipGetErrorString(ei));
  }
  c->graph = g;
  c->graph_steps = steps;
  return NB_OK;
}

int nb_ctx_rebind(nb_ctx* c, const nb_config* cfg) {
  NB_TRY
  NB_API_SHARED;
  if (!c || !cfg || c->cfg.solver != NB_SOLVER_BH || cfg->solver != NB_SOLVER_BH)
    return nb_fail(NB_E_ARG, "nb_ctx_rebind: Barnes-Hut contexts only");
  if (cfg->n < 1 || cfg->n > c->cap_n) return nb_fail(NB_E_ARG, "nb_ctx_rebind: n exceeds the capacity");
  nb_drop_graph(c->stream, c->graph_exec, c->graph);
  NB_HIP(hipStreamSynchronize(c->stream));
  c->cfg = *cfg;
  c->n = (int)cfg->n;
  c->n_pad = (int)((cfg->n + kPadBodies - 1) / kPadBodies * kPadBodies);
  c->G = cfg->G == 0.f ? 1.0 : (double)cfg->G;
  c->cfg.theta = cfg->theta == 0.f ? 0.5f : cfg->theta;
  c->theta2 = c->cfg.theta * c->cfg.theta;
  c->guard = !(cfg->eps2 >= 1e-20f);
  if (c->cfg.leaf_size == 0) {
    int bpl = 0, js = 0, leaf = 0;
    nb_tuned_shape(c->cfg.n, &bpl, &js, &leaf);
    nb_default_shape(c->cfg.n, &bpl, &js, &leaf);
    c->cfg.leaf_size = leaf;
  }
  void* block = c->bh.block;
  c->bh = nb_bh_state{};
  c->bh.block = block;
  nb_bh_carve(c, static_cast<char*>(block), kBhBufCount);
  NB_HIP(nb_launch_fill_i32(reinterpret_cast<int*>(c->bh.pms + c->n), 0, 4 * kSpare, c->stream));
  NB_HIP(nb_launch_fill_i32(c->flags, 0, 4, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  c->uploaded = false;
  c->tree_valid = false;
  c->order.clear();
  c->tim = nb_timing{};
  return NB_OK;
  NB_CATCH("nb_ctx_rebind")
}

extern "C" {

int nb_create(nb_ctx** out, const nb_config* cfg) {
  NB_TRY
  NB_RANGE("nb_create");
  NB_API_SHARED;
  if (!out || !cfg) return nb_fail(NB_E_ARG, "nb_create: null argument");
  *out = nullptr;
  if (cfg->n < 1 || cfg->n >= (int64_t(1) << 30))
    return nb_fail(NB_E_ARG, "nb_create: n must be in [1, 2^30)");
  if (!(cfg->eps2 >= 0.f) || !std::isfinite(cfg->eps2))
    return nb_fail(NB_E_ARG, "nb_create: eps2 must be finite and >= 0");
  if (!std::isfinite(cfg->dt)) return nb_fail(NB_E_ARG, "nb_create: dt must be finite");
  if (cfg->solver != NB_SOLVER_DIRECT && cfg->solver != NB_SOLVER_BH)
    return nb_fail(NB_E_ARG, "nb_create: unknown solver");
  if (cfg->solver == NB_SOLVER_BH && cfg->n >= (int64_t(1) << 27))
    return nb_fail(NB_E_ARG, "nb_create: Barnes-Hut supports n < 2^27");
  // The direct-sum and integration kernels address the [3][n] FP64 state with
  // 32-bit indices (2 n + i), so 3 n must stay below 2^31.
  constexpr int64_t kMaxDirectBodies = ((int64_t(1) << 31) - 1) / 3;  // 715827882
  if (cfg->solver == NB_SOLVER_DIRECT && cfg->n > kMaxDirectBodies)
    return nb_fail(NB_E_ARG, "nb_create: the direct sum supports n <= 715827882");
  if (cfg->integrator != NB_INTEG_KDK_LAGGED && cfg->integrator != NB_INTEG_EULER)
    return nb_fail(NB_E_ARG, "nb_create: unknown integrator");
  if (cfg->leaf_size < 0 || cfg->leaf_size > 64)
    return nb_fail(NB_E_ARG, "nb_create: leaf_size must be in [0, 64]");
  if (cfg->bodies_per_lane != 0 && cfg->bodies_per_lane != 1 && cfg->bodies_per_lane != 2 &&
      cfg->bodies_per_lane != 4)
    return nb_fail(NB_E_ARG, "nb_create: bodies_per_lane must be 0, 1, 2 or 4");
  if (cfg->j_split != 0 && cfg->j_split != 1 && cfg->j_split != 2 && cfg->j_split != 4 &&
      cfg->j_split != 8 && cfg->j_split != 16)
    return nb_fail(NB_E_ARG, "nb_create: j_split must be 0, 1, 2, 4, 8 or 16");
  if (cfg->bodies_per_lane != 0 && cfg->j_split != 0 &&
      !nb_valid_shape(cfg->bodies_per_lane, cfg->j_split))
    return nb_fail(NB_E_ARG, "nb_create: j_split 8 needs bodies_per_lane <= 2, 16 needs 1");
  if (cfg->theta < 0.f || !std::isfinite(cfg->theta))
    return nb_fail(NB_E_ARG, "nb_create: theta must be finite and >= 0");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return nb_fail(NB_E_NOGPU, "nb_create: no HIP device visible");

  nb_ctx* c = new (std::nothrow) nb_ctx();
  if (!c) return nb_fail(NB_E_OOM, "nb_create: host allocation failed");
  c->cfg = *cfg;
  c->cap_n = cfg->n;
  c->n = (int)cfg->n;
  c->n_pad = (int)((cfg->n + kPadBodies - 1) / kPadBodies * kPadBodies);
  c->G = cfg->G == 0.f ? 1.0 : (double)cfg->G;
  const float theta = cfg->theta == 0.f ? 0.5f : cfg->theta;
  c->cfg.theta = theta;
  c->theta2 = theta * theta;
  // The self pair gives dx = 0; its term is exactly 0 unless eps2 is so small
  // that (eps2)^{-3/2} overflows FP32. Below that, use the guarded kernel.
  c->guard = !(cfg->eps2 >= 1e-20f);
  default_direct_shape(c);

  (void)hipGetDevice(&c->device);
  int rc = NB_OK;
  auto fail = [&](int code, const std::string& m) {
    free_ctx(c);
    return nb_fail(code, m);
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(NB_E_HIP, "nb_create: hipStreamCreate failed");
  const size_t n = c->n, np = c->n_pad;
  if (dalloc(&c->pos, 3 * n) || dalloc(&c->vel, 3 * n) || dalloc(&c->acc, 3 * n) ||
      dalloc(&c->force, 3 * n) || dalloc(&c->mass, np) || dalloc(&c->pm, np) ||
      dalloc(&c->flags, 4))
    return fail(NB_E_OOM, "nb_create: device allocation failed (particle state)");
  c->red_blocks = 256;
  if (dalloc(&c->red, 4 * (size_t)c->red_blocks + (np / 256) + 16))
    return fail(NB_E_OOM, "nb_create: device allocation failed (reductions)");
  if (hipMemset(c->flags, 0, 4 * sizeof(int)) != hipSuccess ||
      hipMemset(c->mass, 0, np * sizeof(float)) != hipSuccess ||
      hipMemset(c->pm, 0, np * sizeof(float4)) != hipSuccess ||
      hipMemset(c->force, 0, 3 * n * sizeof(double)) != hipSuccess)
    return fail(NB_E_HIP, "nb_create: hipMemset failed");
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(NB_E_HIP, "nb_create: hipEventCreate failed");
  if (c->cfg.solver == NB_SOLVER_BH) {
    rc = nb_bh_alloc(c);
    if (rc != NB_OK) {
      std::string m = g_last_error;
      return fail(rc, m);
    }
  }
  *out = c;
  return NB_OK;
  NB_CATCH("nb_create")
}

void nb_destroy(nb_ctx* ctx) {
  NB_API_SHARED;
  if (!ctx) return;
  (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
}

static int upload_common(nb_ctx* c, const std::vector<double>& p, const std::vector<double>& v,
                         const std::vector<double>& a, const float* mass) {
  const size_t n = c->n;
  // Drop any captured step: it would be replayed against stale assumptions otherwise.
  nb_drop_graph(c->stream, c->graph_exec, c->graph);
  NB_HIP(hipMemcpyAsync(c->pos, p.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  NB_HIP(hipMemcpyAsync(c->vel, v.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  NB_HIP(hipMemcpyAsync(c->acc, a.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  NB_HIP(hipMemcpyAsync(c->mass, mass, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  NB_HIP(nb_launch_fill_i32(c->flags, 0, 4, c->stream));
  NB_HIP(nb_launch_refresh_pm(c));
  NB_HIP(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  c->tree_valid = false;
  c->order.clear();  // slots are in the caller's order again
  return NB_OK;
}

int nb_upload(nb_ctx* c, const float* pos4, const float* vel4, const float* acc4,
              const float* mass, int64_t n) {
  NB_TRY
  NB_RANGE("nb_upload");
  NB_API_SHARED;
  if (!c || !pos4 || !vel4 || !mass) return nb_fail(NB_E_ARG, "nb_upload: null argument");
  if (n != c->n) return nb_fail(NB_E_ARG, "nb_upload: n differs from the context's n");
  std::vector<double> p(3 * n), v(3 * n), a(3 * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      p[k * n + i] = pos4[4 * i + k];
      v[k * n + i] = vel4[4 * i + k];
      if (acc4) a[k * n + i] = acc4[4 * i + k];
    }
  return upload_common(c, p, v, a, mass);
  NB_CATCH("nb_upload")
}

int nb_upload_f64(nb_ctx* c, const double* pos3, const double* vel3, const double* acc3,
                  const float* mass, int64_t n) {
  NB_TRY
  NB_RANGE("nb_upload_f64");
  NB_API_SHARED;
  if (!c || !pos3 || !vel3 || !mass) return nb_fail(NB_E_ARG, "nb_upload_f64: null argument");
  if (n != c->n) return nb_fail(NB_E_ARG, "nb_upload_f64: n differs from the context's n");
  std::vector<double> p(3 * n), v(3 * n), a(3 * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      p[k * n + i] = pos3[3 * i + k];
      v[k * n + i] = vel3[3 * i + k];
      if (acc3) a[k * n + i] = acc3[3 * i + k];
    }
  return upload_common(c, p, v, a, mass);
  NB_CATCH("nb_upload_f64")
}

int nb_step(nb_ctx* c, int32_t nsteps) {
  NB_TRY
  NB_RANGE("nb_step");
  if (!c) return nb_fail(NB_E_ARG, "nb_step: null context");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_step: no particle system uploaded");
  if (nsteps < 0) return nb_fail(NB_E_ARG, "nb_step: nsteps < 0");
  if (nsteps == 0) return NB_OK;
  if (c->timing) {
    // Timed steps run eagerly with HIP events around each phase.
    NB_API_SHARED;
    for (int s = 0; s < nsteps; ++s) {
      NB_HIP(nb_enqueue_step(c));
      int rc = accumulate_timing(c);
      if (rc) return rc;
    }
    c->tree_valid = c->cfg.solver == NB_SOLVER_BH;
    return NB_OK;
  }
  // The step is captured once per context and replayed (cheap launches for the
  // ~47-kernel Barnes-Hut step); graph and eager launches give identical results.
  const bool use_graph = !(c->cfg.flags & NB_F_NO_GRAPH);
  if (!use_graph) {
    for (int s = 0; s < nsteps; ++s) NB_HIP(nb_enqueue_step(c));
  } else {
    const int rc = nb_prepare_step_graph(c);
    if (rc) return rc;
    // whole graphs, then the remainder as eager launches (identical results)
    int s = 0;
    for (; s + c->graph_steps <= nsteps; s += c->graph_steps)
      NB_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    for (; s < nsteps; ++s) NB_HIP(nb_enqueue_step(c));
  }
  c->tree_valid = c->cfg.solver == NB_SOLVER_BH;
  return NB_OK;
  NB_CATCH("nb_step")
}

int nb_sync(nb_ctx* c) {
  NB_TRY
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_sync: null context");
  NB_HIP(hipStreamSynchronize(c->stream));
  return check_flags(c);
  NB_CATCH("nb_sync")
}

static int force_pass(nb_ctx* c) {
  if (!c->uploaded) return nb_fail(NB_E_ARG, "no particle system uploaded");
  if (c->cfg.solver == NB_SOLVER_BH) {
    NB_HIP(nb_bh_build(c));
    NB_HIP(nb_bh_traverse(c, false));
    c->tree_valid = true;
  } else {
    NB_HIP(nb_launch_direct(c));
  }
  NB_HIP(hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int nb_forces(nb_ctx* c, float* force4) {
  NB_TRY
  NB_RANGE("nb_forces");
  NB_API_SHARED;
  if (!c || !force4) return nb_fail(NB_E_ARG, "nb_forces: null argument");
  int rc = force_pass(c);
  if (rc) return rc;
  const size_t n = c->n;
  std::vector<double> f(3 * n);
  NB_HIP(hipMemcpy(f.data(), c->force, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  to_caller_order(c, f, 3);
  for (size_t i = 0; i < n; ++i) {
    force4[4 * i] = (float)f[i];
    force4[4 * i + 1] = (float)f[n + i];
    force4[4 * i + 2] = (float)f[2 * n + i];
    force4[4 * i + 3] = 0.f;
  }
  return NB_OK;
  NB_CATCH("nb_forces")
}

int nb_forces_f64(nb_ctx* c, double* force3) {
  NB_TRY
  NB_RANGE("nb_forces_f64");
  NB_API_SHARED;
  if (!c || !force3) return nb_fail(NB_E_ARG, "nb_forces_f64: null argument");
  int rc = force_pass(c);
  if (rc) return rc;
  const size_t n = c->n;
  std::vector<double> f(3 * n);
  NB_HIP(hipMemcpy(f.data(), c->force, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  to_caller_order(c, f, 3);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) force3[3 * i + k] = f[k * n + i];
  return NB_OK;
  NB_CATCH("nb_forces_f64")
}

static int download_soa(nb_ctx* c, std::vector<double>& p, std::vector<double>& v,
                        std::vector<double>& a) {
  const size_t n = c->n;
  p.resize(3 * n);
  v.resize(3 * n);
  a.resize(3 * n);
  NB_HIP(hipStreamSynchronize(c->stream));
  NB_HIP(hipMemcpy(p.data(), c->pos, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  NB_HIP(hipMemcpy(v.data(), c->vel, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  NB_HIP(hipMemcpy(a.data(), c->acc, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  to_caller_order(c, p, 3);
  to_caller_order(c, v, 3);
  to_caller_order(c, a, 3);
  return check_flags(c);
}

int nb_download(nb_ctx* c, float* pos4, float* vel4, float* acc4) {
  NB_TRY
  NB_RANGE("nb_download");
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_download: null context");
  std::vector<double> p, v, a;
  const int rc = download_soa(c, p, v, a);
  const size_t n = c->n;
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 4; ++k) {
      if (pos4) pos4[4 * i + k] = k < 3 ? (float)p[k * n + i] : 0.f;
      if (vel4) vel4[4 * i + k] = k < 3 ? (float)v[k * n + i] : 0.f;
      if (acc4) acc4[4 * i + k] = k < 3 ? (float)a[k * n + i] : 0.f;
    }
  return rc;
  NB_CATCH("nb_download")
}

int nb_download_f64(nb_ctx* c, double* pos3, double* vel3, double* acc3) {
  NB_TRY
  NB_RANGE("nb_download_f64");
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_download_f64: null context");
  std::vector<double> p, v, a;
  const int rc = download_soa(c, p, v, a);
  const size_t n = c->n;
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      if (pos3) pos3[3 * i + k] = p[k * n + i];
      if (vel3) vel3[3 * i + k] = v[k * n + i];
      if (acc3) acc3[3 * i + k] = a[k * n + i];
    }
  return rc;
  NB_CATCH("nb_download_f64")
}

int nb_energy(nb_ctx* c, double* ekin, double* epot, double p[3]) {
  NB_TRY
  NB_RANGE("nb_energy");
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_energy: null context");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_energy: no particle system uploaded");
  const int kin_n = 4 * c->red_blocks;
  const int pot_n = (c->n + 255) / 256;
  NB_HIP(nb_launch_energy(c, c->red, c->red + kin_n));
  std::vector<double> h(kin_n + pot_n);
  NB_HIP(hipMemcpyAsync(h.data(), c->red, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  double ek = 0, px = 0, py = 0, pz = 0, ep = 0;
  for (int b = 0; b < c->red_blocks; ++b) {
    ek += h[4 * b];
    px += h[4 * b + 1];
    py += h[4 * b + 2];
    pz += h[4 * b + 3];
  }
  for (int b = 0; b < pot_n; ++b) ep += h[kin_n + b];
  if (ekin) *ekin = ek;
  if (epot) *epot = -0.5 * c->G * ep;
  if (p) { p[0] = px; p[1] = py; p[2] = pz; }
  return check_flags(c);
  NB_CATCH("nb_energy")
}

// Snapshot layout (little-endian): "NBSNAP01", int64 n, float eps2, float dt,
// int32 integrator, int32 solver, then the device's structure-of-arrays buffers
// as stored: positions, velocities, accelerations (double[3][n] each) and
// masses (float[n]).
static const char kSnapMagic[8] = {'N', 'B', 'S', 'N', 'A', 'P', '0', '1'};

int nb_save_snapshot(nb_ctx* c, const char* path) {
  NB_TRY
  NB_API_SHARED;
  if (!c || !path) return nb_fail(NB_E_ARG, "nb_save_snapshot: null argument");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_save_snapshot: nothing uploaded");
  std::vector<double> p, v, a;
  int rc = download_soa(c, p, v, a);
  if (rc) return rc;
  std::vector<float> m(c->n);
  NB_HIP(hipMemcpy(m.data(), c->mass, sizeof(float) * c->n, hipMemcpyDeviceToHost));
  to_caller_order(c, m, 1);
  FILE* f = std::fopen(path, "wb");
  if (!f) return nb_fail(NB_E_ARG, std::string("nb_save_snapshot: cannot open ") + path);
  const int64_t n = c->n;
  const int32_t meta[2] = {c->cfg.integrator, c->cfg.solver};
  bool ok = std::fwrite(kSnapMagic, 1, 8, f) == 8 && std::fwrite(&n, 8, 1, f) == 1 &&
            std::fwrite(&c->cfg.eps2, 4, 1, f) == 1 && std::fwrite(&c->cfg.dt, 4, 1, f) == 1 &&
            std::fwrite(meta, 4, 2, f) == 2 &&
            std::fwrite(p.data(), 8, p.size(), f) == p.size() &&
            std::fwrite(v.data(), 8, v.size(), f) == v.size() &&
            std::fwrite(a.data(), 8, a.size(), f) == a.size() &&
            std::fwrite(m.data(), 4, m.size(), f) == m.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? NB_OK : nb_fail(NB_E_ARG, std::string("nb_save_snapshot: write failed: ") + path);
  NB_CATCH("nb_save_snapshot")
}

int nb_load_snapshot(nb_ctx* c, const char* path) {
  NB_TRY
  NB_API_SHARED;
  if (!c || !path) return nb_fail(NB_E_ARG, "nb_load_snapshot: null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return nb_fail(NB_E_ARG, std::string("nb_load_snapshot: cannot open ") + path);
  char magic[8];
  int64_t n = 0;
  float eps2 = 0, dt = 0;
  int32_t meta[2];
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kSnapMagic, 8) == 0 &&
            std::fread(&n, 8, 1, f) == 1 && std::fread(&eps2, 4, 1, f) == 1 &&
            std::fread(&dt, 4, 1, f) == 1 && std::fread(meta, 4, 2, f) == 2;
  if (!ok || n != c->n) {
    std::fclose(f);
    return nb_fail(NB_E_ARG, "nb_load_snapshot: not a snapshot of this context's size");
  }
  std::vector<double> p(3 * n), v(3 * n), a(3 * n);
  std::vector<float> m(n);
  ok = std::fread(p.data(), 8, p.size(), f) == p.size() &&
       std::fread(v.data(), 8, v.size(), f) == v.size() &&
       std::fread(a.data(), 8, a.size(), f) == a.size() &&
       std::fread(m.data(), 4, m.size(), f) == m.size();
  std::fclose(f);
  if (!ok) return nb_fail(NB_E_ARG, "nb_load_snapshot: truncated file");
  return upload_common(c, p, v, a, m.data());
  NB_CATCH("nb_load_snapshot")
}

static int ensure_tmp(nb_ctx* c) {
  if (!c->tmp && hipMalloc(&c->tmp, sizeof(double) * 3 * (size_t)c->n) != hipSuccess) {
    c->tmp = nullptr;
    return nb_fail(NB_E_OOM, "device allocation failed (staging)");
  }
  return NB_OK;
}

int nb_reorder(nb_ctx* c) {
  NB_TRY
  NB_RANGE("nb_reorder");
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_reorder: null context");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_reorder: nothing uploaded");
  int rc = nb_sort_alloc(c);
  if (rc != NB_OK) return rc;
  rc = ensure_tmp(c);
  if (rc != NB_OK) return rc;
  NB_HIP(nb_morton_sort(c));
  NB_HIP(nb_launch_permute_state(c));
  NB_HIP(nb_launch_refresh_pm(c));
  const size_t n = c->n;
  std::vector<uint32_t> perm(n);
  NB_HIP(hipMemcpyAsync(perm.data(), c->bh.vals[0], sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                        c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  std::vector<int32_t> order(n);
  for (size_t k = 0; k < n; ++k)
    order[k] = c->order.empty() ? (int32_t)perm[k] : c->order[perm[k]];
  c->order.swap(order);
  c->tree_valid = false;
  return check_flags(c);
  NB_CATCH("nb_reorder")
}

int nb_render_view(nb_ctx* c, float* xyzs) {
  NB_TRY
  NB_API_SHARED;
  if (!c || !xyzs) return nb_fail(NB_E_ARG, "nb_render_view: null argument");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_render_view: nothing uploaded");
  int rc = ensure_tmp(c);
  if (rc != NB_OK) return rc;
  float4* view = reinterpret_cast<float4*>(c->tmp);  // 16 B of the 24 B per body staged
  NB_HIP(nb_launch_render_view(c, view));
  const size_t n = c->n;
  std::vector<float> h(4 * n);
  NB_HIP(hipMemcpyAsync(h.data(), view, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  for (size_t s = 0; s < n; ++s) {
    const size_t i = c->order.empty() ? s : (size_t)c->order[s];
    for (int k = 0; k < 4; ++k) xyzs[4 * i + k] = h[4 * s + k];
  }
  return check_flags(c);
  NB_CATCH("nb_render_view")
}

int nb_get_order(nb_ctx* c, int32_t* order) {
  NB_TRY
  if (!c || !order) return nb_fail(NB_E_ARG, "nb_get_order: null argument");
  for (int64_t k = 0; k < c->n; ++k) order[k] = c->order.empty() ? (int32_t)k : c->order[k];
  return NB_OK;
  NB_CATCH("nb_get_order")
}

int nb_set_timing(nb_ctx* c, int32_t enable) {
  NB_TRY
  NB_API_SHARED;
  if (!c) return nb_fail(NB_E_ARG, "nb_set_timing: null context");
  NB_HIP(hipStreamSynchronize(c->stream));
  c->timing = enable != 0;
  c->tim = nb_timing{};
  return NB_OK;
  NB_CATCH("nb_set_timing")
}

int nb_get_timing(nb_ctx* c, nb_timing* out) {
  NB_TRY
  if (!c || !out) return nb_fail(NB_E_ARG, "nb_get_timing: null argument");
  *out = c->tim;
  return NB_OK;
  NB_CATCH("nb_get_timing")
}

int nb_count_interactions(nb_ctx* c, int64_t* interactions) {
  NB_TRY
  NB_API_SHARED;
  if (!c || !interactions) return nb_fail(NB_E_ARG, "nb_count_interactions: null argument");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_count_interactions: nothing uploaded");
  if (c->cfg.solver != NB_SOLVER_BH) {
    *interactions = (int64_t)c->n * (c->n - 1);
    return NB_OK;
  }
  NB_HIP(nb_bh_build(c));
  NB_HIP(nb_bh_traverse(c, true));
  unsigned long long h = 0;
  NB_HIP(hipMemcpyAsync(&h, c->bh.inter, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  *interactions = (int64_t)h;
  c->tim.interactions = (int64_t)h;
  return check_flags(c);
  NB_CATCH("nb_count_interactions")
}

int nb_count_interactions_per_body(nb_ctx* c, int64_t* interactions) {
  NB_TRY
  NB_API_SHARED;
  if (!c || !interactions) return nb_fail(NB_E_ARG, "nb_count_interactions_per_body: null argument");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_count_interactions_per_body: nothing uploaded");
  if (c->cfg.solver != NB_SOLVER_BH) {
    *interactions = (int64_t)c->n * (c->n - 1);
    return NB_OK;
  }
  NB_HIP(nb_bh_build(c));
  NB_HIP(nb_bh_count_per_body(c));
  unsigned long long h = 0;
  NB_HIP(hipMemcpyAsync(&h, c->bh.inter, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  NB_HIP(hipStreamSynchronize(c->stream));
  *interactions = (int64_t)h;
  c->tree_valid = true;
  return check_flags(c);
  NB_CATCH("nb_count_interactions_per_body")
}

int nb_direct_f64(nb_ctx* c, const int64_t* targets, int64_t k, double* acc3) {
  NB_TRY
  NB_RANGE("nb_direct_f64");
  NB_API_SHARED;
  if (!c || (k > 0 && (!targets || !acc3))) return nb_fail(NB_E_ARG, "nb_direct_f64: null argument");
  if (!c->uploaded) return nb_fail(NB_E_ARG, "nb_direct_f64: nothing uploaded");
  if (k < 0 || k > c->n) return nb_fail(NB_E_ARG, "nb_direct_f64: k must be in [0, n]");
  if (k == 0) return NB_OK;
  // caller indices -> device slots (the inverse of the reorder map)
  std::vector<int32_t> slot_of;
  if (!c->order.empty()) {
    slot_of.resize(c->n);
    for (int s = 0; s < c->n; ++s) slot_of[c->order[s]] = s;
  }
  std::vector<int32_t> slots(k);
  for (int64_t t = 0; t < k; ++t) {
    if (targets[t] < 0 || targets[t] >= c->n) return nb_fail(NB_E_ARG, "nb_direct_f64: target out of range");
    slots[t] = slot_of.empty() ? (int32_t)targets[t] : slot_of[targets[t]];
  }
  void* d = nullptr;
  const size_t bytes = (size_t)k * (sizeof(int32_t) + 3 * sizeof(double)) + 256;
  if (hipMalloc(&d, bytes) != hipSuccess) return nb_fail(NB_E_OOM, "nb_direct_f64: device allocation failed");
  int* d_slots = static_cast<int*>(d);
  double* d_acc = reinterpret_cast<double*>(static_cast<char*>(d) + ((size_t)k * sizeof(int32_t) + 255) / 256 * 256);
  hipError_t e = hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = nb_launch_ref_f64(c, d_slots, (int)k, d_acc);
  if (e == hipSuccess)
    e = hipMemcpyAsync(acc3, d_acc, sizeof(double) * 3 * k, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return nb_fail(NB_E_HIP, std::string("nb_direct_f64: ") + hipGetErrorString(e));
  return NB_OK;
  NB_CATCH("nb_direct_f64")
}

int nb_device_info(char* buf, int32_t cap) {
  NB_TRY
  if (!buf || cap < 1) return nb_fail(NB_E_ARG, "nb_device_info: bad buffer");
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    std::snprintf(buf, cap, "no HIP device");
    return nb_fail(NB_E_NOGPU, "nb_device_info: no HIP device");
  }
  std::snprintf(buf, cap, "%s (%s), %d CUs, %.0f MHz, %.1f GiB", prop.name, prop.gcnArchName,
                prop.multiProcessorCount, prop.clockRate / 1000.0,
                prop.totalGlobalMem / 1073741824.0);
  return NB_OK;
  NB_CATCH("nb_device_info")
}

}  // extern "C"
