// nb_batch.hip — batched ensembles: many independent clusters of uneven size
// in one persistent force launch (see the nb_batch_* block of nbody.h).
//
// Layout in HBM: one body pool of `cap` bodies. A cluster occupies a contiguous
// segment [off, off + n_pad) (n_pad = n rounded up to 256, zero-mass padding),
// so its FP32 view pm[off..], FP64 state pos/vel/acc[k * cap + off..] and forces
// are plain slices. Per-slot parameters (offset, n, n_pad, eps2, dt, G) live in
// device arrays indexed by slot.
//
// Work decomposition:
//   force      work item = (slot, i-block of 64*B bodies); items are sorted by
//              cluster size, largest first, and as many workgroups as can be
//              resident at once pull them from an atomic ticket (persistent
//              kernel; every workgroup exits once the ticket passes the item
//              count). The K waves of a
//              workgroup split the cluster's sources; partial sums merge in LDS
//              in wave order, so a cluster's forces do not depend on scheduling
//              or on the other clusters.
//   integrate  one 256-thread block per 256-body block of the pool, via a
//              (slot, first body) table.
// Tables and counts are device-resident, so the captured step graph stays valid
// when clusters are added or removed.
#include <algorithm>
#include <cmath>
#include <map>
#include <new>
#include <vector>

#include "nb_force_core.h"
#include "nb_internal.h"

struct nb_batch {
  nb_batch_config cfg{};
  hipStream_t stream = nullptr;
  int64_t cap = 0;
  int bpl = 1, jsplit = 4;
  int cus = 256;  // compute units: the persistent grid fills every CU to its occupancy
  // device pool
  float4* pm = nullptr;
  double *pos = nullptr, *vel = nullptr, *acc = nullptr, *force = nullptr;
  float* mass = nullptr;
  // device slot table
  int* c_off = nullptr;
  int* c_n = nullptr;
  int* c_npad = nullptr;
  float* c_eps2 = nullptr;
  double* c_dt = nullptr;
  double* c_G = nullptr;
  int* c_quota = nullptr;  // steps the cluster takes per nb_batch_step call (INT32_MAX: all)
  // device work tables
  int2* items = nullptr;   // (slot, i-block)
  int2* blocks = nullptr;  // (slot, first body) per 256-body block
  int* counts = nullptr;   // [0] items, [1] blocks, [2] ticket, [3] flags, [4] step of the call
  double* red = nullptr;   // energy partials
  // host mirror
  std::vector<int> h_off, h_n, h_npad, h_quota;
  std::vector<char> used;
  std::map<int64_t, int64_t> free_list;  // offset -> size
  int64_t used_bodies = 0;
  int64_t interactions = 0;
  bool dirty = true;  // work tables need a rebuild
  // graph, timing
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool timing = false;
  hipEvent_t ev[3] = {};
  nb_timing tim{};
};

namespace {

constexpr int kUnlimited = NB_BATCH_UNLIMITED;

// use_quota: skip the clusters whose quota is used up in this call (steps);
// force-only passes (nb_batch_forces, nb_batch_tune) evaluate every cluster.
template <int B, int K>
__global__ __launch_bounds__(64 * K) void k_batch_force(
    const float4* __restrict__ pm, const int* __restrict__ c_off, const int* __restrict__ c_n,
    const int* __restrict__ c_npad, const float* __restrict__ c_eps2,
    const double* __restrict__ c_G, const int* __restrict__ c_quota, int use_quota,
    const int2* __restrict__ items, int* __restrict__ counts, int64_t cap,
    double* __restrict__ force) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nitems = counts[0];
  const int call_step = use_quota ? counts[4] : -1;
  __shared__ int s_item;
  __shared__ float part[K][6][B][64];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&counts[2], 1);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();  // every lane has read s_item before lane 0 overwrites it
    if (item >= nitems) break;
    const int2 it = items[item];
    const int slot = __builtin_amdgcn_readfirstlane(it.x);
    if (__builtin_amdgcn_readfirstlane(c_quota[slot]) <= call_step) continue;  // block-uniform
    const int off = __builtin_amdgcn_readfirstlane(c_off[slot]);
    const int npad = __builtin_amdgcn_readfirstlane(c_npad[slot]);
    const int n = c_n[slot];
    const float eps2 = c_eps2[slot];
    const int ibase = off + it.y * (64 * B) + lane;
    nbcore::Targets<B> t;
    nbcore::load_targets<B>(pm, ibase, t);
    const int jlen = npad / K;  // npad is a multiple of 256: a multiple of 16 for K <= 16
    const int j0 = off + wave * jlen;
    if (eps2 >= 1e-20f)
      nbcore::accumulate<B, false>(pm, j0, j0 + jlen, eps2, t);
    else
      nbcore::accumulate<B, true>(pm, j0, j0 + jlen, eps2, t);
    if constexpr (K > 1) nbcore::merge_waves<B, K>(part, wave, lane, t);
    if (wave == 0) {
      const double G = c_G[slot];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int i = ibase + 64 * b;
        if (i - off < n) {
          force[i] = G * ((double)t.sx[b] + (double)t.cx[b]);
          force[cap + i] = G * ((double)t.sy[b] + (double)t.cy[b]);
          force[2 * cap + i] = G * ((double)t.sz[b] + (double)t.cz[b]);
        }
      }
    }
    if constexpr (K > 1) __syncthreads();  // LDS partials are reused by the next item
  }
}

__global__ void k_batch_reset(int* __restrict__ counts) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counts[2] = 0;
}

// End of one step: the next step of the same nb_batch_step call.
__global__ void k_batch_advance(int* __restrict__ counts) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && counts[4] < 0x7fffffff) counts[4] += 1;
}

template <bool kEuler>
__global__ __launch_bounds__(256) void k_batch_integrate(
    const int2* __restrict__ blocks, const int* __restrict__ counts, const int* __restrict__ c_off,
    const int* __restrict__ c_n, const double* __restrict__ c_dt, const int* __restrict__ c_quota,
    int64_t cap, double* __restrict__ pos, double* __restrict__ vel, double* __restrict__ acc,
    const double* __restrict__ force, const float* __restrict__ mass, float4* __restrict__ pm) {
  // grid-stride over the pool blocks in use: the launch is sized for the
  // device, not for the pool's capacity, so an almost empty pool costs little
  const int nblocks = counts[1];
  const int step = counts[4];
  for (int blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
    const int2 bl = blocks[blk];
    if (c_quota[bl.x] <= step) continue;  // quota used up in this call
    const int i = bl.y + threadIdx.x;
    if (i - c_off[bl.x] >= c_n[bl.x]) continue;  // padding keeps zero mass
    const double dt = c_dt[bl.x];
    const double h = dt * 0.5;
    double x[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t k = a * cap + i;
      double v = vel[k];
      double p = pos[k];
      const double f = force[k];
      if (kEuler) {
        v = v + f * dt;
        p = p + v * dt;
      } else {
        v = v + acc[k] * h;
        p = p + v * dt;
        v = v + f * h;
      }
      acc[k] = f;
      vel[k] = v;
      pos[k] = p;
      x[a] = p;
    }
    pm[i] = make_float4((float)x[0], (float)x[1], (float)x[2], mass[i]);
  }
}

__global__ __launch_bounds__(256) void k_batch_refresh(int off, int n, int npad, int64_t cap,
                                                       const double* __restrict__ pos,
                                                       const float* __restrict__ mass,
                                                       float4* __restrict__ pm) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= npad) return;
  const int i = off + k;
  pm[i] = k < n ? make_float4((float)pos[i], (float)pos[cap + i], (float)pos[2 * cap + i], mass[i])
                : make_float4(0.f, 0.f, 0.f, 0.f);
}

// FP64 energy of one cluster: block partials of (ekin, sum_i m_i phi_i).
__global__ __launch_bounds__(256) void k_batch_energy(int off, int n, int64_t cap, double eps2,
                                                      const double* __restrict__ pos,
                                                      const double* __restrict__ vel,
                                                      const float* __restrict__ mass,
                                                      double* __restrict__ part) {
  __shared__ double red[2][256];
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int kk = k < n ? k : n - 1;
  const int i = off + kk;
  const double xi = pos[i], yi = pos[cap + i], zi = pos[2 * cap + i];
  double phi = 0.0;
  for (int q = 0; q < n; ++q) {
    const int j = off + q;
    const double dx = pos[j] - xi, dy = pos[cap + j] - yi, dz = pos[2 * cap + j] - zi;
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    phi += (q != kk) ? (double)mass[j] * rsqrt(r2) : 0.0;
  }
  const double m = mass[i];
  const double vx = vel[i], vy = vel[cap + i], vz = vel[2 * cap + i];
  red[0][threadIdx.x] = k < n ? 0.5 * m * (vx * vx + vy * vy + vz * vz) : 0.0;
  red[1][threadIdx.x] = k < n ? m * phi : 0.0;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) part[2 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

template <int B, int K>
hipError_t launch_force(nb_batch* b, int use_quota) {
  // Persistent: as many blocks as can be resident at once, which depends on the
  // shape's registers and LDS (one body per lane fits 8 waves per SIMD, two
  // bodies 4). A fixed 16 waves per CU left the one-body shapes at half
  // occupancy: 1x4 took 16.08 ms on the 512-cluster ensemble, 14.37 ms at full
  // occupancy.
  // (Queried per launch, for the current device: a host-side query, and steps
  // replay a captured graph, so it runs once per capture.)
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_force<B, K>, 64 * K, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = std::max(1, 16 / K);
  const int grid = b->cus * per_cu;
  hipLaunchKernelGGL((k_batch_force<B, K>), dim3(grid), dim3(64 * K), 0, b->stream, b->pm,
                     b->c_off, b->c_n, b->c_npad, b->c_eps2, b->c_G, b->c_quota, use_quota,
                     b->items, b->counts, b->cap, b->force);
  return hipGetLastError();
}

template <int B>
hipError_t launch_force_b(nb_batch* b, int use_quota) {
  switch (b->jsplit) {
    case 1: return launch_force<B, 1>(b, use_quota);
    case 2: return launch_force<B, 2>(b, use_quota);
    case 4: return launch_force<B, 4>(b, use_quota);
    case 8:
      if constexpr (B <= 2) return launch_force<B, 8>(b, use_quota);
      return hipErrorInvalidValue;
    case 16:
      if constexpr (B == 1) return launch_force<B, 16>(b, use_quota);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

hipError_t enqueue_force(nb_batch* b, int use_quota = 0) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->counts);
  switch (b->bpl) {
    case 1: return launch_force_b<1>(b, use_quota);
    case 2: return launch_force_b<2>(b, use_quota);
    case 4: return launch_force_b<4>(b, use_quota);
    default: return hipErrorInvalidValue;
  }
}

hipError_t enqueue_step(nb_batch* b) {
  hipError_t e = enqueue_force(b, 1);
  if (e != hipSuccess) return e;
  if (b->timing) (void)hipEventRecord(b->ev[1], b->stream);
  const unsigned nblk = (unsigned)std::min<int64_t>(b->cap / 256, (int64_t)b->cus * 64);
  if (b->cfg.integrator == NB_INTEG_EULER)
    hipLaunchKernelGGL((k_batch_integrate<true>), dim3(nblk), dim3(256), 0, b->stream, b->blocks,
                       b->counts, b->c_off, b->c_n, b->c_dt, b->c_quota, b->cap, b->pos, b->vel,
                       b->acc, b->force, b->mass, b->pm);
  else
    hipLaunchKernelGGL((k_batch_integrate<false>), dim3(nblk), dim3(256), 0, b->stream, b->blocks,
                       b->counts, b->c_off, b->c_n, b->c_dt, b->c_quota, b->cap, b->pos, b->vel,
                       b->acc, b->force, b->mass, b->pm);
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, b->stream, b->counts);
  return hipGetLastError();
}

// Rebuild the work tables after clusters were added or removed.
int rebuild_tables(nb_batch* b) {
  if (!b->dirty) return NB_OK;
  std::vector<int> slots;
  for (int s = 0; s < (int)b->used.size(); ++s)
    if (b->used[s]) slots.push_back(s);
  std::stable_sort(slots.begin(), slots.end(),
                   [&](int a, int c) { return b->h_npad[a] > b->h_npad[c]; });
  std::vector<int2> items, blocks;
  int64_t inter = 0;
  const int iblock = 64 * b->bpl;
  for (int s : slots) {
    for (int q = 0; q < b->h_npad[s] / iblock; ++q) items.push_back(make_int2(s, q));
    inter += (int64_t)b->h_n[s] * (b->h_n[s] - 1);
  }
  std::vector<int> by_off(slots);
  std::sort(by_off.begin(), by_off.end(), [&](int a, int c) { return b->h_off[a] < b->h_off[c]; });
  for (int s : by_off)
    for (int q = 0; q < b->h_npad[s]; q += 256) blocks.push_back(make_int2(s, b->h_off[s] + q));
  const int counts[2] = {(int)items.size(), (int)blocks.size()};
  if (!items.empty())
    NB_HIP(hipMemcpyAsync(b->items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice,
                          b->stream));
  if (!blocks.empty())
    NB_HIP(hipMemcpyAsync(b->blocks, blocks.data(), blocks.size() * sizeof(int2),
                          hipMemcpyHostToDevice, b->stream));
  NB_HIP(hipMemcpyAsync(b->counts, counts, sizeof(counts), hipMemcpyHostToDevice, b->stream));
  NB_HIP(hipStreamSynchronize(b->stream));  // the host vectors die here
  b->interactions = inter;
  b->dirty = false;
  return NB_OK;
}

void free_batch(nb_batch* b) {
  if (!b) return;
  nb_drop_graph(b->stream, b->graph_exec, b->graph);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  void* ps[] = {b->pm,     b->pos,    b->vel,   b->acc,   b->force, b->mass,  b->c_off,
                b->c_n,    b->c_npad, b->c_eps2, b->c_dt, b->c_G,   b->items, b->blocks,
                b->counts, b->red,    b->c_quota};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int check_slot(nb_batch* b, int32_t id) {
  if (!b) return nb_fail(NB_E_ARG, "nb_batch: null batch");
  if (id < 0 || id >= (int)b->used.size() || !b->used[id])
    return nb_fail(NB_E_ARG, "nb_batch: no such cluster");
  return NB_OK;
}

}  // namespace

extern "C" {

int nb_batch_create(nb_batch** out, const nb_batch_config* cfg) {
  NB_TRY
  NB_API_SHARED;
  if (!out || !cfg) return nb_fail(NB_E_ARG, "nb_batch_create: null argument");
  *out = nullptr;
  if (cfg->max_bodies < 256 || cfg->max_bodies >= (int64_t(1) << 30))
    return nb_fail(NB_E_ARG, "nb_batch_create: max_bodies must be in [256, 2^30)");
  if (cfg->max_clusters < 1 || cfg->max_clusters > (1 << 20))
    return nb_fail(NB_E_ARG, "nb_batch_create: max_clusters must be in [1, 2^20]");
  if (cfg->integrator != NB_INTEG_KDK_LAGGED && cfg->integrator != NB_INTEG_EULER)
    return nb_fail(NB_E_ARG, "nb_batch_create: unknown integrator");
  if (cfg->bodies_per_lane != 0 && cfg->bodies_per_lane != 1 && cfg->bodies_per_lane != 2 &&
      cfg->bodies_per_lane != 4)
    return nb_fail(NB_E_ARG, "nb_batch_create: bodies_per_lane must be 0, 1, 2 or 4");
  if (cfg->j_split != 0 && cfg->j_split != 1 && cfg->j_split != 2 && cfg->j_split != 4 &&
      cfg->j_split != 8 && cfg->j_split != 16)
    return nb_fail(NB_E_ARG, "nb_batch_create: j_split must be 0, 1, 2, 4, 8 or 16");
  if (cfg->bodies_per_lane > 0 && cfg->j_split > 0 &&
      !nb_valid_shape(cfg->bodies_per_lane, cfg->j_split))
    return nb_fail(NB_E_ARG, "nb_batch_create: j_split 8 needs bodies_per_lane <= 2, 16 needs 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return nb_fail(NB_E_NOGPU, "nb_batch_create: no HIP device visible");
  nb_batch* b = new (std::nothrow) nb_batch();
  if (!b) return nb_fail(NB_E_OOM, "nb_batch_create: host allocation failed");
  b->cfg = *cfg;
  b->cap = (cfg->max_bodies + 255) / 256 * 256;
  // Built-in shape, measured on MI355X on the 512-cluster ensemble
  // (N_c log-uniform in [256, 32768]) with occupancy-sized launches: one body
  // per lane and four waves per work item, 14.4 ms per force pass against
  // 14.7 ms for 1x8 and 15.4 ms for 2x4. nb_batch_tune times every shape on the
  // batch's own clusters.
  b->bpl = cfg->bodies_per_lane > 0 ? cfg->bodies_per_lane : 1;
  b->jsplit = cfg->j_split > 0 ? cfg->j_split : 4;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->cus = cus;
  auto fail = [&](int code, const char* m) {
    free_batch(b);
    return nb_fail(code, m);
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(NB_E_HIP, "nb_batch_create: hipStreamCreate failed");
  const size_t cap = b->cap, mc = cfg->max_clusters;
  auto A = [](void** p, size_t bytes) { return hipMalloc(p, bytes < 256 ? 256 : bytes); };
  if (A((void**)&b->pm, cap * sizeof(float4)) || A((void**)&b->pos, 3 * cap * sizeof(double)) ||
      A((void**)&b->vel, 3 * cap * sizeof(double)) || A((void**)&b->acc, 3 * cap * sizeof(double)) ||
      A((void**)&b->force, 3 * cap * sizeof(double)) || A((void**)&b->mass, cap * sizeof(float)) ||
      A((void**)&b->c_off, mc * sizeof(int)) || A((void**)&b->c_n, mc * sizeof(int)) ||
      A((void**)&b->c_npad, mc * sizeof(int)) || A((void**)&b->c_eps2, mc * sizeof(float)) ||
      A((void**)&b->c_dt, mc * sizeof(double)) || A((void**)&b->c_G, mc * sizeof(double)) ||
      A((void**)&b->c_quota, mc * sizeof(int)) ||
      A((void**)&b->items, (cap / 64) * sizeof(int2)) ||
      A((void**)&b->blocks, (cap / 256) * sizeof(int2)) || A((void**)&b->counts, 16 * sizeof(int)) ||
      A((void**)&b->red, (2 * (cap / 256) + 16) * sizeof(double)))
    return fail(NB_E_OOM, "nb_batch_create: device allocation failed");
  if (hipMemset(b->pm, 0, cap * sizeof(float4)) != hipSuccess ||
      hipMemset(b->mass, 0, cap * sizeof(float)) != hipSuccess ||
      hipMemset(b->force, 0, 3 * cap * sizeof(double)) != hipSuccess ||
      hipMemset(b->counts, 0, 16 * sizeof(int)) != hipSuccess)
    return fail(NB_E_HIP, "nb_batch_create: hipMemset failed");
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(NB_E_HIP, "nb_batch_create: event failed");
  b->h_off.assign(mc, 0);
  b->h_n.assign(mc, 0);
  b->h_npad.assign(mc, 0);
  b->used.assign(mc, 0);
  b->h_quota.assign(mc, kUnlimited);
  b->free_list[0] = b->cap;
  *out = b;
  return NB_OK;
  NB_CATCH("nb_batch_create")
}

void nb_batch_destroy(nb_batch* b) {
  NB_API_SHARED;
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  free_batch(b);
}

int nb_batch_add(nb_batch* b, const double* pos3, const double* vel3, const double* acc3,
                 const float* mass, int64_t n, float eps2, float dt, float G, int32_t* id) {
  NB_TRY
  NB_API_SHARED;
  if (!b || !pos3 || !vel3 || !mass || !id) return nb_fail(NB_E_ARG, "nb_batch_add: null argument");
  if (n < 1) return nb_fail(NB_E_ARG, "nb_batch_add: n < 1");
  if (!(eps2 >= 0.f) || !std::isfinite(eps2) || !std::isfinite(dt))
    return nb_fail(NB_E_ARG, "nb_batch_add: bad eps2 or dt");
  const int64_t npad = (n + 255) / 256 * 256;
  int slot = -1;
  for (int s = 0; s < (int)b->used.size(); ++s)
    if (!b->used[s]) { slot = s; break; }
  if (slot < 0) return nb_fail(NB_E_BUDGET, "nb_batch_add: cluster table full");
  auto it = b->free_list.begin();
  for (; it != b->free_list.end(); ++it)
    if (it->second >= npad) break;
  if (it == b->free_list.end()) return nb_fail(NB_E_BUDGET, "nb_batch_add: body pool full");
  const int64_t off = it->first, len = it->second;
  b->free_list.erase(it);
  if (len > npad) b->free_list[off + npad] = len - npad;
  // stage the FP64 state as structure-of-arrays slices of the pool
  const size_t cap = b->cap;
  std::vector<double> p(3 * n), v(3 * n), a(3 * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      p[k * n + i] = pos3[3 * i + k];
      v[k * n + i] = vel3[3 * i + k];
      if (acc3) a[k * n + i] = acc3[3 * i + k];
    }
  std::vector<float> m(npad, 0.f);
  std::copy(mass, mass + n, m.begin());
  const int io = (int)off, in = (int)n, inp = (int)npad;
  const double ddt = dt, dG = G == 0.f ? 1.0 : (double)G;
  auto upload = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    auto put = [&](void* dst, const void* src, size_t bytes) {
      if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream);
    };
    for (int k = 0; k < 3; ++k) {
      put(b->pos + k * cap + off, p.data() + k * n, n * sizeof(double));
      put(b->vel + k * cap + off, v.data() + k * n, n * sizeof(double));
      put(b->acc + k * cap + off, a.data() + k * n, n * sizeof(double));
    }
    put(b->mass + off, m.data(), npad * sizeof(float));
    put(b->c_off + slot, &io, sizeof(int));
    put(b->c_n + slot, &in, sizeof(int));
    put(b->c_npad + slot, &inp, sizeof(int));
    put(b->c_eps2 + slot, &eps2, sizeof(float));
    put(b->c_dt + slot, &ddt, sizeof(double));
    put(b->c_G + slot, &dG, sizeof(double));
    put(b->c_quota + slot, &kUnlimited, sizeof(int));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_refresh, dim3((unsigned)(npad / 256)), dim3(256), 0, b->stream, io,
                       in, inp, b->cap, b->pos, b->mass, b->pm);
    e = hipGetLastError();
    const hipError_t es = hipStreamSynchronize(b->stream);  // host staging buffers die here
    return e != hipSuccess ? e : es;
  };
  const hipError_t ue = upload();
  if (ue != hipSuccess) {
    // give the segment back (coalescing happens on the next remove)
    if (len > npad) b->free_list.erase(off + npad);
    b->free_list[off] = len;
    return nb_fail(NB_E_HIP, std::string("nb_batch_add: upload failed: ") + hipGetErrorString(ue));
  }
  b->used[slot] = 1;
  b->h_off[slot] = io;
  b->h_n[slot] = in;
  b->h_npad[slot] = inp;
  b->h_quota[slot] = kUnlimited;
  b->used_bodies += npad;
  b->dirty = true;
  *id = slot;
  return NB_OK;
  NB_CATCH("nb_batch_add")
}

int nb_batch_remove(nb_batch* b, int32_t id) {
  NB_TRY
  NB_API_SHARED;
  int rc = check_slot(b, id);
  if (rc) return rc;
  NB_HIP(hipStreamSynchronize(b->stream));
  int64_t off = b->h_off[id], len = b->h_npad[id];
  // zero the FP32 view so a stale body can never be read as a source
  NB_HIP(hipMemsetAsync(b->pm + off, 0, len * sizeof(float4), b->stream));
  NB_HIP(hipStreamSynchronize(b->stream));
  b->used[id] = 0;
  b->used_bodies -= len;
  // return the segment to the free list, coalescing neighbours
  auto nx = b->free_list.lower_bound(off);
  if (nx != b->free_list.end() && nx->first == off + len) {
    len += nx->second;
    b->free_list.erase(nx);
  }
  auto pv = b->free_list.lower_bound(off);
  if (pv != b->free_list.begin()) {
    --pv;
    if (pv->first + pv->second == off) {
      off = pv->first;
      len += pv->second;
      b->free_list.erase(pv);
    }
  }
  b->free_list[off] = len;
  b->dirty = true;
  return NB_OK;
  NB_CATCH("nb_batch_remove")
}

int nb_batch_set_quotas(nb_batch* b, const int32_t* ids, const int32_t* steps, int32_t count) {
  NB_TRY
  NB_API_SHARED;
  if (!b || count < 0 || (count > 0 && (!ids || !steps)))
    return nb_fail(NB_E_ARG, "nb_batch_set_quotas: bad argument");
  for (int32_t k = 0; k < count; ++k) {
    int rc = check_slot(b, ids[k]);
    if (rc) return rc;
    if (steps[k] < 0) return nb_fail(NB_E_ARG, "nb_batch_set_quotas: negative quota");
  }
  // the previous upload of the table may still read the host mirror
  NB_HIP(hipStreamSynchronize(b->stream));
  for (int32_t k = 0; k < count; ++k) b->h_quota[ids[k]] = steps[k];
  NB_HIP(hipMemcpyAsync(b->c_quota, b->h_quota.data(), b->h_quota.size() * sizeof(int),
                        hipMemcpyHostToDevice, b->stream));
  return NB_OK;
  NB_CATCH("nb_batch_set_quotas")
}

int nb_batch_step(nb_batch* b, int32_t nsteps) {
  NB_TRY
  NB_RANGE("nb_batch_step");
  if (!b) return nb_fail(NB_E_ARG, "nb_batch_step: null batch");
  if (nsteps < 0) return nb_fail(NB_E_ARG, "nb_batch_step: nsteps < 0");
  {
    NB_API_SHARED;  // the rebuild copies and synchronises; released before any capture
    int rc = rebuild_tables(b);
    if (rc) return rc;
  }
  if (nsteps == 0) return NB_OK;
  // the steps of this call count from 0 for the clusters' quotas
  NB_HIP(nb_launch_fill_i32(b->counts + 4, 0, 1, b->stream));
  if (b->timing) {
    NB_API_SHARED;
    for (int s = 0; s < nsteps; ++s) {
      (void)hipEventRecord(b->ev[0], b->stream);
      NB_HIP(enqueue_step(b));
      (void)hipEventRecord(b->ev[2], b->stream);
      NB_HIP(hipEventSynchronize(b->ev[2]));
      float f = 0, t = 0;
      NB_HIP(hipEventElapsedTime(&f, b->ev[0], b->ev[1]));
      NB_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[2]));
      b->tim.force_ms += f;
      b->tim.force_calls++;
      b->tim.step_ms += t;
      b->tim.step_calls++;
    }
    return NB_OK;
  }
  if (b->cfg.flags & NB_F_NO_GRAPH) {
    for (int s = 0; s < nsteps; ++s) NB_HIP(enqueue_step(b));
    return NB_OK;
  }
  if (!b->graph_exec) {
    NB_API_CAPTURE;
    NB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_step(b);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return nb_fail(NB_E_HIP, "nb_batch_step: graph capture failed");
    }
    const hipError_t ei = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(g);
      b->graph_exec = nullptr;
      return nb_fail(NB_E_HIP, std::string("nb_batch_step: graph instantiation failed: ") +
                                   hipGetErrorString(ei));
    }
    b->graph = g;
  }
  for (int s = 0; s < nsteps; ++s) NB_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  return NB_OK;
  NB_CATCH("nb_batch_step")
}

int nb_batch_sync(nb_batch* b) {
  NB_TRY
  NB_API_SHARED;
  if (!b) return nb_fail(NB_E_ARG, "nb_batch_sync: null batch");
  NB_HIP(hipStreamSynchronize(b->stream));
  return NB_OK;
  NB_CATCH("nb_batch_sync")
}

static int copy_slices(nb_batch* b, int32_t id, const double* src, double* dst) {
  const int64_t n = b->h_n[id], off = b->h_off[id];
  std::vector<double> buf(n);
  for (int k = 0; k < 3; ++k) {
    NB_HIP(hipMemcpy(buf.data(), src + k * b->cap + off, n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) dst[3 * i + k] = buf[i];
  }
  return NB_OK;
}

int nb_batch_forces(nb_batch* b, int32_t id, double* force3) {
  NB_TRY
  NB_RANGE("nb_batch_forces");
  NB_API_SHARED;
  int rc = check_slot(b, id);
  if (rc) return rc;
  if (!force3) return nb_fail(NB_E_ARG, "nb_batch_forces: null output");
  rc = rebuild_tables(b);
  if (rc) return rc;
  NB_HIP(enqueue_force(b));
  NB_HIP(hipStreamSynchronize(b->stream));
  return copy_slices(b, id, b->force, force3);
  NB_CATCH("nb_batch_forces")
}

int nb_batch_download(nb_batch* b, int32_t id, double* pos3, double* vel3, double* acc3) {
  NB_TRY
  NB_API_SHARED;
  int rc = check_slot(b, id);
  if (rc) return rc;
  NB_HIP(hipStreamSynchronize(b->stream));
  if (pos3 && (rc = copy_slices(b, id, b->pos, pos3))) return rc;
  if (vel3 && (rc = copy_slices(b, id, b->vel, vel3))) return rc;
  if (acc3 && (rc = copy_slices(b, id, b->acc, acc3))) return rc;
  return NB_OK;
  NB_CATCH("nb_batch_download")
}

int nb_batch_energy(nb_batch* b, int32_t id, double* ekin, double* epot) {
  NB_TRY
  NB_API_SHARED;
  int rc = check_slot(b, id);
  if (rc) return rc;
  const int n = b->h_n[id], off = b->h_off[id];
  const int blocks = (n + 255) / 256;
  float eps2 = 0.f;
  double G = 1.0;
  NB_HIP(hipMemcpyAsync(&eps2, b->c_eps2 + id, sizeof(float), hipMemcpyDeviceToHost, b->stream));
  NB_HIP(hipMemcpyAsync(&G, b->c_G + id, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  NB_HIP(hipStreamSynchronize(b->stream));
  hipLaunchKernelGGL(k_batch_energy, dim3(blocks), dim3(256), 0, b->stream, off, n, b->cap,
                     (double)eps2, b->pos, b->vel, b->mass, b->red);
  NB_HIP(hipGetLastError());
  std::vector<double> part(2 * blocks);
  NB_HIP(hipMemcpyAsync(part.data(), b->red, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                        b->stream));
  NB_HIP(hipStreamSynchronize(b->stream));
  double ek = 0, ep = 0;
  for (int q = 0; q < blocks; ++q) {
    ek += part[2 * q];
    ep += part[2 * q + 1];
  }
  if (ekin) *ekin = ek;
  if (epot) *epot = -0.5 * G * ep;
  return NB_OK;
  NB_CATCH("nb_batch_energy")
}

int nb_batch_info(nb_batch* b, int32_t* clusters, int64_t* bodies, int64_t* interactions) {
  NB_TRY
  if (!b) return nb_fail(NB_E_ARG, "nb_batch_info: null batch");
  int rc = rebuild_tables(b);
  if (rc) return rc;
  if (clusters) *clusters = (int32_t)std::count(b->used.begin(), b->used.end(), 1);
  if (bodies) *bodies = b->used_bodies;
  if (interactions) *interactions = b->interactions;
  return NB_OK;
  NB_CATCH("nb_batch_info")
}

int nb_batch_set_timing(nb_batch* b, int32_t enable) {
  NB_TRY
  NB_API_SHARED;
  if (!b) return nb_fail(NB_E_ARG, "nb_batch_set_timing: null batch");
  NB_HIP(hipStreamSynchronize(b->stream));
  b->timing = enable != 0;
  b->tim = nb_timing{};
  return NB_OK;
  NB_CATCH("nb_batch_set_timing")
}

int nb_batch_get_timing(nb_batch* b, nb_timing* out) {
  NB_TRY
  if (!b || !out) return nb_fail(NB_E_ARG, "nb_batch_get_timing: null argument");
  *out = b->tim;
  out->interactions = b->interactions;
  return NB_OK;
  NB_CATCH("nb_batch_get_timing")
}

int nb_batch_tune(nb_batch* b, int32_t reps, nb_tuning* out) {
  NB_TRY
  NB_RANGE("nb_batch_tune");
  NB_API_SHARED;
  if (!b) return nb_fail(NB_E_ARG, "nb_batch_tune: null batch");
  if (reps < 1) reps = 3;
  NB_HIP(hipStreamSynchronize(b->stream));
  const int bpl0 = b->bpl, js0 = b->jsplit;
  int best_b = bpl0, best_k = js0;
  double best_ms = -1.0;
  // The force pass leaves the state untouched (the next step recomputes the
  // forces it integrates), so every shape can be timed on the live clusters.
  for (int bb : {1, 2, 4})
    for (int k : {1, 2, 4, 8, 16}) {
      if (!nb_valid_shape(bb, k)) continue;
      b->bpl = bb;
      b->jsplit = k;
      b->dirty = true;
      int rc = rebuild_tables(b);
      if (rc) return rc;
      NB_HIP(enqueue_force(b));  // warm-up
      NB_HIP(hipEventRecord(b->ev[0], b->stream));
      for (int r = 0; r < reps; ++r) NB_HIP(enqueue_force(b));
      NB_HIP(hipEventRecord(b->ev[2], b->stream));
      NB_HIP(hipEventSynchronize(b->ev[2]));
      float ms = 0.f;
      NB_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[2]));
      const double per = ms / reps;
      if (best_ms < 0 || per < best_ms) {
        best_ms = per;
        best_b = bb;
        best_k = k;
      }
    }
  b->bpl = best_b;
  b->jsplit = best_k;
  b->dirty = true;
  nb_drop_graph(b->stream, b->graph_exec, b->graph);
  int rc = rebuild_tables(b);
  if (rc) return rc;
  if (out) {
    *out = nb_tuning{};
    int64_t bodies = 0;
    for (size_t q = 0; q < b->used.size(); ++q)
      if (b->used[q]) bodies += b->h_n[q];
    out->n = bodies;
    out->bodies_per_lane = best_b;
    out->j_split = best_k;
    out->direct_ms = best_ms;
  }
  return NB_OK;
  NB_CATCH("nb_batch_tune")
}

}  // extern "C"
