// nb_internal.h — device-side state and kernel launchers shared by the
// translation units of libnbody (not installed; the public ABI is nbody.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <mutex>
#include <new>
#include <shared_mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "nbody.h"

// ---- error plumbing -------------------------------------------------------
void nb_set_error(const std::string& msg);
int nb_fail(int code, const std::string& msg);
#define NB_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return nb_fail(NB_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Stream capture and the runtime calls that may not overlap it. A graph
// capture (the first nb_step / nb_batch_step of a context) holds this lock
// exclusively; entry points that allocate, free, copy or synchronise hold it
// shared. Different contexts may therefore be used from different threads.
std::shared_mutex& nb_capture_mutex();
#define NB_API_SHARED std::shared_lock<std::shared_mutex> nb_api_guard_(nb_capture_mutex())
#define NB_API_CAPTURE std::unique_lock<std::shared_mutex> nb_capture_guard_(nb_capture_mutex())

// Every int-returning C entry point runs its body inside NB_TRY ... NB_CATCH so
// that no C++ exception (a failed host allocation, a thread that cannot start)
// crosses the C ABI: it becomes a status code and an nb_last_error() message.
#define NB_TRY try {
#define NB_CATCH(fn)                                                                  \
  }                                                                                   \
  catch (const std::bad_alloc&) {                                                     \
    return nb_fail(NB_E_OOM, fn ": host allocation failed");                          \
  }                                                                                   \
  catch (const std::exception& nb_ex_) {                                              \
    return nb_fail(NB_E_HIP, std::string(fn ": ") + nb_ex_.what());                   \
  }                                                                                   \
  catch (...) {                                                                       \
    return nb_fail(NB_E_HIP, fn ": unexpected exception");                            \
  }

// Host trace range over an entry point (roctx): it names the call on a
// `rocprofv3 --marker-trace` timeline and costs a few ns when no tool is attached.
struct nb_trace_range {
  explicit nb_trace_range(const char* name) { roctxRangePushA(name); }
  ~nb_trace_range() { roctxRangePop(); }
  nb_trace_range(const nb_trace_range&) = delete;
  nb_trace_range& operator=(const nb_trace_range&) = delete;
};
#define NB_RANGE(name) nb_trace_range nb_trace_range_(name)

// Padding granule of the FP32 body array: the direct kernel's i-blocks hold
// 64 lanes x up to 4 bodies, and its j-range is split over up to 4 waves in
// multiples of 4 bodies, so every count is a multiple of 256.
constexpr int kPadBodies = 256;

// Octree geometry: 30-bit Morton codes give at most 10 levels below the root.
constexpr int kMaxLevel = 10;

// ---- Barnes-Hut buffer layout --------------------------------------------------
// Every device array of the Barnes-Hut pipeline for n bodies lives in one block:
// a context allocates it once (nb_bh_alloc), and nb_bh_forces_noalloc carves the
// same layout out of a static device array. The Morton-sort buffers come first,
// so a direct-sum context that only reorders (nb_sort_alloc) allocates the
// prefix [0, off[kBhPms]).
constexpr int kBboxBlocks = 1024;                     // bbox partial blocks
constexpr int kSortThreads = 256;                     // 4 waves
constexpr int kSortItems = 8;                         // keys per lane per radix pass
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 keys per block
constexpr int kScanTile = 2048;                       // elements per block in the scans
constexpr int kSpare = 4;                             // zero-mass bodies after the sorted bodies
constexpr int kTopLevels = 4;                         // levels built from per-cell searches
constexpr int kTopCells = 1 + 8 + 64 + 512 + 4096;    // cells of levels 0..kTopLevels
constexpr int kTopEntries = kTopCells + kTopLevels + 1;  // plus an end sentinel per level
constexpr int kCostBuckets = 64;  // half-octaves of estimated record time
// Whether the Barnes-Hut build pairs sibling ranges (and the traversal orders
// its work): by size unless a flag forces it either way.
inline bool nb_bh_pairs(int64_t n, int flags) {
  if (flags & NB_F_BH_SINGLE_RANGES) return false;
  return (flags & NB_F_BH_PAIRS) || n >= NB_BH_PAIR_MIN_BODIES;
}

enum nb_bh_buf {
  kBhBbox, kBhGeom, kBhKey0, kBhKey1, kBhVal0, kBhVal1, kBhCodes, kBhHist, kBhIscan,
  kBhPms, kBhPrefix, kBhScanPart, kBhCom, kBhR2, kBhLink, kBhLevel, kBhChildCnt,
  kBhChildOff, kBhChildBounds, kBhCounters, kBhTopStart, kBhWaves, kBhInter, kBhFlags,
  kBhRecCost, kBhRecBucket, kBhRecOrder, kBhRecHist, kBhBufCount
};

// Node bound: a level holds at most min(8^l, n) nodes and there are 11 levels.
constexpr int64_t nb_tree_bound(int64_t n) {
  int64_t total = 0, w = 1;
  for (int l = 0; l <= kMaxLevel; ++l) {
    total += w < n ? w : n;
    if (w < n) w *= 8;
  }
  return total;
}

struct nb_bh_layout {
  int64_t off[kBhBufCount + 1] = {};  // byte offsets, 256-aligned; off[kBhBufCount] = total
  int64_t sort_blocks = 0, max_nodes = 0;
};

constexpr nb_bh_layout nb_bh_layout_for(int64_t n) {
  nb_bh_layout L{};
  const int64_t sort_blocks = (n + kSortTile - 1) / kSortTile;
  const int64_t hist = 256 * sort_blocks;
  const int64_t scan_parts = ((hist > n ? hist : n) + kScanTile - 1) / kScanTile + 2;
  const int64_t bound = nb_tree_bound(n);
  int64_t bytes[kBhBufCount] = {};
  bytes[kBhBbox] = 4 * 6 * kBboxBlocks;
  bytes[kBhGeom] = 4 * 32;
  bytes[kBhKey0] = bytes[kBhKey1] = bytes[kBhVal0] = bytes[kBhVal1] = bytes[kBhCodes] = 4 * n;
  bytes[kBhHist] = 4 * 2 * hist;
  bytes[kBhIscan] = 4 * scan_parts;
  bytes[kBhPms] = 16 * (n + kSpare);
  bytes[kBhPrefix] = 32 * (n + 1);
  bytes[kBhScanPart] = 32 * ((n + kScanTile - 1) / kScanTile + 2);
  bytes[kBhCom] = 16 * bound;
  bytes[kBhR2] = 4 * bound;
  bytes[kBhLink] = 16 * bound;
  bytes[kBhLevel] = 4 * (kMaxLevel + 2);
  bytes[kBhChildCnt] = bytes[kBhChildOff] = 4 * n;
  bytes[kBhChildBounds] = 4 * 8 * n;
  bytes[kBhCounters] = 4 * 16;
  bytes[kBhTopStart] = 4 * kTopEntries;
  bytes[kBhWaves] = 16 * n;
  bytes[kBhInter] = 8 * 2;
  bytes[kBhFlags] = 4 * 4;
  bytes[kBhRecCost] = bytes[kBhRecBucket] = bytes[kBhRecOrder] = 4 * n;
  bytes[kBhRecHist] = 4 * 2 * kCostBuckets;
  int64_t at = 0;
  for (int k = 0; k < kBhBufCount; ++k) {
    L.off[k] = at;
    at += (bytes[k] + 255) / 256 * 256;
  }
  L.off[kBhBufCount] = at;
  L.sort_blocks = sort_blocks;
  L.max_nodes = bound;
  return L;
}

// Largest system nb_bh_forces_noalloc accepts (its static scratch is sized for it).
constexpr int64_t kNoAllocMaxBodies = NB_NOALLOC_MAX_BODIES;

// Device arrays of the Barnes-Hut pipeline (sized once at nb_create).
struct nb_bh_state {
  void* block = nullptr;        // owning allocation the arrays below are carved from (or null)
  int64_t max_nodes = 0;
  // bbox: per-block partial min/max and the final (min xyz, max xyz, side2[11])
  float* bbox_part = nullptr;   // [2 * 3 * bbox_blocks]
  float* geom = nullptr;        // [6 + 11]: bmin[3], bmax[3], side^2 per level
  int bbox_blocks = 0;
  // Morton keys and body indices, double-buffered for the radix sort
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* codes = nullptr;    // unsorted codes (debug tap)
  int* digit_hist = nullptr;       // radix: [256 * sort_blocks] counts, then offsets
  int sort_blocks = 0;
  int sorted_buf = 0;           // which of keys[]/vals[] holds the final order
  float4* pms = nullptr;        // bodies in sorted order (x, y, z, m)
  double4* prefix = nullptr;    // exclusive FP64 prefix of (m, m x, m y, m z), n+1 entries
  double4* scan_part = nullptr; // scan block partials
  // octree, breadth-first
  float4* node_com = nullptr;   // (cx, cy, cz, m)
  float* node_r2 = nullptr;     // squared opening radius (s / theta + |COM - cell centre|)^2
  int4* node_link = nullptr;    // (first_child, next, start, count << 4 | level)
  int* level_off = nullptr;     // [kMaxLevel + 2] node index where each level begins
  int* child_cnt = nullptr;     // per node of the current level: non-empty children
  int* child_off = nullptr;     // exclusive scan of child_cnt
  int* child_bounds = nullptr;  // per node of the current level: its 7 inner child boundaries [8 n]
  int* iscan_part = nullptr;    // int scan block partials
  int* counters = nullptr;      // [0] wave ranges, [1] traversal ticket
  int4* waves = nullptr;        // records of one or two sibling wave ranges (start, count, start, count)
  int* top_start = nullptr;     // first sorted body of every cell of the top levels
  // traversal order: records by the device time their bodies took in the last pass
  float* rec_cost = nullptr;    // per body (original index): last pass's time of its record / its bodies
  int* rec_bucket = nullptr;    // per record: cost bucket of this pass
  int* rec_order = nullptr;     // tickets -> records, costliest buckets first
  int* rec_hist = nullptr;      // [kCostBuckets] records per bucket, then [kCostBuckets] cursors
  int persist_blocks = 0;       // traversal workgroups (persistent)
  unsigned long long* inter = nullptr;  // interaction counter (diagnostic traversal)
};

struct nb_ctx {
  nb_config cfg{};
  int64_t cap_n = 0;  // bodies the device buffers were sized for (nb_ctx_rebind limit)
  int n = 0;
  int n_pad = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  float theta2 = 0.25f;
  double G = 1.0;

  // particle state, FP64 structure-of-arrays: [x(0..n), y(0..n), z(0..n)]
  double* pos = nullptr;
  double* vel = nullptr;
  double* acc = nullptr;
  double* force = nullptr;   // last force pass, [3][n]
  float* mass = nullptr;     // [n_pad], zero in the padding
  float4* pm = nullptr;      // FP32 force view (x, y, z, m), [n_pad]
  int* flags = nullptr;      // device status word (bit 0: non-finite state)
  double* red = nullptr;     // reduction scratch for diagnostics
  int red_blocks = 0;

  // direct-sum launch shape
  int bpl = 2;     // i-bodies per lane
  int jsplit = 2;  // waves per i-block
  bool guard = false;
  int walk_stack = 0;  // BH traversal stack region cap (0: the full kStackCap); a test hook

  nb_bh_state bh;
  // Morton reordering (nb_reorder): order[slot] = the caller's index of the
  // body held in that slot; empty while slots are in the caller's order.
  std::vector<int32_t> order;
  double* tmp = nullptr;     // [3][n] permutation staging, allocated on first use
  bool uploaded = false;
  bool tree_valid = false;

  // captured step graph: graph_steps consecutive steps (more than one for small
  // direct-sum systems, whose steps are short enough for the graph launch to show)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 1;

  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  nb_timing tim{};
};

// Release a captured step graph. Launches of it may still be queued on `s`,
// and the executable graph owns the kernel-argument memory they read, so the
// stream is drained first.
inline void nb_drop_graph(hipStream_t s, hipGraphExec_t& exec, hipGraph_t& graph) {
  if (exec || graph) (void)hipStreamSynchronize(s);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
}

// ---- launchers (return hipError_t of the launch) ---------------------------
hipError_t nb_launch_direct(nb_ctx* c);  // pm -> force
hipError_t nb_launch_integrate(nb_ctx* c);  // force + state -> state, pm
hipError_t nb_launch_refresh_pm(nb_ctx* c);  // pos, mass -> pm
hipError_t nb_launch_fill_i32(int* p, int v, int64_t count, hipStream_t s);
hipError_t nb_launch_render_view(nb_ctx* c, float4* out);  // (x, y, z, |v|) per slot
hipError_t nb_launch_energy(nb_ctx* c, double* ekin_p, double* epot_p);
// FP64 direct-sum accelerations of the bodies in slots[0..k) -> acc3[3k]
hipError_t nb_launch_ref_f64(nb_ctx* c, const int* slots, int k, double* acc3);

// Barnes-Hut pipeline
int nb_sort_alloc(nb_ctx* c);  // bbox, Morton and radix-sort buffers only
int nb_bh_alloc(nb_ctx* c);    // everything, one allocation
// Point every Barnes-Hut array of c (n bodies) into `base`, laid out by
// nb_bh_layout_for(n); `upto` = kBhBufCount for all of them, kBhPms for the
// sort buffers only. Sets max_nodes, sort_blocks, bbox_blocks, persist_blocks.
void nb_bh_carve(nb_ctx* c, char* base, int upto);
hipError_t nb_morton_sort(nb_ctx* c);  // bbox, Morton codes, stable sort: perm in bh.vals[0]
// slots of pos/vel/acc/force/mass take the bodies of bh.vals[0] (staging in c->tmp)
hipError_t nb_launch_permute_state(nb_ctx* c);
void nb_bh_free(nb_ctx* c);
hipError_t nb_bh_build(nb_ctx* c);     // bbox, Morton, sort, permute, prefix, octree
hipError_t nb_bh_traverse(nb_ctx* c, bool count_interactions);  // -> force
// Per-body traversal with the default (modified) opening test, counting only:
// the interactions the group traversal's test needs body by body -> bh.inter
hipError_t nb_bh_count_per_body(nb_ctx* c);

// Launch-shape defaults: fill the zero fields from the autotuner's table, then
// from the built-in defaults.
bool nb_tuned_shape(int64_t n, int* bpl, int* jsplit, int* leaf);
void nb_default_shape(int64_t n, int* bpl, int* jsplit, int* leaf);
bool nb_valid_shape(int bpl, int jsplit);  // a direct-sum kernel exists for it

// step = forces + integration (no sync)
hipError_t nb_enqueue_step(nb_ctx* c);
// Capture and instantiate the one-step graph if there is none (launches nothing).
int nb_prepare_step_graph(nb_ctx* c);
// Reuse a Barnes-Hut context for another system of the same solver:
// cfg->n <= c->cap_n; the buffers are re-laid out for cfg->n, the captured step
// graph is dropped and the state must be uploaded again. Allocates nothing.
int nb_ctx_rebind(nb_ctx* c, const nb_config* cfg);
