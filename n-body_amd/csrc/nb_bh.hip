// nb_bh.hip — Barnes-Hut force pass on gfx950.
//
// Replaces the reference's GPU octree (ParallelOctreeGPU.cpp:77-300 with the
// shaders boundingBoxPass1/2, resetOctree, expandOctree, countParticlesPerTask,
// prefixSum, addBlockSums, distributeParticles, insertParticles,
// propagateFatherOctree, pruneFather, pruneSubtrees) and the traversal shader
// forceCalcuBarnesHut.glsl:54-118, driven by ParticleSolverBHutGPU.cpp:25-50.
// It uses the reference's dormant Morton encoding (morton.glsl:37-98) as the
// spatial key instead of top-down bucketing.
//
// Pipeline per force pass (all on one stream, no host synchronisation, so the
// whole step can be captured into a hipGraph):
//   1. bbox        per-block min/max, then one block folds the partials and
//                  writes the root box and the cell side^2 of each level.
//   2. encode      30-bit Morton code per body (bit-exact restatement of
//                  morton.glsl: IEEE divide, clamp, 10-bit quantisation).
//   3. sort        LSD radix sort of (code, body index), 4 passes of 8 bits.
//                  Each pass: block digit histograms, a device-wide exclusive
//                  scan, and a stable scatter whose per-wave ranks come from
//                  64-bit ballots (8 ballots give each lane the mask of lanes
//                  holding the same digit). Stable => sorted by (code, index).
//   4. permute     gather (x, y, z, m) into sorted order.
//   5. prefix      FP64 exclusive prefix of (m, m x, m y, m z) over sorted bodies:
//                  any node's mass and centre of mass is a difference of two
//                  prefix entries (cells are contiguous code ranges).
//   6. octree      breadth-first, one level per round: a node with more than
//                  leaf_size bodies (and level < 10) splits into its non-empty
//                  Morton-digit children, found by binary search in its code
//                  range; children get consecutive indices from a scan of child
//                  counts, so numbering is deterministic. Ropes (`next`) point to
//                  the next sibling, or the parent's rope for the last child —
//                  the firstChild/next contract of Octree.cpp:264-294.
//   7. com         per node from the prefix; single-body nodes copy the body.
//   8. traverse    one lane per sorted body, stackless rope walk with the
//                  reference's opening test s^2 < theta^2 d^2 (Octree.cpp:320);
//                  a leaf that is not accepted is evaluated body by body.
#include <algorithm>
#include <cfloat>
#include <type_traits>
#include <vector>

#include "nb_force_core.h"
#include "nb_internal.h"

namespace {

// Block and tile sizes shared with the buffer layout are in nb_internal.h. Sort
// tiles of 2048 keys (8 per lane) measured 40 us faster per 4M-body sort than
// 4096 (more tiles in flight) and than 1024 (larger histograms).
constexpr int kScanThreads = 256;
// Traversal groups: each wave of the traversal takes one "wave range" of at
// most kWaveBodies consecutive sorted bodies. Ranges are cut along the tree: the
// children of a node holding more than kWaveBodies bodies are packed in order
// into ranges of at most kWaveBodies bodies (a child that itself holds more is
// handled at its own level; a level-10 leaf that holds more is split into
// chunks), so every range lies inside one parent cell and its bounding box
// stays tight. (With leaf_size <= kWaveBodies only level-10 leaves can hold
// more than kWaveBodies bodies.) A wave holds 64 / P copies of its range (P = the
// range's body count rounded up to a power of two), so at 32 bodies every range
// is evaluated by at least two slices: against 64-body ranges this measured 8%
// fewer interactions and a 4% faster traversal on the 4M two-disc system (16
// bodies: 14% slower, the walks dominate).
constexpr int kWaveBodies = NB_BH_WAVE_BODIES;
static_assert(kWaveBodies >= 4 && kWaveBodies <= 32 && (kWaveBodies & (kWaveBodies - 1)) == 0,
              "two wave ranges fit one wave");

// ---- 0. wave64 prefix sums on DPP ----------------------------------------------------
// Inclusive prefix over the 64 lanes of a wave, entirely in cross-lane DPP adds:
// shifts by 1, 2, 4 and 8 lanes inside each row of 16 (row_shr; a lane whose
// source lies outside its row adds 0), then lane 15 of rows 0 and 2 is added to
// rows 1 and 3 (row_bcast:15) and lane 31 to rows 2 and 3 (row_bcast:31). Six
// adds and no LDS round trips; the association is fixed, so FP64 prefixes are
// deterministic.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int dpp_shift(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, kRowMask, 0xf, false);
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_shift(double v) {  // the two halves moved alike
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, kCtrl, kRowMask, 0xf, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), kCtrl, kRowMask, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  v += dpp_shift<0x111, 0xf>(v);  // row_shr:1
  v += dpp_shift<0x112, 0xf>(v);  // row_shr:2
  v += dpp_shift<0x114, 0xf>(v);  // row_shr:4
  v += dpp_shift<0x118, 0xf>(v);  // row_shr:8
  v += dpp_shift<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_shift<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Exclusive prefix of v over the threads of a kScanThreads-thread block (wave
// prefixes, then the wave totals added in wave order); total = the block's sum.
// wsum: kScanThreads / 64 ints of LDS, free again when this returns.
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_incl_scan(v);
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    before += w < wave ? wsum[w] : 0;
    total += wsum[w];
  }
  __syncthreads();
  return before + incl - v;
}

// ---- 1. bounding box --------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbox_partial(const float4* __restrict__ pm, int n,
                                                      float* __restrict__ part) {
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float4 p = pm[i];
    mn[0] = fminf(mn[0], p.x); mx[0] = fmaxf(mx[0], p.x);
    mn[1] = fminf(mn[1], p.y); mx[1] = fmaxf(mx[1], p.y);
    mn[2] = fminf(mn[2], p.z); mx[2] = fmaxf(mx[2], p.z);
  }
  __shared__ float red[6][256];
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = mn[a];
    red[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_bbox_final(const float* __restrict__ part, int nparts,
                                                    float* __restrict__ geom) {
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int b = threadIdx.x; b < nparts; b += 256)
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], part[6 * b + a]);
      mx[a] = fmaxf(mx[a], part[6 * b + 3 + a]);
    }
  __shared__ float red[6][256];
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = mn[a];
    red[3 + a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float side = 0.f;
    for (int a = 0; a < 3; ++a) {
      geom[a] = red[a][0];
      geom[3 + a] = red[3 + a][0];
      const float ext = fmaxf(__fsub_rn(red[3 + a][0], red[a][0]), 1e-6f);
      side = fmaxf(side, ext);
    }
    for (int l = 0; l <= kMaxLevel; ++l) {
      const float s = ldexpf(side, -l);
      geom[6 + l] = __fmul_rn(s, s);
    }
    geom[6 + kMaxLevel + 1] = side;  // cubic root extent
  }
}

// ---- 2. Morton encode (morton.glsl:37-98) ------------------------------------
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Inverse of expand_bits on 10-bit values: gathers bits 0, 3, 6, ... of v.
__device__ __forceinline__ uint32_t compact_bits(uint32_t v) {
  v &= 0x49249249u;
  v = (v ^ (v >> 2)) & 0xC30C30C3u;
  v = (v ^ (v >> 4)) & 0x0F00F00Fu;
  v = (v ^ (v >> 8)) & 0xFF0000FFu;
  v = (v ^ (v >> 16)) & 0x000003FFu;
  return v;
}

// morton.glsl:68-75 and :49-51 with the extent given: n = clamp((p - lo) / ext, 0, 1),
// q = uint(clamp(n * 1024, 0, 1023)).
__device__ __forceinline__ uint32_t quantize(float p, float lo, float ext) {
  float nrm = __fdiv_rn(__fsub_rn(p, lo), ext);
  nrm = fminf(fmaxf(nrm, 0.f), 1.f);
  const float g = fminf(fmaxf(__fmul_rn(nrm, 1024.f), 0.f), 1023.f);
  return (uint32_t)g;
}

// cubic = 1: one extent E = max_a max(max_a - min_a, 1e-6) on every axis (a cubic
// root cell, so every cell is a cube of side E / 2^level); cubic = 0: each axis
// over its own extent max(max_a - min_a, 1e-6), exactly morton.glsl's inputs.
__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ pm, int n, int cubic,
                                                const float* __restrict__ geom,
                                                uint32_t* __restrict__ codes,
                                                uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 p = pm[i];
  const float E = geom[6 + kMaxLevel + 1];
  const float ex = cubic ? E : fmaxf(__fsub_rn(geom[3], geom[0]), 1e-6f);
  const float ey = cubic ? E : fmaxf(__fsub_rn(geom[4], geom[1]), 1e-6f);
  const float ez = cubic ? E : fmaxf(__fsub_rn(geom[5], geom[2]), 1e-6f);
  const uint32_t qx = quantize(p.x, geom[0], ex);
  const uint32_t qy = quantize(p.y, geom[1], ey);
  const uint32_t qz = quantize(p.z, geom[2], ez);
  const uint32_t code = expand_bits(qx) * 4u + expand_bits(qy) * 2u + expand_bits(qz);
  codes[i] = code;
  keys[i] = code;
  vals[i] = (uint32_t)i;
}

// ---- generic reduce-then-scan (int32 and FP64x4) ----------------------------
// The element count is either a constant or the size of one octree level,
// read on the device: count = level_off[l + 1] - level_off[l].
// The count is clamped to the launch bound `cap`, so a bad level table can
// never send the scan past its buffers.
__device__ __forceinline__ int scan_count(const int* level_off, int l, int n_const, int cap) {
  const int c = level_off ? level_off[l + 1] - level_off[l] : n_const;
  return c < 0 ? 0 : (c > cap ? cap : c);
}

// Pass 1 of an exclusive int scan: the sum of every kScanTile-element tile.
__global__ __launch_bounds__(kScanThreads) void k_iscan_reduce(const int* __restrict__ in,
                                                               const int* level_off, int l,
                                                               int n_const, int cap,
                                                               int* __restrict__ part) {
  const int n = scan_count(level_off, l, n_const, cap);
  const int base = blockIdx.x * kScanTile;
  int s = 0;
  for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
    const int e = base + k;
    if (e < n) s += in[e];
  }
  __shared__ int wsum[kScanThreads / 64];
  int tile = 0;
  (void)block_excl_scan(s, wsum, tile);
  if (threadIdx.x == 0) part[blockIdx.x] = tile;
}

// Pass 2: every tile adds up the tile sums before it (a few KB of L2 reads per
// tile, which replaces a separate one-block scan of the sums and its launch),
// then rescans its own elements from there (8 consecutive elements per thread,
// one block scan of the thread sums). Block 0 also adds up all nparts tile sums
// into part[nparts] and, for a level, publishes the next level's end:
// level_off[l + 2] = level_off[l + 1] + total.
__global__ __launch_bounds__(kScanThreads) void k_iscan_final(const int* __restrict__ in,
                                                              const int* level_off, int l,
                                                              int n_const, int cap,
                                                              int* __restrict__ part, int nparts,
                                                              int* level_off_mut,
                                                              int* __restrict__ out) {
  const int n = scan_count(level_off, l, n_const, cap);
  const int base = blockIdx.x * kScanTile;
  __shared__ int wsum[kScanThreads / 64];
  if (blockIdx.x == 0) {
    int all = 0;
    for (int b = threadIdx.x; b < nparts; b += kScanThreads) all += part[b];
    int total = 0;
    (void)block_excl_scan(all, wsum, total);
    if (threadIdx.x == 0) {
      part[nparts] = total;
      if (level_off_mut) level_off_mut[l + 2] = level_off[l + 1] + total;
    }
  }
  if (base >= n) return;
  int pre = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanThreads) pre += part[b];
  int start = 0;
  (void)block_excl_scan(pre, wsum, start);
  constexpr int kPer = kScanTile / kScanThreads;  // 8 consecutive elements per thread
  int v[kPer];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = base + threadIdx.x * kPer + k;
    v[k] = e < n ? in[e] : 0;
    s += v[k];
  }
  int total = 0;
  int run = start + block_excl_scan(s, wsum, total);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = base + threadIdx.x * kPer + k;
    if (e < n) out[e] = run;
    run += v[k];
  }
}

// Exclusive scan of at most kSingleScan elements by one workgroup (8 consecutive
// elements per thread), for the small scans: one launch instead of two. Writes
// level_off[l + 2] like k_iscan_final.
constexpr int kSingleScan = 8192;
__global__ __launch_bounds__(1024) void k_iscan_single(const int* __restrict__ in,
                                                       const int* level_off, int l, int n_const,
                                                       int cap, int* __restrict__ out,
                                                       int* level_off_mut) {
  const int n = scan_count(level_off, l, n_const, cap);
  constexpr int kPer = kSingleScan / 1024;
  const int base = threadIdx.x * kPer;
  int v[kPer];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    sum += v[k];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int incl = wave_incl_scan(sum);
  __shared__ int wsum[16];
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0, total = 0;
  for (int w = 0; w < 16; ++w) {  // fixed order
    before += w < wave ? wsum[w] : 0;
    total += wsum[w];
  }
  int run = before + incl - sum;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0 && level_off_mut) level_off_mut[l + 2] = level_off[l + 1] + total;
}

// FP64 prefix of (m, m x, m y, m z): prefix[k] = sum over sorted bodies < k.
__device__ __forceinline__ double4 body_moments(float4 p) {
  const double m = p.w;
  return make_double4(m, m * (double)p.x, m * (double)p.y, m * (double)p.z);
}
__device__ __forceinline__ double4 add4(double4 a, double4 b) {
  return make_double4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// Tiles of kScanTile = 8 rows x 256 bodies: thread t holds the bodies t, 256 + t,
// ..., so every wave load and store is one contiguous 1 KB (float4) or 2 KB
// (double4) run.
__global__ __launch_bounds__(kScanThreads) void k_mscan_reduce(const float4* __restrict__ pms,
                                                               int n, double4* __restrict__ part) {
  constexpr int kRows = kScanTile / kScanThreads;
  const int base = blockIdx.x * kScanTile;
  double4 s = make_double4(0, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int e = base + r * kScanThreads + threadIdx.x;
    if (e < n) s = add4(s, body_moments(pms[e]));
  }
  __shared__ double4 sh[kScanThreads];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kScanThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = add4(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__device__ __forceinline__ double4 shfl_up4(double4 v, int off) {
  return make_double4(__shfl_up(v.x, off), __shfl_up(v.y, off), __shfl_up(v.z, off),
                      __shfl_up(v.w, off));
}
__device__ __forceinline__ double4 wave_incl_scan4(double4 v) {
  return make_double4(wave_incl_scan(v.x), wave_incl_scan(v.y), wave_incl_scan(v.z),
                      wave_incl_scan(v.w));
}

// Exclusive scan of the block sums by one block, in a fixed order: each thread
// folds a contiguous run of partials, the 256 run totals are scanned (wave DPP
// prefixes, then the 4 wave totals in wave order), then each thread rewrites its
// run. part[nparts] receives the total.
__global__ __launch_bounds__(kScanThreads) void k_mscan_partials(double4* __restrict__ part,
                                                                 int nparts) {
  const int per = (nparts + kScanThreads - 1) / kScanThreads;
  const int b0 = min((int)threadIdx.x * per, nparts);
  const int b1 = min(b0 + per, nparts);
  double4 tot = make_double4(0, 0, 0, 0);
  for (int b = b0; b < b1; ++b) tot = add4(tot, part[b]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double4 incl = wave_incl_scan4(tot);
  __shared__ double4 wtot[kScanThreads / 64];
  if (lane == 63) wtot[wave] = incl;
  __syncthreads();
  double4 run = make_double4(0, 0, 0, 0), grand = make_double4(0, 0, 0, 0);
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wave) run = add4(run, wtot[w]);
    grand = add4(grand, wtot[w]);
  }
  const double4 below = shfl_up4(incl, 1);  // inclusive value of the lane below
  if (lane > 0) run = add4(run, below);
  for (int b = b0; b < b1; ++b) {
    const double4 v = part[b];
    part[b] = run;
    run = add4(run, v);
  }
  if (threadIdx.x == 0) part[nparts] = grand;
}

// Exclusive prefix of each tile, row by row: a row of 256 bodies is scanned by
// wave DPP prefixes and a fixed-order combine of the 4 wave totals; rows follow
// each other in order. Deterministic (no atomics, fixed association).
__global__ __launch_bounds__(kScanThreads) void k_mscan_final(const float4* __restrict__ pms,
                                                              int n,
                                                              const double4* __restrict__ part,
                                                              double4* __restrict__ prefix) {
  constexpr int kRows = kScanTile / kScanThreads;
  constexpr int kWaves = kScanThreads / 64;
  const int base = blockIdx.x * kScanTile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ double4 wtot[kRows][kWaves];
  double4 incl[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int e = base + r * kScanThreads + threadIdx.x;
    const double4 x0 = e < n ? body_moments(pms[e]) : make_double4(0, 0, 0, 0);
    const double4 x = wave_incl_scan4(x0);
    if (lane == 63) wtot[r][wave] = x;
    const double4 below = shfl_up4(x, 1);  // inclusive value of the lane below
    incl[r] = lane > 0 ? below : make_double4(0, 0, 0, 0);
  }
  __syncthreads();
  double4 run = part[blockIdx.x];  // exclusive start of this tile
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    double4 before = run;  // tile start + earlier rows + earlier waves of this row
    for (int w = 0; w < kWaves; ++w) {
      if (w < wave) before = add4(before, wtot[r][w]);
      run = add4(run, wtot[r][w]);
    }
    const int e = base + r * kScanThreads + threadIdx.x;
    if (e < n) prefix[e] = add4(before, incl[r]);  // incl[r]: exclusive within the wave
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) prefix[n] = part[gridDim.x];
}

// ---- 3. radix sort ------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             int n, int shift, int nblocks,
                                                             int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortItems; ++r) {
    const int e = base + r * kSortThreads + threadIdx.x;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Mask of the active lanes whose 8-bit digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
  uint64_t m = active;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// Stable scatter of one 8-bit digit pass. The tile is first reordered in LDS by
// digit (stable ranks from per-wave ballots, waves and rounds in element
// order), then written out in LDS order: consecutive threads write the runs of
// equal digits to consecutive global addresses instead of 64 scattered lines
// per wave store.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, int n, int shift,
    int nblocks, const int* __restrict__ offsets, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out) {
  __shared__ int whist[4][256];
  __shared__ int wrun[4][256];
  __shared__ int local_start[256];  // first tile slot of each digit
  __shared__ int global_start[256];  // first output index of each digit for this tile
  __shared__ uint32_t sk[kSortTile], sv[kSortTile];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int w = 0; w < 4; ++w) whist[w][threadIdx.x] = 0;
  __syncthreads();

  // Wave w owns tile elements [w*kSortTile/4, (w+1)*kSortTile/4), 64 consecutive per round.
  const int tile0 = blockIdx.x * kSortTile;
  const int base = tile0 + wave * (kSortTile / 4);
  const uint64_t lt = (uint64_t(1) << lane) - 1;
  // The rounds read their keys again in phase 2 (coalesced, 4 B per key)
  // instead of holding 16 keys and values per lane: the kernel then needs a few
  // dozen VGPRs instead of hundreds, and many waves per SIMD hide the latency.
  // phase 1: per-wave digit counts
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int e = base + r * 64 + lane;
    const uint64_t active = __ballot(e < n);
    const uint32_t d = ((e < n ? keys_in[e] : 0u) >> shift) & 255u;
    const uint64_t m = match_digit(d, active);
    if (e < n && (m & lt) == 0) whist[wave][d] += __popcll(m);
  }
  __syncthreads();
  {
    // tile-local digit starts: exclusive scan of the tile's digit counts
    const int d = threadIdx.x;
    const int cnt = whist[0][d] + whist[1][d] + whist[2][d] + whist[3][d];
    const int incl = wave_incl_scan(cnt);
    __shared__ int wsum[4];
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int start = incl - cnt;
    for (int w = 0; w < wave; ++w) start += wsum[w];
    local_start[d] = start;
    global_start[d] = offsets[d * nblocks + blockIdx.x];
    int run = start;
    for (int w = 0; w < 4; ++w) {
      wrun[w][d] = run;
      run += whist[w][d];
    }
  }
  __syncthreads();
  // phase 2: stable tile-local ranks, reorder in LDS
#pragma unroll 1
  for (int r = 0; r < kSortItems; ++r) {
    const int e = base + r * 64 + lane;
    const uint64_t active = __ballot(e < n);
    const uint32_t key = e < n ? keys_in[e] : 0u;
    const uint32_t val = e < n ? vals_in[e] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    const uint64_t m = match_digit(d, active);
    if (e < n) {
      const int slot = wrun[wave][d] + __popcll(m & lt);
      if (slot >= 0 && slot < kSortTile) {  // always true for a consistent histogram
        sk[slot] = key;
        sv[slot] = val;
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (e < n && (m & lt) == 0) wrun[wave][d] += __popcll(m);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // phase 3: write out in tile order
  const int count = min(kSortTile, n - tile0);
#pragma unroll 4
  for (int i = threadIdx.x; i < count; i += kSortThreads) {
    const uint32_t key = sk[i];
    const uint32_t d = (key >> shift) & 255u;
    const int dst = global_start[d] + (i - local_start[d]);
    if (dst >= 0 && dst < n) {
      keys_out[dst] = key;
      vals_out[dst] = sv[i];
    }
  }
}

// ---- 4. permute -------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(const float4* __restrict__ pm,
                                                const uint32_t* __restrict__ perm, int n,
                                                float4* __restrict__ pms) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const uint32_t p = perm[i];
    pms[i] = p < (uint32_t)n ? pm[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Morton reordering of the resident state: dst[k] = src[perm[k]] per component.
__global__ __launch_bounds__(256) void k_permute_d3(const double* __restrict__ src,
                                                    const uint32_t* __restrict__ perm, int n,
                                                    double* __restrict__ dst) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const uint32_t p = perm[i];
    if (p < (uint32_t)n) {
      dst[i] = src[p];
      dst[n + i] = src[n + p];
      dst[2 * n + i] = src[2 * n + p];
    }
  }
}

__global__ __launch_bounds__(256) void k_permute_f32(const float* __restrict__ src,
                                                     const uint32_t* __restrict__ perm, int n,
                                                     float* __restrict__ dst) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const uint32_t p = perm[i];
    if (p < (uint32_t)n) dst[i] = src[p];
  }
}

// ---- 6. octree ------------------------------------------------------------------------
__device__ __forceinline__ int link_count(int w) { return w >> 4; }
__device__ __forceinline__ int link_level(int w) { return w & 15; }

// First index in [lo, hi) whose digit (codes >> shift) & 7 is >= d (digits are sorted).
__device__ __forceinline__ int digit_lower_bound(const uint32_t* __restrict__ codes, int lo,
                                                 int hi, int shift, uint32_t d) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (((codes[mid] >> shift) & 7u) < d) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Boundaries of the 8 digit cells of the sorted range [s, e): bound[d] = first
// index whose digit is >= d (bound[0] = s, bound[8] = e). The seven binary
// searches run in lockstep, so each step issues seven independent loads and the
// latency is log2(e - s) loads rather than seven times that (dense nodes near
// the disc centres hold 10^5 bodies). Counting the digits of short ranges with
// independent loads instead was measured slower.
__device__ __forceinline__ void digit_bounds(const uint32_t* __restrict__ codes, int s, int e,
                                             int shift, int (&bound)[9]) {
  int lo[7], hi[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    lo[d] = s;
    hi[d] = e;
  }
  for (int len = e - s; len > 0; len >>= 1) {  // ceil(log2) + 1 rounds cover every search
    uint32_t v[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) v[d] = lo[d] < hi[d] ? codes[(lo[d] + hi[d]) >> 1] : 0u;
#pragma unroll
    for (int d = 0; d < 7; ++d)
      if (lo[d] < hi[d]) {
        const int mid = (lo[d] + hi[d]) >> 1;
        if (((v[d] >> shift) & 7u) < (uint32_t)(d + 1)) lo[d] = mid + 1;
        else hi[d] = mid;
      }
  }
  // a search whose interval is not empty yet (at most one more halving) finishes here
#pragma unroll
  for (int d = 0; d < 7; ++d)
    while (lo[d] < hi[d]) {
      const int mid = (lo[d] + hi[d]) >> 1;
      if (((codes[mid] >> shift) & 7u) < (uint32_t)(d + 1)) lo[d] = mid + 1;
      else hi[d] = mid;
    }
  bound[0] = s;
  bound[8] = e;
#pragma unroll
  for (int d = 0; d < 7; ++d) bound[d + 1] = lo[d];
}

// Counts the non-empty children of every internal node of level l and keeps
// their boundaries (bound[1..7]; bound[0] and bound[8] are the node's own range)
// for k_tree_emit, which would otherwise repeat the seven binary searches.
__global__ __launch_bounds__(256) void k_tree_count(const uint32_t* __restrict__ codes, int n,
                                                    const int4* __restrict__ link,
                                                    const int* __restrict__ level_off, int l,
                                                    int leaf, int cap,
                                                    int* __restrict__ child_cnt,
                                                    int* __restrict__ child_bounds) {
  const int lo = level_off[l];
  const int hi = min(level_off[l + 1], lo + cap);  // cap = launch bound of this level
  const int shift = 27 - 3 * l;
  for (int k = lo + blockIdx.x * 256 + threadIdx.x; k < hi; k += gridDim.x * 256) {
    const int4 nd = link[k];
    const int cnt = link_count(nd.w);
    int nc = 0;
    if (cnt > leaf && nd.z >= 0 && nd.z + cnt <= n) {
      int bound[9];
      digit_bounds(codes, nd.z, nd.z + cnt, shift, bound);
#pragma unroll
      for (int d = 0; d < 8; ++d) nc += bound[d + 1] > bound[d] ? 1 : 0;
#pragma unroll
      for (int d = 1; d < 8; ++d) child_bounds[8 * (k - lo) + d] = bound[d];
    }
    child_cnt[k - lo] = nc;
  }
}

// Wave ranges of one node's children (see kWaveBodies): calls out(start, count)
// for each range in order; returns how many there are.
template <typename Out>
__device__ __forceinline__ int pack_ranges(const int* starts, const int* ends, int nc, int l,
                                           int leaf, Out out) {
  int r = 0, run_s = 0, run_n = 0;
  for (int q = 0; q < nc; ++q) {
    const int cn = ends[q] - starts[q];
    if (cn > kWaveBodies) {
      if (run_n) { out(r++, run_s, run_n); }
      run_n = 0;
      if (l + 1 == kMaxLevel || cn <= leaf)  // a leaf: chunks (internal: its own level)
        for (int o = 0; o < cn; o += kWaveBodies) out(r++, starts[q] + o, min(kWaveBodies, cn - o));
    } else {
      if (run_n + cn > kWaveBodies) {
        out(r++, run_s, run_n);
        run_n = 0;
      }
      if (run_n == 0) run_s = starts[q];
      run_n += cn;
    }
  }
  if (run_n) out(r++, run_s, run_n);
  return r;
}

// P lanes per copy of a range of `count` bodies: a power of two >= count, >= 4.
__device__ __forceinline__ int copy_lanes(int count) {
  return count <= 4 ? 4 : 1 << (32 - __clz(count - 1));
}

// Number of records emit_records makes of one node's wave ranges.
__device__ __forceinline__ int count_records(const int* starts, const int* ends, int nc, int l,
                                             int leaf, bool pairs) {
  if (!pairs) return pack_ranges(starts, ends, nc, l, leaf, [](int, int, int) {});
  int per_width[4] = {0, 0, 0, 0};  // ranges of width 4, 8, 16, 32
  pack_ranges(starts, ends, nc, l, leaf, [&](int, int, int c) {
    const int k = 29 - __clz(copy_lanes(c));
#pragma unroll
    for (int j = 0; j < 4; ++j) per_width[j] += j == k ? 1 : 0;
  });
  return ((per_width[0] + 1) >> 1) + ((per_width[1] + 1) >> 1) + ((per_width[2] + 1) >> 1) +
         ((per_width[3] + 1) >> 1);
}

// One node's wave ranges (pack_ranges) as records of two sibling ranges of the
// same copy width (a pair shares one lane layout, so neither range idles
// lanes), (start, count, start, count), or of one range, (start, count, 0, 0):
// each range waits for the next of its width, in order. Writes count_records()
// records from record `first` on. There are at most as many records as ranges,
// and at most n ranges. Without `pairs` every range is a record of its own.
__device__ __forceinline__ void emit_records(int4* __restrict__ waves, int first, int n,
                                             const int* starts, const int* ends, int nc, int l,
                                             int leaf, bool pairs) {
  int2 wait[4];  // one waiting range per width 4, 8, 16, 32
#pragma unroll
  for (int k = 0; k < 4; ++k) wait[k] = make_int2(0, 0);
  int r = 0;
  auto put = [&](int4 v) {
    if (first + r < n) waves[first + r] = v;
    ++r;
  };
  if (!pairs) {
    pack_ranges(starts, ends, nc, l, leaf, [&](int, int a, int c) { put(make_int4(a, c, 0, 0)); });
    return;
  }
  pack_ranges(starts, ends, nc, l, leaf, [&](int, int a, int c) {
    const int k = 29 - __clz(copy_lanes(c));  // 4 -> 0, ..., 32 -> 3
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (j == k) {
        if (wait[j].y > 0) {
          put(make_int4(wait[j].x, wait[j].y, a, c));
          wait[j] = make_int2(0, 0);
        } else {
          wait[j] = make_int2(a, c);
        }
      }
  });
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (wait[j].y > 0) put(make_int4(wait[j].x, wait[j].y, 0, 0));
}

__global__ __launch_bounds__(256) void k_tree_emit(const uint32_t* __restrict__ codes, int n,
                                                   int4* __restrict__ link,
                                                   const int* __restrict__ level_off, int l,
                                                   int leaf, int cap,
                                                   const int* __restrict__ child_off,
                                                   const int* __restrict__ child_bounds,
                                                   int64_t max_nodes, int* __restrict__ flags,
                                                   int4* __restrict__ waves,
                                                   int* __restrict__ wctl, int pairs) {
  const int lo = level_off[l];
  const int hi = min(level_off[l + 1], lo + cap);
  const int next_lo = level_off[l + 1];
  const int lane = threadIdx.x & 63;
  // Rounds are uniform across the block, so every lane of a wave reaches the
  // wave-wide range allocation below.
  for (int base = lo + blockIdx.x * 256; base < hi; base += gridDim.x * 256) {
    const int k = base + threadIdx.x;
    int starts[8], ends[8], nc = 0;
    int4 nd = make_int4(-1, -1, -1, 0);
    bool internal = false;
    if (k < hi) {
      nd = link[k];
      const int cnt = link_count(nd.w);
      internal = cnt > leaf && nd.z >= 0 && nd.z + cnt <= n;
    }
    if (internal) {
      int bound[9];  // as k_tree_count found them
      bound[0] = nd.z;
      bound[8] = nd.z + link_count(nd.w);
#pragma unroll
      for (int d = 1; d < 8; ++d) bound[d] = child_bounds[8 * (k - lo) + d];
#pragma unroll
      for (int d = 0; d < 8; ++d)
        if (bound[d + 1] > bound[d]) {
          starts[nc] = bound[d];
          ends[nc] = bound[d + 1];
          ++nc;
        }
      const int first = next_lo + child_off[k - lo];
      if ((int64_t)first + nc > max_nodes) {
        atomicOr(flags, 2);
        internal = false;
      } else {
        for (int q = 0; q < nc; ++q) {
          const int idx = first + q;
          const int nxt = q + 1 < nc ? idx + 1 : nd.y;
          link[idx] = make_int4(-1, nxt, starts[q], ((ends[q] - starts[q]) << 4) | (l + 1));
        }
        link[k].x = first;
      }
    }
    // wave ranges of the children of nodes holding more than kWaveBodies bodies,
    // in records of two: one atomic per wave reserves the records of all its lanes
    const bool packs = internal && link_count(nd.w) > kWaveBodies;
    const int nrec = packs ? count_records(starts, ends, nc, l, leaf, pairs) : 0;
    const int incl = wave_incl_scan(nrec);
    const int total = __shfl(incl, 63);
    int wbase = 0;
    if (lane == 63 && total > 0) wbase = atomicAdd(&wctl[0], total);
    wbase = __shfl(wbase, 63);
    if (nrec > 0) emit_records(waves, wbase + incl - nrec, n, starts, ends, nc, l, leaf, pairs);
  }
}

// ---- 6a. the top levels from per-cell searches ---------------------------------------
// Levels 0..kTopLevels are built without the level loop: a cell of level l is a
// Morton prefix of 3l bits, so its body range in the sorted codes is found by a
// lower-bound search for its smallest code, independently for every cell.
// A cell is a node when it is the root, or it holds bodies and its parent holds
// more than leaf_size; nodes are numbered breadth-first with children in digit
// order, exactly as the level loop numbers them. The loop then continues from
// level kTopLevels.

__host__ __device__ constexpr int top_cell_base(int l) {  // first cell of level l
  return l == 0 ? 0 : top_cell_base(l - 1) + (1 << (3 * (l - 1)));
}
__host__ __device__ constexpr int top_entry_base(int l) {  // first entry of level l
  return top_cell_base(l) + l;
}

// top_start[top_entry_base(l) + c] = first sorted body of cell c of level l
// (c = 8^l gives n).
__global__ __launch_bounds__(256) void k_top_bounds(const uint32_t* __restrict__ codes, int n,
                                                    int* __restrict__ top_start) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= kTopEntries) return;
  int l = 0;
  while (l < kTopLevels && e >= top_entry_base(l + 1)) ++l;
  const uint32_t c = (uint32_t)(e - top_entry_base(l));
  if (c == (1u << (3 * l))) {
    top_start[e] = n;
    return;
  }
  const uint32_t key = c << (30 - 3 * l);
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (codes[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  top_start[e] = lo;
}

// One workgroup: node links of levels 0..kTopLevels, level offsets, the wave
// ranges of the nodes above kTopLevels, and the range counter / traversal ticket.
__global__ __launch_bounds__(1024) void k_top_build(const int* __restrict__ top_start, int n,
                                                    int leaf, int4* __restrict__ link,
                                                    int* __restrict__ level_off,
                                                    int4* __restrict__ waves,
                                                    int* __restrict__ wctl,
                                                    int* __restrict__ rec_hist, int pairs) {
  __shared__ int s_idx[kTopCells];   // exclusive scan of existence, then node index
  __shared__ int s_next[kTopCells];  // rope of each node
  __shared__ int s_part[1024];
  __shared__ int s_ranges;
  const int tid = threadIdx.x;
  auto cell_count = [&](int l, int c) {
    const int e = top_entry_base(l) + c;
    return top_start[e + 1] - top_start[e];
  };
  auto exists = [&](int l, int c) {
    return l == 0 || (cell_count(l, c) > 0 && cell_count(l - 1, c >> 3) > leaf);
  };
  // 1. existence flags and their exclusive scan (cells in level order)
  constexpr int kPer = (kTopCells + 1023) / 1024;
  int flag[kPer];
  int sum = 0;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int g = tid * kPer + q;
    int f = 0;
    if (g < kTopCells) {
      int l = 0;
      while (l < kTopLevels && g >= top_cell_base(l + 1)) ++l;
      f = exists(l, g - top_cell_base(l)) ? 1 : 0;
    }
    flag[q] = f;
    sum += f;
  }
  s_part[tid] = sum;
  if (tid == 0) s_ranges = 0;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the per-thread sums
    const int v = tid >= off ? s_part[tid - off] : 0;
    __syncthreads();
    s_part[tid] += v;
    __syncthreads();
  }
  int run = s_part[tid] - sum;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int g = tid * kPer + q;
    if (g < kTopCells) s_idx[g] = flag[q] ? run : -1;
#pragma unroll
    for (int l = 0; l <= kTopLevels; ++l)  // level_off[l] = nodes in the levels above l
      if (g == top_cell_base(l)) level_off[l] = run;
    run += flag[q];
  }
  if (tid == 1023) level_off[kTopLevels + 1] = s_part[1023];
  __syncthreads();
  if (tid < kCostBuckets) rec_hist[tid] = 0;  // this pass's traversal order starts empty
  if (tid == 0) {
    wctl[1] = 0;  // traversal ticket
    if (n <= kWaveBodies) {
      waves[0] = make_int4(0, n, 0, 0);
      s_ranges = 1;
    } else if (n <= leaf) {
      // the root is itself a leaf, so no node packs ranges below it: chunk it,
      // two full chunks per record, a shorter last chunk on its own
      int r = 0;
      for (int o = 0; o < n; o += kWaveBodies) {
        const int k = min(kWaveBodies, n - o), o2 = o + kWaveBodies;
        if (pairs && k == kWaveBodies && n - o2 >= kWaveBodies) {
          waves[r++] = make_int4(o, k, o2, kWaveBodies);
          o = o2;
        } else {
          waves[r++] = make_int4(o, k, 0, 0);
        }
      }
      s_ranges = r;
    }
  }
  // 2. links, level by level (a node's rope falls back to its parent's)
  for (int l = 0; l <= kTopLevels; ++l) {
    const int base = top_cell_base(l), cells = 1 << (3 * l);
    for (int c = tid; c < cells; c += 1024) {
      const int i = s_idx[base + c];
      if (i < 0) continue;
      const int cnt = cell_count(l, c);
      const int start = top_start[top_entry_base(l) + c];
      int nxt = l == 0 ? -1 : s_next[top_cell_base(l - 1) + (c >> 3)];
      for (int sib = c + 1; sib < (c | 7) + 1 && l > 0; ++sib)
        if (s_idx[base + sib] >= 0) {
          nxt = s_idx[base + sib];
          break;
        }
      s_next[base + c] = nxt;
      int first = -1;
      const bool internal = cnt > leaf && l < kMaxLevel;
      if (internal && l < kTopLevels) {
        const int cb = top_cell_base(l + 1) + 8 * c;
        for (int d = 0; d < 8; ++d)
          if (s_idx[cb + d] >= 0) {
            first = s_idx[cb + d];
            break;
          }
      }
      link[i] = make_int4(first, nxt, start, (cnt << 4) | l);
      // wave ranges of the children of a node above kTopLevels holding more
      // than kWaveBodies bodies (level kTopLevels is done by the level loop)
      if (internal && l < kTopLevels && cnt > kWaveBodies) {
        int starts[8], ends[8], nc = 0;
        const int eb = top_entry_base(l + 1) + 8 * c;
        for (int d = 0; d < 8; ++d)
          if (top_start[eb + d + 1] > top_start[eb + d]) {
            starts[nc] = top_start[eb + d];
            ends[nc] = top_start[eb + d + 1];
            ++nc;
          }
        const int nrec = count_records(starts, ends, nc, l, leaf, pairs);
        emit_records(waves, atomicAdd(&s_ranges, nrec), n, starts, ends, nc, l, leaf, pairs);
      }
    }
    __syncthreads();
  }
  if (tid == 0) wctl[0] = s_ranges;
}

// ---- 7. centre of mass and opening radius -------------------------------------------
// Squared opening radius of the default traversal: r = s_l / theta + delta, where
// s_l is the cell side of the reference's test (Octree.cpp:311-320) and delta the
// distance from the COM to the cell's geometric centre. This is the offset term of
// Barnes' modified criterion (J. Comput. Phys. 87, 1990; the default of his
// treecode 1.4): a cell whose mass sits near one face can hold bodies much closer
// to a target than s_l / theta suggests. The cell is the 3l-bit prefix of its first
// body's code, so its centre is bmin + (index + 1/2) * extent / 2^l per axis.
// Every operation is one IEEE operation rounded to nearest, in a fixed order and
// with contraction into FMAs off (HIP's __fadd_rn and __fmul_rn are plain
// operators that -ffp-contract=fast may fuse, and its __fsqrt_rn is the native
// approximation; division and sqrtf are correctly rounded by default), so the
// CPU build computes the same bits from the same COM. Massless cells get -1:
// always accepted, never listed.
__device__ __forceinline__ float open_radius2(float4 cm, uint32_t code, int l,
                                              const float* __restrict__ geom, int cubic,
                                              float theta) {
#pragma clang fp contract(off)
  if (!(cm.w > 0.f)) return -1.f;
  const float E = geom[6 + kMaxLevel + 1];
  const uint32_t q = code >> (30 - 3 * l);
  const float cmv[3] = {cm.x, cm.y, cm.z};
  float d2 = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float ext = cubic ? E : fmaxf(geom[3 + a] - geom[a], 1e-6f);
    const float h = ldexpf(ext, -l);
    const uint32_t idx = compact_bits(q >> (2 - a));
    const float ctr = geom[a] + ((float)idx + 0.5f) * h;
    const float d = cmv[a] - ctr;
    d2 = a == 0 ? d * d : d2 + d * d;
  }
  const float rc = ldexpf(E, -l) / theta + sqrtf(d2);
  return rc * rc;
}

__global__ __launch_bounds__(256) void k_tree_com(const int4* __restrict__ link, int n,
                                                  int64_t max_nodes,
                                                  const int* __restrict__ level_off,
                                                  const double4* __restrict__ prefix,
                                                  const float4* __restrict__ pms,
                                                  const uint32_t* __restrict__ codes,
                                                  const float* __restrict__ geom, int cubic,
                                                  float theta, float4* __restrict__ com,
                                                  float* __restrict__ r2) {
  const int total = (int)min((int64_t)level_off[kMaxLevel + 1], max_nodes);
  for (int k = blockIdx.x * 256 + threadIdx.x; k < total; k += gridDim.x * 256) {
    const int4 nd = link[k];
    const int cnt = link_count(nd.w);
    if (nd.z < 0 || cnt < 1 || nd.z + cnt > n) {
      com[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      r2[k] = -1.f;
      continue;
    }
    float4 cm;
    if (cnt == 1) {
      cm = pms[nd.z];
    } else {
      const double4 a = prefix[nd.z], b = prefix[nd.z + cnt];
      const double m = b.x - a.x;
      const double inv = m > 0.0 ? 1.0 / m : 0.0;
      cm = make_float4((float)((b.y - a.y) * inv), (float)((b.z - a.z) * inv),
                       (float)((b.w - a.w) * inv), (float)m);
    }
    com[k] = cm;
    r2[k] = open_radius2(cm, codes[nd.z], link_level(nd.w), geom, cubic, theta);
  }
}

// ---- 8. traversal ---------------------------------------------------------------------------
// kModified: open with the default traversal's radius (r2[i] < d^2) instead of
// the reference's test; used only to count what that test needs per body.
template <bool kCount, bool kModified = false>
__global__ __launch_bounds__(256) void k_traverse(const float4* __restrict__ pms,
                                                  const uint32_t* __restrict__ perm, int n,
                                                  const float4* __restrict__ com,
                                                  const int4* __restrict__ link,
                                                  const int* __restrict__ level_off,
                                                  int64_t max_nodes,
                                                  const float* __restrict__ geom, float theta2,
                                                  float eps2, double G,
                                                  double* __restrict__ force,
                                                  unsigned long long* __restrict__ inter,
                                                  int* __restrict__ flags,
                                                  const float* __restrict__ r2 = nullptr) {
  __shared__ float side2[kMaxLevel + 1];
  if (threadIdx.x <= kMaxLevel) side2[threadIdx.x] = geom[6 + threadIdx.x];
  __syncthreads();
  const int t = blockIdx.x * 256 + threadIdx.x;
  // A rope walk visits each node at most once; the bound and the index check
  // turn a corrupt tree into an error flag instead of a hang or a fault.
  const int total = (int)min((int64_t)level_off[kMaxLevel + 1], max_nodes);
  unsigned long long cnt = 0;
  if (t < n) {
    const float4 me = pms[t];
    float ax = 0.f, ay = 0.f, az = 0.f;
    int i = 0;
    for (int it = 0; i >= 0; ++it) {
      if (i >= total || it > total) {
        atomicOr(flags, 4);
        break;
      }
      const float4 c = com[i];
      const int4 nd = link[i];
      const float dx = c.x - me.x, dy = c.y - me.y, dz = c.z - me.z;
      const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      if (kModified ? r2[i] < d2 : side2[link_level(nd.w)] < theta2 * d2) {
        if (c.w > 0.f && d2 > 0.f) {
          const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
          const float s = (c.w * inv) * (inv * inv);
          ax = fmaf(dx, s, ax);
          ay = fmaf(dy, s, ay);
          az = fmaf(dz, s, az);
          if (kCount) ++cnt;
        }
        i = nd.y;
      } else if (nd.x < 0) {
        const int e = min(nd.z + link_count(nd.w), n);
        for (int b = max(nd.z, 0); b < e; ++b) {
          const float4 q = pms[b];
          const float bx = q.x - me.x, by = q.y - me.y, bz = q.z - me.z;
          const float b2 = fmaf(bz, bz, fmaf(by, by, bx * bx));
          if (b2 > 0.f) {
            const float inv = __builtin_amdgcn_rsqf(b2 + eps2);
            const float s = (q.w * inv) * (inv * inv);
            ax = fmaf(bx, s, ax);
            ay = fmaf(by, s, ay);
            az = fmaf(bz, s, az);
            if (kCount) ++cnt;
          }
        }
        i = nd.y;
      } else {
        i = nd.x;
      }
    }
    const uint32_t o = perm[t];
    if (!kModified && o < (uint32_t)n) {
      force[o] = G * (double)ax;
      force[n + o] = G * (double)ay;
      force[2 * n + o] = G * (double)az;
    }
  }
  if (kCount) {
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(inter, red[0]);
  }
}

// Wave traversal (the default). A wave takes one range of at most
// NB_BH_WAVE_BODIES Morton-consecutive bodies of one cell (emitted by the build)
// and its lanes share every decision: a cell is accepted for the whole range
// when its opening radius (side / theta + |COM - cell centre|) is less than
// dmin, the distance from its COM to the range's bounding box; otherwise it is
// opened, and a leaf that is not accepted is evaluated body by body by every
// lane. Per body this accepts a subset of what the per-body test accepts (at
// least as accurate), and nothing diverges. For latency, each wave keeps a
// stack of opened nodes in LDS as (first child, level of the children << 27 |
// end of the body range). One iteration pops up to 8 of them and lets 8 lanes
// per node examine its (at most 8) children at once, so one memory round trip
// covers 64 nodes. The walk only appends indices to an interaction list in LDS
// (accepted cells, flagged, and the bodies of unaccepted leaves) and pushes
// opened children; no source data is read while walking. Whenever the list
// fills up, and at the end, the wave gathers the sources 64 at a time into LDS
// and streams them through the direct-sum inner loop (LDS broadcast reads, 13
// VALU per interaction). From NB_BH_PAIR_MIN_BODIES bodies on (or with
// NB_F_BH_PAIRS) the build pairs sibling ranges of equal copy width: such a
// pair walks once for both ranges and leaves to each range only the cells the
// two decide differently (walk() below), which on the 4M two-disc system takes
// 30% fewer walk iterations and stages 34% fewer sources for exactly the same
// interactions.
constexpr int kListWaves = 4;
// LDS (26.7 KB per block) and 80 VGPRs allow six blocks per CU. A 384-entry
// list (instead of 512) is what makes the sixth fit; it measured about 1% faster
// on the two-disc 4M system.
constexpr int kListBlocksPerCU = 6;
// The deepest stack seen on the two-disc 4M system (400 sampled ranges, leaf 8)
// is 73 entries. A pair's deferred entries share the region from its top: a CPU
// replay of the pair walk on 4M-body two-disc, Plummer and uniform-ball systems
// (leaf 1) needs at most 245 stack plus deferred entries. Overflow is detected
// and reported (NB_E_TREE_OVERFLOW).
constexpr int kStackCap = 512;
constexpr int kListCap = 384;                 // source indices per wave
constexpr uint32_t kCellBit = 0x80000000u;    // list entry refers to a cell, not a body

// Stream the list through the direct-sum inner loop: 64 entries at a time,
// each lane gathers one source (cell COM or body) into the LDS staging area,
// then the lanes read the staged sources by broadcast. The bodies of a range are
// held S = 64 / P times (P = its body count rounded up to a power of two, at
// least 4; S = 1 above 32 bodies): the lanes of slice s = lane / P take the
// staged entries s, s + S, s + 2S, ..., so every lane does useful work, and the
// slices' partial sums are added at the end of the range. Entries are staged
// slice-major (entry j at (j % S) * (64 / S) + j / S), so a slice reads its
// entries from consecutive addresses: four reads per step share one base
// address with immediate offsets, and the trip count is wave-uniform.
// Compensated accumulation: each flush sums its (at most kListCap / 2 per lane
// at two slices) terms in plain FP32 and folds the partial sums into the
// running (sum, error) pairs with TwoSum, as the direct sum folds every 256
// sources.
template <bool kCount, bool kGuard>
__device__ __forceinline__ void list_flush(const uint32_t* list, int lc, float4* stage, int lane,
                                           int slice, int log2S, const float4* __restrict__ pms,
                                           const float4* __restrict__ com, float4 me, float eps2_s,
                                           float& sx, float& sy, float& sz, float& cx, float& cy,
                                           float& cz, unsigned long long& cnt) {
  float ax = 0.f, ay = 0.f, az = 0.f;
  float eps2 = eps2_s;  // a VGPR copy: all-VGPR operand forms issue faster
  asm volatile("" : "+v"(eps2));
  const int S = 1 << log2S;
  const int per = 64 >> log2S;  // entries per slice
  const int put = (lane & (S - 1)) * per + (lane >> log2S);  // slice-major slot of entry `lane`
  const float4* mine = stage + slice * per;
  // lane's source of the chunk starting at e (a zero-mass pad past the list)
  auto gather = [&](int e) {
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    if (e + lane < lc) {
      const uint32_t id = list[e + lane];
      g = (id & kCellBit) ? com[id & ~kCellBit] : pms[id];
    }
    return g;
  };
  float4 q = gather(0);
  for (int e = 0; e < lc; e += 64) {
    const int m = min(64, lc - e);
    stage[put] = q;
    __builtin_amdgcn_wave_barrier();
    // the next chunk's gather is in flight while this chunk is evaluated
    if (e + 64 < lc) q = gather(e + 64);
    // T = ceil(m / S) rounded up to 4 entries per slice; S <= 16 divides 64, so
    // 64 / S is a multiple of 4 and T <= 64 / S: every read stays inside this
    // slice's block; entries at or past m are zero-mass pads.
    const int T = __builtin_amdgcn_readfirstlane((((m + S - 1) >> log2S) + 3) & ~3);
    for (int t = 0; t < T; t += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 src = mine[t + u];  // few distinct addresses: LDS broadcast
        const float dx = src.x - me.x, dy = src.y - me.y, dz = src.z - me.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        float s = (src.w * inv) * (inv * inv);
        if (kGuard) s = r2 > 0.f ? s : 0.f;
        ax = fmaf(dx, s, ax);
        ay = fmaf(dy, s, ay);
        az = fmaf(dz, s, az);
        if (kCount) cnt += (src.w > 0.f && fmaf(dz, dz, fmaf(dy, dy, dx * dx)) > 0.f) ? 1 : 0;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  nbcore::two_sum_fold(sx, cx, ax);
  nbcore::two_sum_fold(sy, cy, ay);
  nbcore::two_sum_fold(sz, cz, az);
}

// Shared LDS of one traversing wave. The stack region holds the walk's stack
// from the bottom and a pair's deferred entries from the top.
struct WaveLds {
  int2* stack;
  uint32_t* list;
  float4* stage;
  const int* lvl;      // first node of each level (kMaxLevel + 2 entries)
  int cap;             // entries of the stack region in use (<= kStackCap)
};

// An axis-aligned box held in scalar registers (wave-uniform).
struct Box {
  float lx, ly, lz, hx, hy, hz;
};

// Squared distance from c to the box (0 inside).
__device__ __forceinline__ float box_dist2(const Box& b, const float4& c) {
  const float ex = fmaxf(fmaxf(b.lx - c.x, c.x - b.hx), 0.f);
  const float ey = fmaxf(fmaxf(b.ly - c.y, c.y - b.hy), 0.f);
  const float ez = fmaxf(fmaxf(b.lz - c.z, c.z - b.hz), 0.f);
  return fmaf(ez, ez, fmaf(ey, ey, ex * ex));
}

// Walk the tree for the bodies held by the wave's lanes and accumulate their
// accelerations (lanes of slice s take staged entries s, s + S, ...).
//   kShared: the shared walk of a pair of ranges with boxes A and B. A cell
//     both accept is evaluated for both ranges; a cell exactly one accepts is
//     deferred: its (first sibling, 8-bit sibling mask) entry goes to the top of
//     the stack region and nd counts them. Cells neither accepts are opened (or,
//     as leaves, evaluated body by body) for both.
//   otherwise: the walk of one range with box U, from the root when nd == 0,
//     else from the nd deferred entries (read, not consumed: the pair's other
//     range walks them too).
// For each range this examines and evaluates exactly the cells its own walk
// from the root would: the shared walk only descends where both ranges open,
// and a deferred cell is decided by each range's own test.
template <bool kCount, bool kGuard, bool kShared>
__device__ void walk(const WaveLds& L, int lane, int slice, int log2S, float4 me, int n,
                     const float4* __restrict__ pms, const float4* __restrict__ com,
                     const int4* __restrict__ link, const float* __restrict__ r2, float eps2,
                     int* __restrict__ flags, const Box& U, const Box& A, const Box& B, int& nd,
                     float& ax, float& ay, float& az, float& cx, float& cy, float& cz,
                     unsigned long long& cnt) {
  const uint64_t lt = (uint64_t(1) << lane) - 1;
  int2* stack = L.stack;
  uint32_t* list = L.list;
#define NB_FLUSH()                                                                              \
  do {                                                                                          \
    list_flush<kCount, kGuard>(list, lc, L.stage, lane, slice, log2S, pms, com, me, eps2, ax, ay, \
                               az, cx, cy, cz, cnt);                                           \
    lc = 0;                                                                                     \
  } while (0)
  int sp = 0, lc = 0, dpos = 0;
  if (kShared || nd == 0) {
    // a virtual parent whose only child is the root
    if (lane == 0) stack[0] = make_int2(0, (int)((1u << 27) | (uint32_t)n));
    sp = 1;
    __builtin_amdgcn_wave_barrier();
  }
  const int q = lane >> 3, ci = lane & 7;
  for (int iter = 0; sp > 0 || (!kShared && dpos < nd); ++iter) {
    if (iter > L.lvl[kMaxLevel + 1]) {  // every node is examined at most once
      if (lane == 0) atomicOr(flags, 4);
      break;
    }
    // room for this iteration's accepted cells (at most 64), flushed while
    // nothing of the iteration is live
    if (lc + 64 > kListCap) NB_FLUSH();
    int child = -1, efirst = 0;
    int4 cl = make_int4(-1, -1, 0, 0);
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    float rr = 0.f;
    if (!kShared && sp < 8 && dpos < nd) {
      // deferred entries: up to 32 of them, at most 64 cells, listed in the
      // staging area (free while walking) so that lane j takes the j-th cell
      int2 d = make_int2(0, 0);
      const bool have = lane < 32 && dpos + lane < nd;
      if (have) d = stack[L.cap - 1 - (dpos + lane)];
      const int pc = __popc((uint32_t)d.y);
      const int incl = wave_incl_scan(pc);
      const uint64_t fit = __ballot(have && incl <= 64);
      const int kd = __popcll(fit);  // >= 1: one entry has at most 8 cells
      int* cells = reinterpret_cast<int*>(L.stage);
      if ((fit >> lane) & 1) {
        int o = incl - pc;
        for (uint32_t mm = (uint32_t)d.y; mm; mm &= mm - 1) cells[o++] = d.x + __ffs(mm) - 1;
      }
      __builtin_amdgcn_wave_barrier();
      const int T = __builtin_amdgcn_readlane(incl, kd - 1);
      if (lane < T) {
        child = cells[lane];
        cl = link[child];
        c = com[child];
        rr = r2[child];
      }
      __builtin_amdgcn_wave_barrier();
      dpos += kd;
    } else {
      const int k = min(sp, 8);
      sp -= k;
      if (q < k) {
        // entry: first child, and (slots << 27 | parent's body end); the slots
        // stop at the end of the children's level, so every load below is in
        // bounds and the three are issued together (a child is one whose bodies
        // start before the parent's end)
        const int2 e = stack[sp + q];
        efirst = e.x;
        const uint32_t w = (uint32_t)e.y;
        const int slots = (int)(w >> 27);
        const int pend = (int)(w & ((1u << 27) - 1));
        if (ci < slots) {
          const int cand = e.x + ci;
          cl = link[cand];
          c = com[cand];
          rr = r2[cand];
          if (cl.z < pend) child = cand;
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    bool accept = false, defer = false, leaf = false, open = false;
    if (child >= 0) {
      // the group's modified test (see open_radius2); a pair shares what both accept
      if (kShared) {
        const bool a = rr < box_dist2(A, c), b = rr < box_dist2(B, c);
        accept = a && b;
        defer = a != b;
      } else {
        accept = rr < box_dist2(U, c);
      }
      leaf = !accept && !defer && cl.x < 0;
      open = !accept && !defer && cl.x >= 0;
    }
    // accepted cells with mass -> list
    const bool acc = accept && c.w > 0.f;
    const uint64_t ba = __ballot(acc);
    if (acc) list[lc + __popcll(ba & lt)] = (uint32_t)child | kCellBit;
    lc += __popcll(ba);
    // deferred cells -> one entry per stack entry that has any
    int nd_add = 0;
    bool head = false;
    uint32_t dmask = 0;
    uint64_t bg = 0;
    if (kShared) {
      const uint64_t bd = __ballot(defer);
      if (bd) {
        const uint32_t half = q < 4 ? (uint32_t)bd : (uint32_t)(bd >> 32);
        dmask = (half >> (lane & 24)) & 0xffu;  // bits 8q .. 8q + 7
        head = ci == 0 && dmask != 0;
        bg = __ballot(head);
        nd_add = __popcll(bg);
      }
    }
    // opened cells -> stack
    const uint64_t bo = __ballot(open);
    if (sp + __popcll(bo) + nd + nd_add > L.cap) {
      if (lane == 0) atomicOr(flags, 8);
      break;
    }
    if (kShared && head) stack[L.cap - 1 - (nd + __popcll(bg & lt))] = make_int2(efirst, (int)dmask);
    nd += nd_add;
    if (open) {
      const uint32_t pend = (uint32_t)(cl.z + link_count(cl.w));
      const int slots = max(0, min(8, L.lvl[link_level(cl.w) + 2] - cl.x));
      stack[sp + __popcll(bo & lt)] = make_int2(cl.x, (int)(((uint32_t)slots << 27) | pend));
    }
    sp += __popcll(bo);
    // leaves that are not accepted -> the indices of their bodies
    uint64_t bl = __ballot(leaf);
    while (bl) {
      const int src = __ffsll((unsigned long long)bl) - 1;
      bl &= bl - 1;
      const int ls = __builtin_amdgcn_readlane(cl.z, src);
      const int lcount = __builtin_amdgcn_readlane(link_count(cl.w), src);
      for (int off = 0; off < lcount; off += 64) {
        const int m = min(64, lcount - off);
        if (lc + 64 > kListCap) NB_FLUSH();
        if (lane < m) list[lc + lane] = (uint32_t)min(ls + off + lane, n - 1);
        lc += m;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  NB_FLUSH();
#undef NB_FLUSH
}

// The bounding box of the bodies held by the lanes, reduced over every lane bit
// except `keep` (0: over all lanes; else each half of the lanes split by that
// bit gets its own box), then read from lanes `from0` and `from1`.
__device__ __forceinline__ void lane_boxes(float4 me, int keep, int from0, int from1, Box& b0,
                                           Box& b1) {
  float lx = me.x, ly = me.y, lz = me.z, hx = me.x, hy = me.y, hz = me.z;
  // lanes lane ^ off: ds_swizzle's bit mode inside each half of the wave, one
  // ds_bpermute across the halves
  auto step = [&](auto xor_lanes) {
    lx = fminf(lx, xor_lanes(lx));
    ly = fminf(ly, xor_lanes(ly));
    lz = fminf(lz, xor_lanes(lz));
    hx = fmaxf(hx, xor_lanes(hx));
    hy = fmaxf(hy, xor_lanes(hy));
    hz = fmaxf(hz, xor_lanes(hz));
  };
  auto swz = [](auto off_c) {
    return [](float v) {
      constexpr int kPattern = (decltype(off_c)::value << 10) | 0x1f;  // xor, or 0, and 0x1f
      return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), kPattern));
    };
  };
  if (keep != 32) step([](float v) { return __shfl_xor(v, 32); });
  if (keep != 16) step(swz(std::integral_constant<int, 16>{}));
  if (keep != 8) step(swz(std::integral_constant<int, 8>{}));
  if (keep != 4) step(swz(std::integral_constant<int, 4>{}));
  if (keep != 2) step(swz(std::integral_constant<int, 2>{}));
  if (keep != 1) step(swz(std::integral_constant<int, 1>{}));
  // uniform over each group: keep the boxes in scalar registers
  auto at = [](float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  };
  b0 = Box{at(lx, from0), at(ly, from0), at(lz, from0), at(hx, from0), at(hy, from0), at(hz, from0)};
  b1 = Box{at(lx, from1), at(ly, from1), at(lz, from1), at(hx, from1), at(hy, from1), at(hz, from1)};
}

// Persistent: every wave takes records from a ticket until none is left, so the
// cost of uneven ranges balances dynamically. A record is one wave range or a
// pair of sibling ranges (x, y) and (z, w), w > 0. A pair first walks once for
// both (its lanes: [slice][range][slot], P1 lanes per copy of a range), then,
// if cells were deferred, each range walks them with its own lanes ([slice][slot]);
// each body's force is the sum of the two parts in FP64, in that order. The
// counting variant (diagnostics only) needs more registers and is allowed five
// blocks per CU; its surplus blocks start when others retire and take what
// tickets are left.
template <bool kCount, bool kGuard>
__global__ __launch_bounds__(64 * kListWaves, kCount ? 5 : kListBlocksPerCU) void k_traverse_list(
    const float4* __restrict__ pms, const uint32_t* __restrict__ perm, int n,
    const float4* __restrict__ com, const int4* __restrict__ link,
    const int* __restrict__ level_off, int64_t max_nodes, const float* __restrict__ r2,
    float eps2, double G, double* __restrict__ force,
    unsigned long long* __restrict__ inter, int* __restrict__ flags,
    const int4* __restrict__ waves, int* __restrict__ wctl, const int* __restrict__ order,
    float* __restrict__ cost, int stack_cap) {
  __shared__ int2 s_stack[kListWaves][kStackCap];
  __shared__ uint32_t s_list[kListWaves][kListCap];
  __shared__ float4 s_stage[kListWaves][64];
  __shared__ int s_lvl[kMaxLevel + 2];
  if (threadIdx.x < kMaxLevel + 2)
    s_lvl[threadIdx.x] = (int)min((int64_t)level_off[threadIdx.x], max_nodes);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const WaveLds L{s_stack[wave], s_list[wave], s_stage[wave], s_lvl,
                  stack_cap > 0 && stack_cap < kStackCap ? stack_cap : kStackCap};
  const int nrec = min(wctl[0], n);
  unsigned long long cnt = 0;
  for (;;) {
    int w = 0;
    if (lane == 0) w = atomicAdd(&wctl[1], 1);
    w = __builtin_amdgcn_readfirstlane(w);
    if (w >= nrec) break;
    const uint64_t t_start = wall_clock64();
    const int ri = order ? order[w] : w;
    const int4 rec = waves[(unsigned)ri < (unsigned)nrec ? ri : w];  // wave-uniform: scalar registers
    const bool pair = rec.w > 0;
    // shared walk: a single range is held 64 / P1 times, a pair 32 / P1 times
    const int P1 = copy_lanes(pair ? max(rec.y, rec.w) : rec.y);
    const int grp = pair ? 2 * P1 : P1;
    const int slot = lane & (P1 - 1);
    const bool second = pair && (lane & P1);
    const int rs = second ? rec.z : rec.x, rn = second ? rec.w : rec.y;
    const bool valid = slot < rn;
    const int t = min(rs + (valid ? slot : 0), n - 1);  // lanes past a range copy its first body
    const float4 me = pms[t];
    const int log2S = 6 - (31 - __clz(grp)), slice = lane / grp;
    Box A, B;
    lane_boxes(me, pair ? P1 : 0, 0, pair ? P1 : 0, A, B);
    float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    unsigned long long c = 0;
    int nd = 0;
    if (pair)
      walk<kCount, kGuard, true>(L, lane, slice, log2S, me, n, pms, com, link, r2, eps2, flags, A,
                                 A, B, nd, ax, ay, az, cx, cy, cz, c);
    else
      walk<kCount, kGuard, false>(L, lane, slice, log2S, me, n, pms, com, link, r2, eps2, flags,
                                  A, A, B, nd, ax, ay, az, cx, cy, cz, c);
    // each slice's compensated sum in FP64, then the slices added (butterfly
    // over the slice bits, fixed order)
    double fx = (double)ax + (double)cx, fy = (double)ay + (double)cy,
           fz = (double)az + (double)cz;
    for (int off = grp; off < 64; off <<= 1) {
      fx += __shfl_xor(fx, off);
      fy += __shfl_xor(fy, off);
      fz += __shfl_xor(fz, off);
    }
    if (valid) cnt += c;
    if (nd == 0) {
      if (valid && lane < grp) {
        const uint32_t o = perm[t];
        if (o < (uint32_t)n) {
          force[o] = G * fx;
          force[n + o] = G * fy;
          force[2 * n + o] = G * fz;
        }
      }
    }
    // the deferred cells, range by range
    for (int h = 0; h < (nd > 0 ? 2 : 0); ++h) {
      const int hs = h ? rec.z : rec.x, hn = h ? rec.w : rec.y;
      const int P = copy_lanes(hn);
      const int s2 = lane & (P - 1);
      const bool v2 = s2 < hn;
      const int t2 = min(hs + (v2 ? s2 : 0), n - 1);
      const float4 me2 = pms[t2];
      ax = ay = az = cx = cy = cz = 0.f;
      c = 0;
      walk<kCount, kGuard, false>(L, lane, lane / P, 6 - (31 - __clz(P)), me2, n, pms, com, link,
                                  r2, eps2, flags, h ? B : A, A, B, nd, ax, ay, az, cx, cy, cz, c);
      double gx = (double)ax + (double)cx, gy = (double)ay + (double)cy,
             gz = (double)az + (double)cz;
      for (int off = P; off < 64; off <<= 1) {
        gx += __shfl_xor(gx, off);
        gy += __shfl_xor(gy, off);
        gz += __shfl_xor(gz, off);
      }
      // the shared part of body s2 of range h sits in lane h * P1 + s2
      const double px = __shfl(fx, h * P1 + s2), py = __shfl(fy, h * P1 + s2),
                   pz = __shfl(fz, h * P1 + s2);
      if (v2) cnt += c;
      if (v2 && lane < P) {
        const uint32_t o = perm[t2];
        if (o < (uint32_t)n) {
          force[o] = G * (px + gx);
          force[n + o] = G * (py + gy);
          force[2 * n + o] = G * (pz + gz);
        }
      }
    }
    // the record's device time, shared out over its bodies for the next pass's order
    const float share = (float)(wall_clock64() - t_start) / (float)(rec.y + rec.w);
    if (cost && valid && lane < grp) {
      const uint32_t o = perm[t];
      if (o < (uint32_t)n) cost[o] = share;
    }
  }
  if (kCount) {
    __shared__ unsigned long long red[64 * kListWaves];
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 64 * kListWaves / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(inter, red[0]);
  }
}

// ---- 8. traversal order ---------------------------------------------------------------
// A persistent launch ends when its last wave finishes its last record, so the
// records are taken costliest first: each record's estimate is the device time
// its bodies' records took in the previous pass (stored per body, by original
// index, so it survives the rebuild), bucketed by half-octaves. Within a bucket
// the order is that of the atomics. The order changes only which wave takes a
// record and when; every record's arithmetic, and so every force, is the same.

// Half-octave of an estimate; anything not a finite value >= 1 (no history yet,
// or memory not written by a pass) goes last.
__device__ __forceinline__ int cost_bucket(float est) {
  if (!(est >= 1.f) || !(est <= FLT_MAX)) return 0;
  const uint32_t u = __float_as_uint(est);
  return min(kCostBuckets - 1, 2 * ((int)(u >> 23) - 127) + (int)((u >> 22) & 1));
}

// Records are handled in chunks of kOrderChunk by a grid of at most kOrderBlocks
// blocks; each block counts in LDS and makes one global atomic per bucket per
// chunk, so the few busy buckets see few atomics.
constexpr int kOrderChunk = 1024;
constexpr int kOrderBlocks = 256;
// Below this the three ordering launches cost more than the tail they save
// (Plummer steps with single ranges, ordered against not: 0.577 against 0.583
// ms at 65,536 bodies, 0.858 against 0.881 at 262,144, 1.267 against 1.296 at
// 524,288).
constexpr int kOrderMinBodies = 1 << 16;

__global__ __launch_bounds__(256) void k_rec_bucket(const int4* __restrict__ waves,
                                                    const int* __restrict__ wctl, int n,
                                                    const float* __restrict__ cost,
                                                    const uint32_t* __restrict__ perm,
                                                    int* __restrict__ bucket,
                                                    int* __restrict__ hist) {
  __shared__ int h[kCostBuckets];
  if (threadIdx.x < kCostBuckets) h[threadIdx.x] = 0;
  __syncthreads();
  const int nrec = min(wctl[0], n);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < nrec; i += gridDim.x * 256) {
    const int4 r = waves[i];  // starts are sorted body indices < n
    float est = cost[perm[r.x]] * (float)r.y;
    if (r.w > 0) est += cost[perm[r.z]] * (float)r.w;
    const int b = cost_bucket(est);
    bucket[i] = b;
    atomicAdd(&h[b], 1);
  }
  __syncthreads();
  if (threadIdx.x < kCostBuckets && h[threadIdx.x] > 0) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// One wave: cursor[b] = records in buckets above b (costliest first); the
// counts are cleared for the next pass's k_rec_bucket.
__global__ __launch_bounds__(64) void k_rec_offsets(int* __restrict__ hist) {
  const int lane = threadIdx.x;
  const int v = hist[kCostBuckets - 1 - lane];
  hist[kCostBuckets + kCostBuckets - 1 - lane] = wave_incl_scan(v) - v;
  hist[kCostBuckets - 1 - lane] = 0;
}

__global__ __launch_bounds__(256) void k_rec_scatter(const int* __restrict__ wctl, int n,
                                                     const int* __restrict__ bucket,
                                                     int* __restrict__ hist,
                                                     int* __restrict__ order) {
  __shared__ int h[kCostBuckets], base[kCostBuckets];
  constexpr int kPer = kOrderChunk / 256;
  const int nrec = min(wctl[0], n);
  for (int c0 = blockIdx.x * kOrderChunk; c0 < nrec; c0 += gridDim.x * kOrderChunk) {
    if (threadIdx.x < kCostBuckets) h[threadIdx.x] = 0;
    __syncthreads();
    int bk[kPer], pos[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = c0 + k * 256 + threadIdx.x;
      bk[k] = i < nrec ? bucket[i] : -1;
      pos[k] = bk[k] >= 0 ? atomicAdd(&h[bk[k]], 1) : 0;
    }
    __syncthreads();
    if (threadIdx.x < kCostBuckets && h[threadIdx.x] > 0)
      base[threadIdx.x] = atomicAdd(&hist[kCostBuckets + threadIdx.x], h[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int at = bk[k] >= 0 ? base[bk[k]] + pos[k] : nrec;
      if (at < nrec) order[at] = c0 + k * 256 + threadIdx.x;
    }
    __syncthreads();
  }
}

__global__ void k_zero_u64(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

int level_bound(int n, int l) {
  int64_t w = 1;
  for (int k = 0; k < l && w < n; ++k) w *= 8;
  return (int)(w < n ? w : n);
}

// Enqueue an exclusive int scan: out[0..count) from in, count constant or per level.
hipError_t iscan(const int* in, int* out, int* part, const int* level_off, int l, int n_const,
                 int bound, int* level_off_mut, hipStream_t s) {
  const int blocks = (bound + kScanTile - 1) / kScanTile;
  if (blocks < 1) return hipSuccess;
  if (bound <= kSingleScan) {
    hipLaunchKernelGGL(k_iscan_single, dim3(1), dim3(1024), 0, s, in, level_off, l, n_const, bound,
                       out, level_off_mut);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_iscan_reduce, dim3(blocks), dim3(kScanThreads), 0, s, in, level_off, l,
                     n_const, bound, part);
  hipLaunchKernelGGL(k_iscan_final, dim3(blocks), dim3(kScanThreads), 0, s, in, level_off, l,
                     n_const, bound, part, blocks, level_off_mut, out);
  return hipGetLastError();
}

}  // namespace

void nb_bh_carve(nb_ctx* c, char* base, int upto) {
  nb_bh_state& b = c->bh;
  const nb_bh_layout L = nb_bh_layout_for(c->n);
  auto at = [&](int k) -> void* { return k < upto ? base + L.off[k] : nullptr; };
  b.bbox_part = (float*)at(kBhBbox);
  b.geom = (float*)at(kBhGeom);
  b.keys[0] = (uint32_t*)at(kBhKey0);
  b.keys[1] = (uint32_t*)at(kBhKey1);
  b.vals[0] = (uint32_t*)at(kBhVal0);
  b.vals[1] = (uint32_t*)at(kBhVal1);
  b.codes = (uint32_t*)at(kBhCodes);
  b.digit_hist = (int*)at(kBhHist);
  b.iscan_part = (int*)at(kBhIscan);
  b.pms = (float4*)at(kBhPms);
  b.prefix = (double4*)at(kBhPrefix);
  b.scan_part = (double4*)at(kBhScanPart);
  b.node_com = (float4*)at(kBhCom);
  b.node_r2 = (float*)at(kBhR2);
  b.node_link = (int4*)at(kBhLink);
  b.level_off = (int*)at(kBhLevel);
  b.child_cnt = (int*)at(kBhChildCnt);
  b.child_off = (int*)at(kBhChildOff);
  b.child_bounds = (int*)at(kBhChildBounds);
  b.counters = (int*)at(kBhCounters);
  b.top_start = (int*)at(kBhTopStart);
  b.waves = (int4*)at(kBhWaves);
  b.inter = (unsigned long long*)at(kBhInter);
  b.rec_cost = (float*)at(kBhRecCost);
  b.rec_bucket = (int*)at(kBhRecBucket);
  b.rec_order = (int*)at(kBhRecOrder);
  b.rec_hist = (int*)at(kBhRecHist);
  b.bbox_blocks = kBboxBlocks;
  b.sort_blocks = (int)L.sort_blocks;
  b.max_nodes = upto > kBhPms ? L.max_nodes : 0;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b.persist_blocks = kListBlocksPerCU * cus;  // resident traversal blocks
}

int nb_sort_alloc(nb_ctx* c) {
  NB_TRY
  nb_bh_state& b = c->bh;
  if (b.keys[0]) return NB_OK;
  const nb_bh_layout L = nb_bh_layout_for(c->n);
  if (hipMalloc(&b.block, (size_t)L.off[kBhPms]) != hipSuccess) {
    b.block = nullptr;
    return nb_fail(NB_E_OOM, "device allocation failed (Morton sort)");
  }
  nb_bh_carve(c, (char*)b.block, kBhPms);
  if (hipMemset(b.geom, 0, sizeof(float) * 32) != hipSuccess)
    return nb_fail(NB_E_HIP, "hipMemset failed (Morton sort)");
  return NB_OK;
  NB_CATCH("nb_sort_alloc")
}

int nb_bh_alloc(nb_ctx* c) {
  NB_TRY
  nb_bh_state& b = c->bh;
  const nb_bh_layout L = nb_bh_layout_for(c->n);
  if (hipMalloc(&b.block, (size_t)L.off[kBhBufCount]) != hipSuccess) {
    b.block = nullptr;
    return nb_fail(NB_E_OOM, "nb_create: device allocation failed (octree)");
  }
  nb_bh_carve(c, (char*)b.block, kBhBufCount);
  // the batched leaf loop may read up to kSpare bodies past the end: zero mass
  if (hipMemset(b.pms + c->n, 0, sizeof(float4) * kSpare) != hipSuccess ||
      hipMemset(b.geom, 0, sizeof(float) * 32) != hipSuccess ||
      hipMemset(b.rec_cost, 0, sizeof(float) * c->n) != hipSuccess)
    return nb_fail(NB_E_HIP, "nb_create: hipMemset failed (octree)");
  return NB_OK;
  NB_CATCH("nb_bh_alloc")
}

void nb_bh_free(nb_ctx* c) {
  nb_bh_state& b = c->bh;
  if (b.block) (void)hipFree(b.block);
  b = nb_bh_state{};
}

hipError_t nb_morton_sort(nb_ctx* c) {
  nb_bh_state& b = c->bh;
  const int n = c->n;
  hipStream_t s = c->stream;
  const int g256 = (n + 255) / 256;
  // 1. bbox
  hipLaunchKernelGGL(k_bbox_partial, dim3(b.bbox_blocks), dim3(256), 0, s, c->pm, n, b.bbox_part);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, b.bbox_part, b.bbox_blocks, b.geom);
  // 2. encode
  const int cubic = (c->cfg.flags & NB_F_BH_AXIS_BOX) ? 0 : 1;
  hipLaunchKernelGGL(k_morton, dim3(g256), dim3(256), 0, s, c->pm, n, cubic, b.geom, b.codes, b.keys[0],
                     b.vals[0]);
  // 3. sort: 4 passes of 8 bits, ping-pong 0 -> 1 -> 0 -> 1 -> 0
  const int nb = b.sort_blocks;
  const int hist = 256 * nb;
  int* offs = b.digit_hist + hist;
  for (int pass = 0; pass < 4; ++pass) {
    const int in = pass & 1, out = in ^ 1;
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, b.keys[in], n, 8 * pass,
                       nb, b.digit_hist);
    hipError_t e = iscan(b.digit_hist, offs, b.iscan_part, nullptr, 0, hist, hist, nullptr, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, b.keys[in],
                       b.vals[in], n, 8 * pass, nb, offs, b.keys[out], b.vals[out]);
  }
  b.sorted_buf = 0;
  return hipGetLastError();
}

hipError_t nb_launch_permute_state(nb_ctx* c) {
  nb_bh_state& b = c->bh;
  const int n = c->n;
  hipStream_t s = c->stream;
  const int g256 = (n + 255) / 256;
  const uint32_t* perm = b.vals[0];
  double* arrays[4] = {c->pos, c->vel, c->acc, c->force};
  for (double* a : arrays) {
    hipLaunchKernelGGL(k_permute_d3, dim3(g256), dim3(256), 0, s, a, perm, n, c->tmp);
    hipError_t e = hipMemcpyAsync(a, c->tmp, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  float* mtmp = reinterpret_cast<float*>(c->tmp);
  hipLaunchKernelGGL(k_permute_f32, dim3(g256), dim3(256), 0, s, c->mass, perm, n, mtmp);
  return hipMemcpyAsync(c->mass, mtmp, sizeof(float) * n, hipMemcpyDeviceToDevice, s);
}

hipError_t nb_bh_build(nb_ctx* c) {
  nb_bh_state& b = c->bh;
  const int n = c->n;
  hipStream_t s = c->stream;
  const int g256 = (n + 255) / 256;
  hipError_t es = nb_morton_sort(c);
  if (es != hipSuccess) return es;
  const uint32_t* codes_sorted = b.keys[0];
  const uint32_t* perm = b.vals[0];
  // 4. permute
  hipLaunchKernelGGL(k_gather, dim3(g256), dim3(256), 0, s, c->pm, perm, n, b.pms);
  // 5. FP64 prefix of mass moments
  const int mblocks = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_mscan_reduce, dim3(mblocks), dim3(kScanThreads), 0, s, b.pms, n, b.scan_part);
  hipLaunchKernelGGL(k_mscan_partials, dim3(1), dim3(kScanThreads), 0, s, b.scan_part, mblocks);
  hipLaunchKernelGGL(k_mscan_final, dim3(mblocks), dim3(kScanThreads), 0, s, b.pms, n, b.scan_part,
                     b.prefix);
  // 6. octree: the top levels from per-cell searches, then level by level
  const int leaf = c->cfg.leaf_size;
  hipLaunchKernelGGL(k_top_bounds, dim3((kTopEntries + 255) / 256), dim3(256), 0, s, codes_sorted,
                     n, b.top_start);
  const int pairs = nb_bh_pairs(n, c->cfg.flags) ? 1 : 0;
  hipLaunchKernelGGL(k_top_build, dim3(1), dim3(1024), 0, s, b.top_start, n, leaf, b.node_link,
                     b.level_off, b.waves, b.counters, b.rec_hist, pairs);
  for (int l = kTopLevels; l < kMaxLevel; ++l) {
    const int bound = level_bound(n, l);
    const int grid = std::max(1, std::min((bound + 255) / 256, 4096));
    hipLaunchKernelGGL(k_tree_count, dim3(grid), dim3(256), 0, s, codes_sorted, n, b.node_link,
                       b.level_off, l, leaf, bound, b.child_cnt, b.child_bounds);
    hipError_t e = iscan(b.child_cnt, b.child_off, b.iscan_part, b.level_off, l, 0, bound,
                         b.level_off, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tree_emit, dim3(grid), dim3(256), 0, s, codes_sorted, n, b.node_link,
                       b.level_off, l, leaf, bound, b.child_off, b.child_bounds, b.max_nodes,
                       c->flags, b.waves, b.counters, pairs);
  }
  // 7. centre of mass
  const int cgrid = (int)std::max<int64_t>(1, std::min<int64_t>((b.max_nodes + 255) / 256, 8192));
  const int cubic = (c->cfg.flags & NB_F_BH_AXIS_BOX) ? 0 : 1;
  hipLaunchKernelGGL(k_tree_com, dim3(cgrid), dim3(256), 0, s, b.node_link, n, b.max_nodes,
                     b.level_off, b.prefix, b.pms, codes_sorted, b.geom, cubic, c->cfg.theta,
                     b.node_com, b.node_r2);
  return hipGetLastError();
}

hipError_t nb_bh_traverse(nb_ctx* c, bool count_interactions) {
  nb_bh_state& b = c->bh;
  const int n = c->n;
  const int g256 = (n + 255) / 256;
  const bool group = !(c->cfg.flags & NB_F_BH_PER_BODY);
  if (count_interactions)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, b.inter);
#define NB_TRAVERSE_ARGS                                                                       \
  b.pms, b.vals[0], n, b.node_com, b.node_link, b.level_off, b.max_nodes, b.geom, c->theta2, \
      c->cfg.eps2, c->G, c->force, b.inter, c->flags
  const bool guard = c->guard;
  const int gl = b.persist_blocks;  // persistent blocks pulling wave ranges
  // work costliest first from kOrderMinBodies on (and wherever pairs are
  // forced); smaller systems take their records in order and keep no times
  const bool ordered = group && (n >= kOrderMinBodies || (c->cfg.flags & NB_F_BH_PAIRS) != 0);
  if (ordered) {
    const int og = std::max(1, std::min((n + kOrderChunk - 1) / kOrderChunk, kOrderBlocks));
    hipLaunchKernelGGL(k_rec_bucket, dim3(og), dim3(256), 0, c->stream, b.waves, b.counters, n,
                       b.rec_cost, b.vals[0], b.rec_bucket, b.rec_hist);
    hipLaunchKernelGGL(k_rec_offsets, dim3(1), dim3(64), 0, c->stream, b.rec_hist);
    hipLaunchKernelGGL(k_rec_scatter, dim3(og), dim3(256), 0, c->stream, b.counters, n,
                       b.rec_bucket, b.rec_hist, b.rec_order);
  }
#define NB_LIST_ARGS                                                                     \
  b.pms, b.vals[0], n, b.node_com, b.node_link, b.level_off, b.max_nodes, b.node_r2,     \
      c->cfg.eps2, c->G, c->force, b.inter, c->flags, b.waves, b.counters,                \
      ordered ? b.rec_order : nullptr, ordered ? b.rec_cost : nullptr, c->walk_stack
  if (group && count_interactions && guard)
    hipLaunchKernelGGL((k_traverse_list<true, true>), dim3(gl), dim3(64 * kListWaves), 0, c->stream,
                       NB_LIST_ARGS);
  else if (group && count_interactions)
    hipLaunchKernelGGL((k_traverse_list<true, false>), dim3(gl), dim3(64 * kListWaves), 0, c->stream,
                       NB_LIST_ARGS);
  else if (group && guard)
    hipLaunchKernelGGL((k_traverse_list<false, true>), dim3(gl), dim3(64 * kListWaves), 0, c->stream,
                       NB_LIST_ARGS);
  else if (group)
    hipLaunchKernelGGL((k_traverse_list<false, false>), dim3(gl), dim3(64 * kListWaves), 0, c->stream,
                       NB_LIST_ARGS);
#undef NB_LIST_ARGS
  else if (count_interactions)
    hipLaunchKernelGGL((k_traverse<true>), dim3(g256), dim3(256), 0, c->stream, NB_TRAVERSE_ARGS);
  else
    hipLaunchKernelGGL((k_traverse<false>), dim3(g256), dim3(256), 0, c->stream, NB_TRAVERSE_ARGS);
#undef NB_TRAVERSE_ARGS
  return hipGetLastError();
}

hipError_t nb_bh_count_per_body(nb_ctx* c) {
  nb_bh_state& b = c->bh;
  const int n = c->n;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, b.inter);
  hipLaunchKernelGGL((k_traverse<true, true>), dim3((n + 255) / 256), dim3(256), 0, c->stream, b.pms,
                     b.vals[0], n, b.node_com, b.node_link, b.level_off, b.max_nodes, b.geom,
                     c->theta2, c->cfg.eps2, c->G, c->force, b.inter, c->flags, b.node_r2);
  return hipGetLastError();
}

// ---- debug taps ---------------------------------------------------------------------------------
extern "C" int nb_debug_set_walk_stack(nb_ctx* c, int32_t entries) {
  NB_TRY
  if (!c) return nb_fail(NB_E_ARG, "nb_debug_set_walk_stack: null context");
  if (entries < 0 || entries > kStackCap)
    return nb_fail(NB_E_ARG, "nb_debug_set_walk_stack: entries must be 0..512");
  // a captured step holds the old value as a kernel argument
  nb_drop_graph(c->stream, c->graph_exec, c->graph);
  c->walk_stack = entries;
  return NB_OK;
  NB_CATCH("nb_debug_set_walk_stack")
}

extern "C" int nb_debug_morton(nb_ctx* c, uint32_t* codes, uint32_t* perm) {
  NB_TRY
  if (!c) return nb_fail(NB_E_ARG, "nb_debug_morton: null context");
  if (c->cfg.solver != NB_SOLVER_BH) return nb_fail(NB_E_ARG, "nb_debug_morton: not a BH context");
  if (!c->tree_valid) return nb_fail(NB_E_ARG, "nb_debug_morton: no BH force pass has run");
  NB_HIP(hipStreamSynchronize(c->stream));
  if (codes) NB_HIP(hipMemcpy(codes, c->bh.codes, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost));
  if (perm) NB_HIP(hipMemcpy(perm, c->bh.vals[0], sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost));
  return NB_OK;
  NB_CATCH("nb_debug_morton")
}

extern "C" int nb_debug_tree(nb_ctx* c, nb_node_view* nodes, int64_t cap, int64_t* node_count,
                             float* bbox) {
  NB_TRY
  if (!c) return nb_fail(NB_E_ARG, "nb_debug_tree: null context");
  if (c->cfg.solver != NB_SOLVER_BH) return nb_fail(NB_E_ARG, "nb_debug_tree: not a BH context");
  if (!c->tree_valid) return nb_fail(NB_E_ARG, "nb_debug_tree: no BH force pass has run");
  NB_HIP(hipStreamSynchronize(c->stream));
  int off[kMaxLevel + 2];
  NB_HIP(hipMemcpy(off, c->bh.level_off, sizeof(off), hipMemcpyDeviceToHost));
  const int64_t total = off[kMaxLevel + 1];
  if (node_count) *node_count = total;
  if (bbox) NB_HIP(hipMemcpy(bbox, c->bh.geom, sizeof(float) * 6, hipMemcpyDeviceToHost));
  const int64_t m = std::min<int64_t>(cap, total);
  if (nodes && m > 0) {
    std::vector<int4> link(m);
    std::vector<float4> com(m);
    std::vector<float> r2(m);
    NB_HIP(hipMemcpy(link.data(), c->bh.node_link, sizeof(int4) * m, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(com.data(), c->bh.node_com, sizeof(float4) * m, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpy(r2.data(), c->bh.node_r2, sizeof(float) * m, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < m; ++k) {
      nb_node_view& v = nodes[k];
      v.com[0] = com[k].x;
      v.com[1] = com[k].y;
      v.com[2] = com[k].z;
      v.mass = com[k].w;
      v.first_child = link[k].x;
      v.next = link[k].y;
      v.body_start = link[k].z;
      v.body_count = link[k].w >> 4;
      v.level = link[k].w & 15;
      v.open_r2 = r2[k];
    }
  }
  return NB_OK;
  NB_CATCH("nb_debug_tree")
}
