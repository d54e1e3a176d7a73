// nb_force_core.h — the direct-sum inner loop shared by the single-system
// kernel (nb_direct.hip) and the batched-ensemble kernel (nb_batch.hip).
//
// A lane owns B target bodies held in VGPRs; the sources [j0, j1) are read at
// wave-uniform addresses, so they come through the scalar cache into SGPRs and
// feed the VALU as scalar operands. One interaction is
//   3 v_sub + 3 v_fma (r^2) + v_rsq + 3 v_mul + 3 v_fma.
// Sums are plain FP32 within a chunk of kChunk sources and folded into a
// (sum, error) pair with Knuth's TwoSum once per chunk.
#pragma once
#include <hip/hip_runtime.h>

namespace nbcore {

constexpr int kChunk = 256;

__device__ __forceinline__ void two_sum_fold(float& s, float& c, float x) {
  const float t = s + x;
  const float bp = t - s;
  const float e = (s - (t - bp)) + (x - bp);
  s = t;
  c += e;
}

template <int B>
struct Targets {
  float x[B], y[B], z[B];
  float sx[B], sy[B], sz[B];  // compensated sums
  float cx[B], cy[B], cz[B];  // their error terms
};

// Accumulate the sources [j0, j1) (wave-uniform, j1 - j0 a multiple of 16).
template <int B, bool kGuard>
__device__ __forceinline__ void accumulate(const float4* __restrict__ pm, int j0, int j1,
                                           float eps2_s, Targets<B>& t) {
  // eps2 is uniform and would stay an SGPR operand; a VGPR copy (the empty asm
  // hides its uniformity) turns the r^2 seed v_fma into an all-VGPR form.
  float eps2 = eps2_s;
  asm volatile("" : "+v"(eps2));
  for (int jc = j0; jc < j1; jc += kChunk) {
    const int je = min(jc + kChunk, j1);
    float px[B], py[B], pz[B];
#pragma unroll
    for (int b = 0; b < B; ++b) px[b] = py[b] = pz[b] = 0.f;
    auto step = [&](const float4* qq) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 q = qq[u];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const float dx = q.x - t.x[b];
          const float dy = q.y - t.y[b];
          const float dz = q.z - t.z[b];
          const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
          const float inv = __builtin_amdgcn_rsqf(r2);
          float s = (q.w * inv) * (inv * inv);
          if (kGuard) s = r2 > 0.f ? s : 0.f;
          px[b] = fmaf(dx, s, px[b]);
          py[b] = fmaf(dy, s, py[b]);
          pz[b] = fmaf(dz, s, pz[b]);
        }
      }
    };
    // Two register sets of four sources: the scalar loads of one set are
    // issued before the other set's arithmetic, so the load latency overlaps
    // the math instead of stalling every four sources (measured 10% faster at
    // N = 2^18 than loading each set right before its use). The memory clobber
    // keeps the compiler from sinking the loads next to their first use.
    // je - jc is a multiple of 8 (ranges are multiples of 16 sources).
    float4 qa[4], qb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) qa[u] = pm[jc + u];
    for (int j = jc; j < je; j += 8) {
#pragma unroll
      for (int u = 0; u < 4; ++u) qb[u] = pm[j + 4 + u];  // wave-uniform: scalar loads
      asm volatile("" ::: "memory");
      step(qa);
      const int jn = j + 8 < je ? j + 8 : j;  // the last step reloads its own (in range)
#pragma unroll
      for (int u = 0; u < 4; ++u) qa[u] = pm[jn + u];
      asm volatile("" ::: "memory");
      step(qb);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      two_sum_fold(t.sx[b], t.cx[b], px[b]);
      two_sum_fold(t.sy[b], t.cy[b], py[b]);
      two_sum_fold(t.sz[b], t.cz[b], pz[b]);
    }
  }
}

template <int B>
__device__ __forceinline__ void load_targets(const float4* __restrict__ pm, int ibase,
                                             Targets<B>& t) {
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const float4 p = pm[ibase + 64 * b];
    t.x[b] = p.x;
    t.y[b] = p.y;
    t.z[b] = p.z;
    t.sx[b] = t.sy[b] = t.sz[b] = 0.f;
    t.cx[b] = t.cy[b] = t.cz[b] = 0.f;
  }
}

// Merge the K wave partials of one i-block in wave order 0..K-1 through LDS.
// All K waves must call it; afterwards wave 0 holds the merged sums.
template <int B, int K>
__device__ __forceinline__ void merge_waves(float (*part)[6][B][64], int wave, int lane,
                                            Targets<B>& t) {
#pragma unroll
  for (int b = 0; b < B; ++b) {
    part[wave][0][b][lane] = t.sx[b];
    part[wave][1][b][lane] = t.sy[b];
    part[wave][2][b][lane] = t.sz[b];
    part[wave][3][b][lane] = t.cx[b];
    part[wave][4][b][lane] = t.cy[b];
    part[wave][5][b][lane] = t.cz[b];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int b = 0; b < B; ++b) {
      t.sx[b] = part[0][0][b][lane];
      t.sy[b] = part[0][1][b][lane];
      t.sz[b] = part[0][2][b][lane];
      t.cx[b] = part[0][3][b][lane];
      t.cy[b] = part[0][4][b][lane];
      t.cz[b] = part[0][5][b][lane];
#pragma unroll
      for (int w = 1; w < K; ++w) {
        two_sum_fold(t.sx[b], t.cx[b], part[w][0][b][lane]);
        two_sum_fold(t.sy[b], t.cy[b], part[w][1][b][lane]);
        two_sum_fold(t.sz[b], t.cz[b], part[w][2][b][lane]);
        t.cx[b] += part[w][3][b][lane];
        t.cy[b] += part[w][4][b][lane];
        t.cz[b] += part[w][5][b][lane];
      }
    }
  }
}

}  // namespace nbcore
