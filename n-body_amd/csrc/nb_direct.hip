// nb_direct.hip — O(N^2) direct-sum force pass and the FP64 integrator.
//
// Replaces the reference's tiled GLSL force shader (forceCalculationOptimized.glsl:28-76,
// dispatched by ParticleSolverGPU.cpp:31-35) and the update shader
// (updateParticles.glsl:34-58, ParticleSolverGPU.cpp:37-40).
//
// Direct-sum design (gfx950, wave64, vector ALU):
//   * Each lane owns B target bodies (i), held in VGPRs. A workgroup of K waves
//     shares one i-block of 64*B bodies and splits the source range (j) into K
//     contiguous slices, one per wave.
//   * Source bodies are read at wave-uniform addresses, so they arrive through
//     the scalar data cache into SGPRs (s_load_dwordx16 = 4 bodies) and feed the
//     VALU directly as scalar operands: no LDS staging, no barriers in the j-loop.
//   * One interaction = 3 v_sub + 3 v_fma (r^2) + v_rsq + 3 v_mul + 3 v_fma.
//     The self pair needs no branch: dx = 0 contributes exactly 0 when eps2 > 0.
//   * Compensated accumulation: plain FP32 partial sums over a chunk of
//     256 sources, folded into a (sum, error) pair with Knuth's TwoSum once
//     per chunk, so compensation costs < 1% of the ALU work. The K wave partials
//     are merged in LDS in fixed wave order and widened to FP64; G is applied in
//     FP64. Results are deterministic (no atomics, fixed order).
#include "nb_force_core.h"
#include "nb_internal.h"

namespace {

template <int B, int K, bool kGuard>
__global__ __launch_bounds__(64 * K) void k_direct(const float4* __restrict__ pm, int n,
                                                   int n_pad, float eps2, double G,
                                                   double* __restrict__ force) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ibase = blockIdx.x * (64 * B) + lane;
  nbcore::Targets<B> t;
  nbcore::load_targets<B>(pm, ibase, t);
  const int jlen = n_pad / K;  // multiple of 64 (n_pad is a multiple of 256)
  nbcore::accumulate<B, kGuard>(pm, wave * jlen, wave * jlen + jlen, eps2, t);
  if constexpr (K > 1) {
    __shared__ float part[K][6][B][64];
    nbcore::merge_waves<B, K>(part, wave, lane, t);
    if (wave != 0) return;
  }
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int i = ibase + 64 * b;
    if (i < n) {
      force[i] = G * ((double)t.sx[b] + (double)t.cx[b]);
      force[n + i] = G * ((double)t.sy[b] + (double)t.cy[b]);
      force[2 * n + i] = G * ((double)t.sz[b] + (double)t.cz[b]);
    }
  }
}

template <int B, int K>
hipError_t launch_bk(const nb_ctx* c) {
  const int blocks = c->n_pad / (64 * B);
  if (c->guard)
    hipLaunchKernelGGL((k_direct<B, K, true>), dim3(blocks), dim3(64 * K), 0, c->stream, c->pm,
                       c->n, c->n_pad, c->cfg.eps2, c->G, c->force);
  else
    hipLaunchKernelGGL((k_direct<B, K, false>), dim3(blocks), dim3(64 * K), 0, c->stream, c->pm,
                       c->n, c->n_pad, c->cfg.eps2, c->G, c->force);
  return hipGetLastError();
}

// Shapes: B in {1, 2, 4}; K in {1, 2, 4}, plus K = 8 for B <= 2 and K = 16 for
// B = 1 (small systems need the extra waves; the LDS merge holds K*6*B*64 floats).
template <int B>
hipError_t launch_b(const nb_ctx* c) {
  switch (c->jsplit) {
    case 1: return launch_bk<B, 1>(c);
    case 2: return launch_bk<B, 2>(c);
    case 4: return launch_bk<B, 4>(c);
    case 8:
      if constexpr (B <= 2) return launch_bk<B, 8>(c);
      return hipErrorInvalidValue;
    case 16:
      if constexpr (B == 1) return launch_bk<B, 16>(c);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

// ---- FP64 integration -----------------------------------------------------
// Lagged kick-drift-kick (ParticleSystem.h:34-51):
//   v += a*dt/2; x += v*dt; a = F (evaluated at the pre-drift x); v += a*dt/2
// and semi-implicit Euler (updateParticles.glsl:44-55): v += F*dt; x += v*dt.
// Also refreshes the FP32 force view pm = (x, y, z, m) and flags non-finite state.
template <bool kEuler>
__global__ __launch_bounds__(256) void k_integrate(int n, double dt, double* __restrict__ pos,
                                                   double* __restrict__ vel,
                                                   double* __restrict__ acc,
                                                   const double* __restrict__ force,
                                                   const float* __restrict__ mass,
                                                   float4* __restrict__ pm, int* __restrict__ flags) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double h = dt * 0.5;
  double x[3];
  bool finite = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int k = a * n + i;
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
    finite = finite && isfinite(p) && isfinite(v);
  }
  pm[i] = make_float4((float)x[0], (float)x[1], (float)x[2], mass[i]);
  if (!finite) atomicOr(flags, 1);
}

__global__ __launch_bounds__(256) void k_refresh_pm(int n, int n_pad, const double* __restrict__ pos,
                                                    const float* __restrict__ mass,
                                                    float4* __restrict__ pm) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_pad) return;
  if (i < n)
    pm[i] = make_float4((float)pos[i], (float)pos[n + i], (float)pos[2 * n + i], mass[i]);
  else
    pm[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Render view: (x, y, z, |v|) per body, the inputs of the reference's point
// shaders (vertexShader.glsl:3-9 reads positions and velocities;
// fragmentShader.glsl colours by length(velocity)).
__global__ __launch_bounds__(256) void k_render_view(int n, const double* __restrict__ pos,
                                                     const double* __restrict__ vel,
                                                     float4* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
  out[i] = make_float4((float)pos[i], (float)pos[n + i], (float)pos[2 * n + i],
                       (float)sqrt(vx * vx + vy * vy + vz * vz));
}

__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ p, int v, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = v;
}

// ---- FP64 diagnostics -------------------------------------------------------
// Block partials of (ekin, px, py, pz) and of sum_i m_i phi_i, where
// phi_i = sum_{j != i} m_j / sqrt(r_ij^2 + eps2) in FP64. Partials are summed on
// the host in block order, so the result is deterministic.
__global__ __launch_bounds__(256) void k_energy_kin(int n, const double* __restrict__ vel,
                                                    const float* __restrict__ mass,
                                                    double* __restrict__ part) {
  __shared__ double red[4][256];
  double e = 0, px = 0, py = 0, pz = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double m = mass[i];
    const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
    e += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    px += m * vx;
    py += m * vy;
    pz += m * vz;
  }
  red[0][threadIdx.x] = e;
  red[1][threadIdx.x] = px;
  red[2][threadIdx.x] = py;
  red[3][threadIdx.x] = pz;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_energy_pot(int n, const double* __restrict__ pos,
                                                    const float* __restrict__ mass, double eps2,
                                                    double* __restrict__ part) {
  __shared__ double red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const double xi = pos[ii], yi = pos[n + ii], zi = pos[2 * n + ii];
  double phi = 0.0;
  for (int j = 0; j < n; ++j) {  // wave-uniform j
    const double dx = pos[j] - xi, dy = pos[n + j] - yi, dz = pos[2 * n + j] - zi;
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    const double t = (j != ii) ? (double)mass[j] * rsqrt(r2) : 0.0;
    phi += t;
  }
  red[threadIdx.x] = i < n ? (double)mass[i] * phi : 0.0;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// FP64 direct-sum acceleration of selected bodies: the error reference for the
// FP32 solvers, computed on the device from the same FP32 (x, y, z, m) view the
// force passes read, widened to FP64. One workgroup per target; thread t adds
// the sources t, t + 256, ... in order, and the 256 partials are added in a
// fixed tree, so the result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_ref_f64(const float4* __restrict__ pm, int n,
                                                 const int* __restrict__ slots, double eps2,
                                                 double G, double* __restrict__ acc3) {
  __shared__ double red[3][256];
  const int i = slots[blockIdx.x];
  const float4 me = pm[i];
  const double xi = me.x, yi = me.y, zi = me.z;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) {
    const float4 q = pm[j];
    const double dx = (double)q.x - xi, dy = (double)q.y - yi, dz = (double)q.z - zi;
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    const double s = j == i ? 0.0 : (double)q.w / (r2 * sqrt(r2));
    ax += s * dx;
    ay += s * dy;
    az += s * dz;
  }
  red[0][threadIdx.x] = ax;
  red[1][threadIdx.x] = ay;
  red[2][threadIdx.x] = az;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) acc3[3 * blockIdx.x + threadIdx.x] = G * red[threadIdx.x][0];
}

}  // namespace

hipError_t nb_launch_ref_f64(nb_ctx* c, const int* slots, int k, double* acc3) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ref_f64, dim3(k), dim3(256), 0, c->stream, c->pm, c->n, slots,
                     (double)c->cfg.eps2, c->G, acc3);
  return hipGetLastError();
}

hipError_t nb_launch_direct(nb_ctx* c) {
  switch (c->bpl) {
    case 1: return launch_b<1>(c);
    case 2: return launch_b<2>(c);
    case 4: return launch_b<4>(c);
    default: return hipErrorInvalidValue;
  }
}

hipError_t nb_launch_integrate(nb_ctx* c) {
  const int blocks = (c->n + 255) / 256;
  if (c->cfg.integrator == NB_INTEG_EULER)
    hipLaunchKernelGGL((k_integrate<true>), dim3(blocks), dim3(256), 0, c->stream, c->n,
                       (double)c->cfg.dt, c->pos, c->vel, c->acc, c->force, c->mass, c->pm,
                       c->flags);
  else
    hipLaunchKernelGGL((k_integrate<false>), dim3(blocks), dim3(256), 0, c->stream, c->n,
                       (double)c->cfg.dt, c->pos, c->vel, c->acc, c->force, c->mass, c->pm,
                       c->flags);
  return hipGetLastError();
}

hipError_t nb_launch_refresh_pm(nb_ctx* c) {
  hipLaunchKernelGGL(k_refresh_pm, dim3(c->n_pad / 256), dim3(256), 0, c->stream, c->n, c->n_pad,
                     c->pos, c->mass, c->pm);
  return hipGetLastError();
}

hipError_t nb_launch_render_view(nb_ctx* c, float4* out) {
  hipLaunchKernelGGL(k_render_view, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, c->pos,
                     c->vel, out);
  return hipGetLastError();
}

hipError_t nb_launch_fill_i32(int* p, int v, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, p, v,
                     count);
  return hipGetLastError();
}

// ekin_p receives red_blocks * 4 partials, epot_p receives ceil(n/256) partials.
hipError_t nb_launch_energy(nb_ctx* c, double* ekin_p, double* epot_p) {
  hipLaunchKernelGGL(k_energy_kin, dim3(c->red_blocks), dim3(256), 0, c->stream, c->n, c->vel,
                     c->mass, ekin_p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_energy_pot, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n,
                     c->pos, c->mass, (double)c->cfg.eps2, epot_p);
  return hipGetLastError();
}
