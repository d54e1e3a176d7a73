/*
 * oracle.h — CPU restatement of the reference's N-body hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine
 * and the timed CPU baseline in bench.py; no product code links, loads or
 * calls it. It is compiled twice from oracle.c:
 *   liboracle.so      -O2, IEEE arithmetic (-ffp-contract=off): the checker;
 *   liboracle_fast.so -O3 -ffast-math -fopenmp: the v2/v5 speed baseline, built
 *                     with the reference's flags (CMakeLists.txt:49,83-91) except
 *                     -march=x86-64-v3 in place of -march=native (portable .so).
 *
 * Every function cites the reference code it restates. Layouts follow the
 * reference: pos4/vel4/acc4/force4 are float[4n] (w = 0), mass4 is float[4n]
 * with the mass in .x (ParticleSystem.cpp:18). FP64 paths use double[3n].
 */
#ifndef NB_ORACLE_H
#define NB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FP64 direct sum (the "FP64 oracle"):
 * acc[t] = G sum_{j != i_t} m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^{3/2}, i_t = targets[t]
 * (targets == NULL: all bodies, nt = n). Closed form of
 * ParticleSolverCPUSequential.cpp:26-45 evaluated in double. OpenMP over targets;
 * each target's sum runs over j in ascending order, so the result is independent
 * of the thread count. */
void or_direct_f64(int64_t n, const double* pos3, const double* mass, double eps2, double G,
                   int64_t nt, const int64_t* targets, double* acc3, int nthreads);

/* v1 / v2 force pass, FP32, same operation order as ParticleSolverCPUSequential.cpp:26-45
 * (glm vec4 subtract, length2 = (x*x+y*y)+(z*z+w*w), powf(l2+eps2,1.5f), (G*m)/d, v*s).
 * nthreads = 1 is v1; nthreads > 1 is v2's `omp parallel for schedule(static)`
 * (ParticleSolverCPUParallel.cpp:14,19). Targets [t0, t1) only (bounded samples). */
void or_pp_forces(int64_t n, const float* pos4, const float* mass4, float eps2, float G,
                  float* force4, int64_t t0, int64_t t1, int nthreads);

/* Lagged kick-drift-kick, FP32 (ParticleSystem.h:34-51), bodies [0, n). */
void or_kdk_update(int64_t n, float* pos4, float* vel4, float* acc4, const float* force4,
                   float dt, int nthreads);
/* Semi-implicit Euler, FP32 (updateParticles.glsl:44-55). */
void or_euler_update(int64_t n, float* pos4, float* vel4, const float* force4, float dt);

/* One v1/v2 step: forces for all bodies, then the lagged KDK update
 * (ParticleSolverCPUSequential.cpp:15-24; ParticleSolverCPUParallel.cpp:12-22). */
void or_pp_step(int64_t n, float* pos4, float* vel4, float* acc4, const float* mass4,
                float* force4, float eps2, float G, float dt, int nthreads);

/* FP64 lagged KDK on double state (the engine's integrator, checked against the
 * FP32 restatement above): same update order as ParticleSystem.h:34-51. */
void or_kdk_update_f64(int64_t n, double* pos3, double* vel3, double* acc3,
                       const double* force3, double dt);
void or_euler_update_f64(int64_t n, double* pos3, double* vel3, double* acc3,
                         const double* force3, double dt);

/* FP64 energy / momentum: ekin = sum m v^2 / 2, epot = -G sum_{i<j} m_i m_j / sqrt(r^2+eps2). */
void or_energy_f64(int64_t n, const double* pos3, const double* vel3, const double* mass,
                   double eps2, double G, double* ekin, double* epot, double p[3], int nthreads);

/* ---- v5 Barnes-Hut restatement (Octree.cpp:9-350), int32 node links ---- */
/* Builds the sequential pointer octree, propagates, prunes, then traverses for
 * every body. theta is unsquared (the reference stores 0.5*0.5). Returns 0, or -5
 * if the node pool (pool_factor * n nodes; the reference uses n<200 ? 200n : 8n,
 * pass pool_factor <= 0 for that) overflows. *node_count receives the nodes used.
 * Deviation: the bounding-box max starts at -FLT_MAX (the reference's
 * numeric_limits<float>::min() is the smallest positive float, Octree.cpp:27). */
int or_v5_forces(int64_t n, const float* pos4, const float* mass4, float eps2, float G,
                 float theta, float* force4, double pool_factor, int64_t* node_count,
                 int nthreads);

/* ---- v6 parallel Barnes-Hut restatement (ParallelOctreeCPU.cpp:7-505,
 * ParticleSolverBHutCPUParallel.cpp:18-41), the timed CPU baseline ----
 * Top tree of depth <= 3 (at most 512 tasks), bodies bucketed into the deepest
 * top cells under one omp_lock per cell, one sequential subtree per task built
 * with `schedule(dynamic)`, merge rule d^2 < 1e-3 or 10 extra levels, top-level
 * COM, prune, then the force loop with `schedule(dynamic)` for targets [t0, t1).
 * Deviations: int32 node links (the reference's float links corrupt trees
 * above ~65.6k bodies, D3) and subtree pools sized by need (16 nodes per body
 * of the task plus 64) instead of n/2 nodes per task (48 GiB at n = 2^22).
 * Returns 0 or -5 (pool overflow); build_s/force_s receive the phase times. */
int or_v6_forces(int64_t n, const float* pos4, const float* mass4, float eps2, float G,
                 float theta, float* force4, int64_t t0, int64_t t1, int nthreads,
                 double* build_s, double* force_s);

/* ---- CPU build of the engine's Morton / sort / octree (topology reference) ---- */
/* Exact min/max of the FP32 positions; max starts at -FLT_MAX. */
void or_bbox(int64_t n, const float* pos4, float bmin[3], float bmax[3]);
/* 30-bit Morton code per morton.glsl:37-75: n = clamp((p-min)/ext,0,1);
 * q = uint(clamp(n*1024, 0, 1023)); code = expand(qx)*4 + expand(qy)*2 + expand(qz).
 * cubic = 0: ext = max(max-min, 1e-6) per axis (morton.glsl's inputs);
 * cubic = 1: one ext = the largest of those on every axis (the engine's default). */
uint32_t or_expand_bits(uint32_t v);
void or_morton(int64_t n, const float* pos4, const float bmin[3], const float bmax[3], int cubic,
               uint32_t* codes);
/* Stable sort of (code, index): perm[k] = index of the k-th smallest (code, index). */
void or_sort_codes(int64_t n, const uint32_t* codes, uint32_t* perm, uint32_t* codes_sorted);

typedef struct or_tree {
  int64_t count;       /* nodes used */
  int32_t* first_child;
  int32_t* next;
  int32_t* start;
  int32_t* nbody;
  int32_t* level;
  float* com4;         /* (cx, cy, cz, m) */
} or_tree;
/* Breadth-first octree over sorted codes: a node with more than leaf_size bodies
 * and level < 10 splits into its non-empty Morton-digit children (ascending digit).
 * Returns the node count; arrays must hold max_nodes entries (or_tree_bound). */
int64_t or_tree_bound(int64_t n);
int64_t or_build_tree(int64_t n, const uint32_t* codes_sorted, int leaf_size, int64_t max_nodes,
                      int32_t* first_child, int32_t* next, int32_t* start, int32_t* nbody,
                      int32_t* level);
/* Mass and centre of mass per node from FP64 prefix sums over the sorted bodies;
 * single-body nodes copy the body exactly. pos4s/mass are in sorted order. */
void or_tree_com(int64_t nodes, const int32_t* start, const int32_t* nbody, int64_t n,
                 const float* pos4s, float* com4);
/* Stackless rope traversal (Octree.cpp:303-338 contract): accept a cell when
 * s_l^2 < theta^2 d^2 (use its COM), otherwise descend, and evaluate a leaf that is
 * not accepted body by body; pairs with d^2 == 0 are skipped. s_l = root side / 2^l
 * with root side = max over axes of max(max-min, 1e-6). Writes force4 in ORIGINAL
 * order via perm. Returns the number of interactions evaluated. */
int64_t or_bh_forces(int64_t n, const float* pos4s, const uint32_t* perm, const float bmin[3],
                     const float bmax[3], int64_t nodes, const int32_t* first_child,
                     const int32_t* next, const int32_t* start, const int32_t* nbody,
                     const int32_t* level, const float* com4, float eps2, float G, float theta,
                     float* force4, int nthreads);
/* Wave ranges of the engine's traversal: for every node holding more than cap
 * bodies, its children are packed in order into ranges of at most cap bodies; a
 * child holding more packs its own children at its level if it is internal, and
 * is cut into chunks of cap if it is a leaf. A system of at most cap bodies is
 * one range. Every body lies in exactly one range. Returns the number of ranges. */
int64_t or_bh_wave_ranges(int64_t n, int64_t nodes, const int32_t* first_child,
                          const int32_t* start, const int32_t* nbody, const int32_t* level,
                          int cap, int32_t* range_start, int32_t* range_count);
/* Inverse of or_expand_bits on 10-bit values (bits 0, 3, 6, ... gathered). */
uint32_t or_compact_bits(uint32_t v);
/* Squared opening radius of every node for the engine's default traversal:
 * r = s_l / theta + delta, with s_l = root side / 2^l (the cell side of
 * Octree.cpp:311-320) and delta = |COM - geometric centre of the cell| (the
 * offset term of Barnes' modified criterion, J. Barnes, "A modified tree code:
 * don't laugh; it runs", J. Comput. Phys. 87 (1990); also the default of his
 * treecode 1.4). The cell of a node is the 3l-bit prefix of its first body's
 * code; its centre is bmin + (index + 1/2) * extent / 2^l per axis (extent: the
 * cubic box's largest side, or each axis's own as in or_morton). Massless nodes
 * get -1 (never opened, never listed). Every operation is one IEEE FP32 op in
 * this order, which the GPU restates with explicit round-to-nearest intrinsics. */
void or_tree_open_r2(int64_t nodes, const int32_t* start, const int32_t* level, const float* com4,
                     const uint32_t* codes_sorted, const float bmin[3], const float bmax[3],
                     int cubic, float theta, float* r2);
/* Group traversal (the engine's default on the GPU): the bodies of one wave
 * range share every decision: a cell is accepted for the group when
 * r^2 < dmin^2, r the node's opening radius (or_tree_open_r2) and dmin the
 * distance from the cell's COM to the range's bounding box, else the group
 * descends; a leaf that is not accepted is evaluated body by body for every
 * member. Since r > s_l / theta and dmin <= d for every member, each body
 * accepts a subset of the cells the reference's per-body test s_l^2 < theta^2 d^2
 * (or_bh_forces) accepts on the same tree. group is the range capacity
 * (NB_BH_WAVE_BODIES = 32 on the GPU; 1 gives a per-body traversal with the
 * modified test). codes_sorted and cubic locate each node's cell. */
int64_t or_bh_forces_group(int64_t n, const float* pos4s, const uint32_t* perm,
                           const float bmin[3], const float bmax[3], int64_t nodes,
                           const int32_t* first_child, const int32_t* next, const int32_t* start,
                           const int32_t* nbody, const int32_t* level, const float* com4,
                           const uint32_t* codes_sorted, int cubic, float eps2, float G,
                           float theta, int group, float* force4, int nthreads);
/* The whole engine BH force pass on the CPU: bbox, Morton, sort, tree, COM,
 * traversal. Node arrays are optional outputs (NULL to skip). */
int64_t or_bh_pipeline(int64_t n, const float* pos4, float eps2, float G, float theta,
                       int leaf_size, float* force4, uint32_t* codes, uint32_t* perm,
                       int64_t* node_count, int nthreads);
/* As or_bh_pipeline with the group traversal (group = 0: per-body traversal) and
 * the choice of Morton box (cubic, see or_morton); or_bh_pipeline is group 0,
 * cubic 0. */
int64_t or_bh_pipeline_group(int64_t n, const float* pos4, float eps2, float G, float theta,
                             int leaf_size, int group, int cubic, float* force4, uint32_t* codes,
                             uint32_t* perm, int64_t* node_count, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
