/*
 * nbody.h — C ABI of the MI355X (gfx950) N-body force engine.
 *
 * One opaque context (nb_ctx) owns every device buffer of one particle
 * system: FP64 positions/velocities/accelerations, the FP32 (x,y,z,m)
 * force view, the octree and all scratch. State stays resident in HBM
 * between calls; host memory is only touched by nb_upload / nb_download*.
 *
 * Each entry point replaces a piece of the reference's GPU solver plumbing:
 *   ParticleSolver::updateParticlePositions  include/ParticleSolver.h:10
 *   ParticleSolverGPU (v4 tiled direct sum)   src/ParticleSolver/ParticleSolverGPU/ParticleSolverGPU.cpp:18-41
 *   ParticleSolverBHutGPU (v7 Barnes-Hut)     src/ParticleSolver/ParticleSolverBHutGPU/ParticleSolverBHutGPU.cpp:9-61
 *   ParticleSystem SSBO set-up (bindings 0-4) src/ParticleSystem/ParticleSystem.cpp:73-83
 *   OpenGLBuffer create/map/read-back        src/OpenGLBuffer/OpenGLBuffer.cpp:12-165
 *
 * Conventions (reference semantics):
 *   - "force" is the acceleration a_i = G sum_{j!=i} m_j r_ij / (|r_ij|^2 + eps2)^{3/2}
 *     with no m_i factor (ParticleSolverCPUSequential.cpp:26-45, ParticleSystem.h:43-46).
 *   - Host arrays use the reference's per-attribute vec4 layout: pos4/vel4/acc4
 *     are float[4*n] with w ignored on input and written as 0 on output;
 *     masses are a plain float[n] (the reference keeps them in vec4.x).
 *   - theta is given unsquared (0.5), as in Octree.cpp:15 and
 *     ParticleSolverBHutGPU.h:19; the reference's test s^2 < theta^2 d^2
 *     (Octree.cpp:320) is NB_F_BH_PER_BODY, the default adds the offset of the
 *     cell's mass from its centre (see NB_F_BH_PER_BODY below).
 *
 * Threading: a context is used by one host thread at a time; distinct
 * contexts may be used concurrently (the library keeps a context's one-time
 * step-graph capture from overlapping other threads' allocations and copies).
 * nb_step is asynchronous on the
 * context's stream; nb_download*, nb_forces, nb_energy and nb_sync wait.
 * Every function returns 0 on success or a negative NB_E_* code; the
 * calling thread's nb_last_error() then describes the failure.
 */
#ifndef NBODY_H
#define NBODY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define NB_OK 0
#define NB_E_ARG (-1)           /* bad argument / size / state */
#define NB_E_HIP (-2)           /* HIP runtime error */
#define NB_E_OOM (-3)           /* device (or host) allocation failed */
#define NB_E_BUDGET (-4)        /* request exceeds the HBM admission budget */
#define NB_E_TREE_OVERFLOW (-5) /* octree node pool exhausted */
#define NB_E_NONFINITE (-6)     /* a NaN/Inf appeared in the particle state */
#define NB_E_NOGPU (-7)         /* no HIP device visible */

/* ---- configuration ---------------------------------------------------- */
typedef enum nb_solver {
  NB_SOLVER_DIRECT = 0, /* O(N^2) tiled particle-particle (replaces v4) */
  NB_SOLVER_BH = 1      /* Barnes-Hut octree (replaces v7) */
} nb_solver;

typedef enum nb_integrator {
  /* CPU solvers' lagged kick-drift-kick: v+=a*dt/2; x+=v*dt; a=F(x_pre); v+=a*dt/2
   * (ParticleSystem.h:34-51). Default. */
  NB_INTEG_KDK_LAGGED = 0,
  /* GPU solvers' semi-implicit Euler: v+=F*dt; x+=v*dt (updateParticles.glsl:44-55). */
  NB_INTEG_EULER = 1
} nb_integrator;

typedef struct nb_config {
  int64_t n;          /* number of bodies: direct sum 1 <= n <= 715827882 ((2^31 - 1) / 3,
                         32-bit state indices); Barnes-Hut 1 <= n < 2^27 */
  float eps2;         /* squared Plummer softening (reference "-s") */
  float dt;           /* time step (reference "-t") */
  float G;            /* gravitational constant; 0 means 1.0 (reference G=1) */
  float theta;        /* BH opening angle, unsquared; 0 means 0.5 */
  int32_t solver;     /* nb_solver */
  int32_t integrator; /* nb_integrator */
  int32_t leaf_size;  /* BH bodies per leaf, 1..64; 0 = tuned default */
  int32_t bodies_per_lane; /* direct: i-bodies per lane (1,2,4); 0 = tuned default */
  int32_t j_split;    /* direct: waves sharing one i-block (1,2,4; 8 with bodies_per_lane
                         <= 2; 16 with 1); 0 = tuned default */
  int32_t flags;      /* NB_F_* */
} nb_config;

#define NB_F_NO_GRAPH 0x1 /* launch steps eagerly instead of replaying a captured hipGraph */
/* Barnes-Hut opening test. Default: the bodies of a group of at most
 * NB_BH_WAVE_BODIES Morton-consecutive bodies of one parent cell share every
 * decision; a cell is accepted when its COM lies farther than
 * side / theta + |COM - cell centre| from the group's bounding box (Barnes'
 * modified criterion: the reference's test plus the offset of the mass from the
 * cell centre). Each body therefore accepts a subset of what the reference's
 * test s^2 < theta^2 d^2 (Octree.cpp:320) accepts on the same tree. Measured
 * against the FP64 direct sum on the two-disc system at theta = 0.5 it is more
 * accurate than the reference's v5 solver (Octree.cpp:81-338) at 2^16 and 2^22
 * bodies (tests/test_gpu_bh_v5.py). NB_F_BH_PER_BODY instead opens cells per body
 * with exactly the reference's test (divergent, slower, less accurate on flat
 * systems). */
#define NB_F_BH_PER_BODY 0x4
/* Largest group of bodies that share one walk in the default Barnes-Hut traversal. */
#define NB_BH_WAVE_BODIES 32
/* Barnes-Hut: quantize each axis over its own extent, exactly the inputs of the
 * reference's morton.glsl:68-75 (cells become boxes with the root's aspect
 * ratio), instead of the default cubic root cell (extent = the largest axis). */
#define NB_F_BH_AXIS_BOX 0x8
/* Barnes-Hut range pairing. The build can pair sibling wave ranges of equal
 * copy width so that a pair walks the tree once for both (exactly the same
 * interactions per range, summed in another order). By default it pairs from
 * NB_BH_PAIR_MIN_BODIES bodies on; below that there are too few ranges per
 * wave for a pair's two phases to pay.
 * NB_F_BH_PAIRS pairs at any size, NB_F_BH_SINGLE_RANGES never (it wins if
 * both are set). Pairing is usually faster on large systems, but not always:
 * on a uniform 4M-body ball at leaf size 4 it was slower. nb_tune_leaf times
 * every leaf size both ways on the caller's bodies and sets one of the two. */
#define NB_F_BH_SINGLE_RANGES 0x10
#define NB_F_BH_PAIRS 0x20
#define NB_BH_PAIR_MIN_BODIES (1 << 20)

typedef struct nb_ctx nb_ctx;

/* Create a context and allocate every device buffer for cfg->n bodies. */
int nb_create(nb_ctx** out, const nb_config* cfg);
void nb_destroy(nb_ctx* ctx);
const char* nb_last_error(void);

/* Upload the particle system (replaces ParticleSystem.cpp:75-79 glBufferData at
 * bindings 0-3). acc4 may be NULL (zero initial acceleration, Particle.cpp:4). */
int nb_upload(nb_ctx* ctx, const float* pos4, const float* vel4, const float* acc4,
              const float* mass, int64_t n);
/* FP64 variant: pos3/vel3/acc3 are double[3*n] (x,y,z interleaved). */
int nb_upload_f64(nb_ctx* ctx, const double* pos3, const double* vel3, const double* acc3,
                  const float* mass, int64_t n);

/* Advance nsteps full steps (forces + integration), asynchronously
 * (replaces ParticleSolver::updateParticlePositions, ParticleSolver.h:10). */
int nb_step(nb_ctx* ctx, int32_t nsteps);
int nb_sync(nb_ctx* ctx);

/* Force-only pass at the current positions, no integration; writes float[4*n]
 * (fx,fy,fz,0) — the reference's forces buffer, binding 4. */
int nb_forces(nb_ctx* ctx, float* force4_out);
/* FP64 view of the same forces (compensated FP32 sum widened): double[3*n]. */
int nb_forces_f64(nb_ctx* ctx, double* force3_out);

int nb_download(nb_ctx* ctx, float* pos4, float* vel4, float* acc4);
int nb_download_f64(nb_ctx* ctx, double* pos3, double* vel3, double* acc3);

/* FP64 diagnostics: kinetic and softened potential energy, linear momentum.
 * epot = -G sum_{i<j} m_i m_j / sqrt(r^2 + eps2). */
int nb_energy(nb_ctx* ctx, double* ekin, double* epot, double p[3]);

/* Binary snapshot of the device state: n, eps2, dt, FP64 positions, velocities,
 * accelerations and FP32 masses. Loading it into a context with the same n and
 * continuing reproduces an uninterrupted run bit for bit. The reference's only
 * persistence is the text format (ParticleSystem.cpp:205-215 writer,
 * ParticleSystemFile.cpp:10-90 reader). */
int nb_save_snapshot(nb_ctx* ctx, const char* path);
int nb_load_snapshot(nb_ctx* ctx, const char* path);

/* Morton reordering of the resident state (replaces ParticleSystem::gpuSort,
 * ParticleSystem.cpp:106-171, which is compiled but never called there): the
 * device slots of positions, velocities, accelerations, forces and masses are
 * permuted into the stable (Morton code, slot) order of the current positions,
 * so that bodies close in space are close in memory. Either solver. Uploads,
 * downloads, forces and snapshots keep the caller's body order; only the debug
 * taps below and nb_get_order see slots. An upload returns the slots to the
 * caller's order. */
int nb_reorder(nb_ctx* ctx);
/* order[slot] = the caller's index of the body in that slot (identity until
 * nb_reorder). */
int nb_get_order(nb_ctx* ctx, int32_t* order);

/* Headless render consumer: float[4*n] = (x, y, z, |v|) per body in the
 * caller's order, FP32, from the FP64 state. These are the inputs of the
 * reference's point renderer (vertexShader.glsl:3-9 reads positions and
 * velocities at bindings 0/1; fragmentShader.glsl colours by length(velocity)),
 * which this engine leaves to the consumer. */
int nb_render_view(nb_ctx* ctx, float* xyzs);

/* ---- allocation-free Barnes-Hut force pass --------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc): the force pass of a Barnes-Hut context (the kernels of
 * nb_forces / nb_forces_f64 with NB_SOLVER_BH, same results bit for bit) on
 * caller-owned device buffers, with every intermediate array (Morton keys, sort
 * buffers, octree, wave ranges, counters) in one static device array of the
 * library, about 21 MB per GPU, held until the process ends. It calls no
 * hipMalloc, hipFree, hipMemset or allocating library routine and does not
 * synchronise: it only enqueues kernels on `stream` (a hipStream_t; NULL = the
 * null stream) and returns.
 *   pm4      device float[4 n]: (x, y, z, m) per body
 *   force3   device double[3][n] (structure of arrays: fx[0..n), fy, fz):
 *            G sum_j m_j r_ij / (|r_ij|^2 + eps2)^{3/2}, the FP32 sums widened
 *   status   device int32 (may be NULL): zeroed, then OR-ed with 2 / 4 / 8 on
 *            octree overflow / a walk leaving the node range / stack overflow
 *   leaf_size 1..64 (0 = 8), theta (0 = 0.5), flags NB_F_BH_AXIS_BOX /
 *            NB_F_BH_PER_BODY as for nb_create.
 * 1 <= n <= NB_NOALLOC_MAX_BODIES, else NB_E_ARG with nothing enqueued. Every
 * call on one GPU uses that GPU's one scratch array, which keeps the last
 * call's contents: calls that overlap on one GPU (two streams, or two host
 * threads) race and must be ordered by the caller. */
#define NB_NOALLOC_MAX_BODIES 65536
int nb_bh_forces_noalloc(const float* pm4, int64_t n, float eps2, float G, float theta,
                         int32_t leaf_size, int32_t flags, double* force3, int32_t* status,
                         void* stream);

/* ---- Barnes-Hut debug taps (valid after a BH nb_step or nb_forces) ---- */
/* codes[n]: 30-bit Morton code of the body in slot i (unsorted order);
 * perm[n]: slot at sorted position k (sorted by (code, slot)). */
int nb_debug_morton(nb_ctx* ctx, uint32_t* codes, uint32_t* perm);
/* Flattened octree, breadth-first numbering, children in Morton-digit order. */
typedef struct nb_node_view {
  float com[3];         /* centre of mass */
  float mass;
  int32_t first_child;  /* -1 for a leaf */
  int32_t next;         /* rope: next node in pre-order after this subtree, -1 at end */
  int32_t body_start;   /* first sorted body of the node's range */
  int32_t body_count;   /* bodies in the range */
  int32_t level;        /* 0 = root; cell side = root side / 2^level */
  float open_r2;        /* squared opening radius of the default traversal,
                           (side / theta + |com - cell centre|)^2; -1 if massless */
} nb_node_view;
/* Writes min(cap, node_count) nodes; *node_count receives the total. The
 * root bounding box is written to bbox[6] = (min xyz, max xyz) if non-NULL. */
int nb_debug_tree(nb_ctx* ctx, nb_node_view* nodes, int64_t cap, int64_t* node_count,
                  float* bbox);

/* Cap the default traversal's per-wave LDS stack region (its stack of opened
 * nodes plus a pair's deferred entries) at `entries`, 1..512; 0 restores 512.
 * For tests of the overflow path: a walk that would exceed the cap stops, and
 * the pass returns NB_E_TREE_OVERFLOW. No counterpart in the reference. */
int nb_debug_set_walk_stack(nb_ctx* ctx, int32_t entries);

/* ---- timing ------------------------------------------------------------ */
typedef struct nb_timing {
  double force_ms;      /* summed device time of force kernels (direct or traversal) */
  int64_t force_calls;
  double step_ms;       /* summed device time of whole steps (first to last kernel) */
  int64_t step_calls;
  double build_ms;      /* BH: bbox + encode + sort + build + COM */
  int64_t build_calls;
  int64_t interactions; /* BH: counted pair interactions of the last counted pass */
} nb_timing;
/* Enable (1) or disable (0) HIP-event timing of each step's phases; resets totals. */
int nb_set_timing(nb_ctx* ctx, int32_t enable);
int nb_get_timing(nb_ctx* ctx, nb_timing* out);
/* BH: count body-cell + body-body interactions at the current state (diagnostic
 * traversal, not on the step path). */
int nb_count_interactions(nb_ctx* ctx, int64_t* interactions);
/* BH: the interactions the default traversal's opening test needs when every
 * body walks the same tree on its own (the floor of the group traversal's work;
 * the group shares decisions among up to NB_BH_WAVE_BODIES bodies and does more).
 * Direct sum: n (n - 1). Diagnostic, not on the step path. */
int nb_count_interactions_per_body(nb_ctx* ctx, int64_t* interactions);
/* FP64 direct-sum accelerations of the bodies targets[0..k) (caller indices)
 * from the context's FP32 (x, y, z, m) view widened to FP64, G applied:
 * acc3 = double[3k], interleaved. The error reference for both solvers'
 * FP32 forces at any n (the CPU oracle pins it). Synchronises. */
int nb_direct_f64(nb_ctx* ctx, const int64_t* targets, int64_t k, double* acc3);

/* ---- batched ensembles ---------------------------------------------------
 * Many independent particle systems ("clusters") of uneven size, each with its
 * own eps2, dt and G, advanced together: one persistent force launch takes
 * (cluster, body-block) work items from a device work queue, one launch
 * integrates every cluster, and multi-step calls replay a captured hipGraph.
 * Clusters live in one HBM body pool (max_bodies, each cluster padded to a
 * multiple of 256) and can be added and removed between steps; a cluster's
 * results do not depend on which other clusters share the launch (fixed
 * per-cluster summation order, no cross-cluster tiles).
 * The reference has no counterpart (it simulates one system per process,
 * ParticleSimulation.cpp:4-14); this is the engine's batching extension. */
typedef struct nb_batch nb_batch;
typedef struct nb_batch_config {
  int64_t max_bodies;      /* body pool capacity */
  int32_t max_clusters;
  int32_t integrator;      /* nb_integrator, shared by all clusters */
  int32_t bodies_per_lane; /* 0 = built-in 1 (nb_batch_tune picks per ensemble) */
  int32_t j_split;         /* 0 = built-in 4 */
  int32_t flags;           /* NB_F_NO_GRAPH */
  int32_t pad_;
} nb_batch_config;
int nb_batch_create(nb_batch** out, const nb_batch_config* cfg);
void nb_batch_destroy(nb_batch* b);
/* Admit a cluster (FP64 state, interleaved xyz; acc3 may be NULL). *id receives
 * its slot. NB_E_BUDGET when the pool or the slot table cannot hold it now. */
int nb_batch_add(nb_batch* b, const double* pos3, const double* vel3, const double* acc3,
                 const float* mass, int64_t n, float eps2, float dt, float G, int32_t* id);
int nb_batch_remove(nb_batch* b, int32_t id);
/* Advance every cluster nsteps steps, except that a cluster whose quota q is
 * below nsteps advances q steps (quotas: nb_batch_set_quotas). */
int nb_batch_step(nb_batch* b, int32_t nsteps);
/* Quota of steps per nb_batch_step call for the clusters ids[0..count):
 * steps[k] >= 0, NB_BATCH_UNLIMITED (every cluster's quota when it is added)
 * for no limit. Quotas persist until changed. A serving loop gives short jobs
 * many steps per call and long ones one; a cluster's trajectory does not depend
 * on how its steps are split over calls. */
#define NB_BATCH_UNLIMITED 0x7fffffff
int nb_batch_set_quotas(nb_batch* b, const int32_t* ids, const int32_t* steps, int32_t count);
int nb_batch_sync(nb_batch* b);
/* Force-only pass for every cluster (state untouched); writes cluster id's
 * forces to force3 (double[3n], interleaved). */
int nb_batch_forces(nb_batch* b, int32_t id, double* force3);
int nb_batch_download(nb_batch* b, int32_t id, double* pos3, double* vel3, double* acc3);
/* FP64 kinetic and softened potential energy of one cluster. */
int nb_batch_energy(nb_batch* b, int32_t id, double* ekin, double* epot);
/* Active clusters, pool bodies in use (padded), pair interactions per step. */
int nb_batch_info(nb_batch* b, int32_t* clusters, int64_t* bodies, int64_t* interactions);
int nb_batch_set_timing(nb_batch* b, int32_t enable);
int nb_batch_get_timing(nb_batch* b, nb_timing* out);

/* ---- serving ---------------------------------------------------------------
 * A request loop over many simulation jobs: jobs are submitted at any time,
 * admitted in arrival order while their device memory fits the HBM budget and
 * their estimated step time fits the tick budget (a bounded look-ahead lets
 * small jobs pass a large one that does not fit yet), and every admitted job
 * advances at least one step per nb_server_tick. The tick budget keeps a tick
 * short, so admitted jobs do not share the device with every queued job at
 * once. Direct-sum jobs (n < bh_threshold) share one nb_batch launch;
 * Barnes-Hut jobs run their own contexts on their own streams, so their kernels
 * overlap on the device. Barnes-Hut contexts are sized in power-of-two
 * capacity classes and kept in a pool when their job ends, so a steady stream
 * allocates and frees no device memory. Finished jobs keep their FP64 final
 * state on the host until fetched. A job's result is the same alone or in any
 * mix.
 * The reference has no counterpart (one interactive simulation per process,
 * main.cpp:73-153); this is the engine's serving extension. */
typedef struct nb_server nb_server;
typedef struct nb_server_config {
  int64_t hbm_budget;    /* device bytes jobs may hold; 0 = 60% of free memory */
  int64_t direct_pool;   /* bodies in the shared direct-sum pool; 0 = 2^24 */
  int32_t max_jobs;      /* jobs running at once; 0 = 1024 */
  int32_t bh_threshold;  /* n >= this runs Barnes-Hut; 0 = 65536 */
  float theta;           /* Barnes-Hut opening angle; 0 = 0.5 */
  int32_t leaf_size;     /* Barnes-Hut bodies per leaf; 0 = default */
  int32_t ic_threads;    /* host threads preparing initial conditions; 0 = 4 */
  int32_t lookahead;     /* queued jobs considered per admission pass; 0 = 16 */
  float tick_budget_ms;  /* estimated device time of one tick's steps; 0 = 10 ms.
                            A job larger than the budget runs when it is alone. */
  int32_t policy;        /* NB_SERVE_ARRIVAL (0) or NB_SERVE_SHORTEST */
} nb_server_config;

/* Scheduling policy. ARRIVAL: admission in arrival order among the first
 * `lookahead` queued jobs (a job that does not fit yet lets later ones pass)
 * while the admitted jobs' single steps fit the tick budget; every admitted job
 * advances at least one step per tick, and the budget left over goes to extra
 * steps of the jobs with the least remaining work (up to 32 per tick), so short
 * jobs finish in a few ticks. SHORTEST: preemptive shortest remaining work
 * first: admission in order of estimated total work (steps x step time) up to
 * four ticks' budget, and each tick's budget goes to the jobs with the least
 * remaining work, as many steps each as fit; the others wait resident. It
 * finishes the most jobs per second and can keep a large job waiting while
 * smaller ones keep arriving. */
#define NB_SERVE_ARRIVAL 0
#define NB_SERVE_SHORTEST 1

#define NB_IC_PLUMMER 0 /* a = 50, total mass 1e4, truncated at 10 a */
#define NB_IC_BALL 1    /* uniform ball, R = 100, m = 1 */

typedef struct nb_job {
  int64_t n;
  int32_t steps;
  int32_t ic;            /* NB_IC_* */
  uint64_t seed;
  float eps2, dt, G;     /* G = 0 means 1 */
  int32_t pad_;
} nb_job;

#define NB_JOB_QUEUED 0
#define NB_JOB_RUNNING 1
#define NB_JOB_DONE 2
#define NB_JOB_FAILED (-1)

typedef struct nb_job_status {
  int32_t state;         /* NB_JOB_* */
  int32_t solver;        /* nb_solver chosen for the job */
  int32_t steps_done, steps;
  double t_submit, t_admit, t_done;  /* seconds since nb_server_create */
} nb_job_status;

int nb_server_create(nb_server** out, const nb_server_config* cfg);
void nb_server_destroy(nb_server* s);
int nb_server_submit(nb_server* s, const nb_job* job, int64_t* id);
/* Admit what fits, advance every running job one step, retire finished jobs.
 * running/finished (optional) receive the running count and the jobs retired. */
int nb_server_tick(nb_server* s, int32_t* running, int32_t* finished);
/* Ids of the jobs retired (done or failed) since the last poll, oldest first:
 * up to cap of them are moved into ids, *count receives how many. */
int nb_server_poll(nb_server* s, int64_t* ids, int32_t cap, int32_t* count);
int nb_server_status(nb_server* s, int64_t id, nb_job_status* out);
/* FP64 final state of a finished job (double[3n] each, interleaved). */
int nb_server_result(nb_server* s, int64_t id, double* pos3, double* vel3);
int nb_server_forget(nb_server* s, int64_t id);
double nb_server_now(nb_server* s);
/* Where the serving loop's time went, summed over its ticks (host wall time,
 * ms): selection, the idle wait for a first prepared job, starting jobs,
 * enqueueing steps, waiting for the device, retiring; plus tick and job-step
 * counts and the idle Barnes-Hut contexts in the pool. */
typedef struct nb_server_stats {
  int64_t ticks, steps, pooled_contexts;
  double select_ms, idle_wait_ms, start_ms, enqueue_ms, device_wait_ms, retire_ms;
} nb_server_stats;
int nb_server_get_stats(nb_server* s, nb_server_stats* out);

/* ---- autotuning ---------------------------------------------------------
 * Times candidate launch shapes on the current device with a seeded Plummer
 * sphere of n bodies and installs the fastest as the default of later
 * nb_create / nb_batch_create calls for systems in the same power-of-two size
 * class (explicit non-zero config fields still win). Candidates: direct sum
 * bodies_per_lane x j_split in {1,2,4} x {1,2,4,8,16} (the shapes a kernel
 * exists for: 8 needs bodies_per_lane <= 2, 16 needs 1); Barnes-Hut leaf_size in
 * {1,2,4,8,16,32}. The reference hard-codes its work-group sizes (64, 320,
 * 1024) in the shaders. */
typedef struct nb_tuning {
  int64_t n;
  int32_t bodies_per_lane;
  int32_t j_split;
  int32_t leaf_size;
  int32_t bh_single_ranges;  /* nb_tune_leaf: 1 if single ranges timed faster than pairs
                                (NB_F_BH_SINGLE_RANGES set), 0 if pairs (NB_F_BH_PAIRS set) */
  double direct_ms;  /* force pass at the chosen shape (0 if not tuned) */
  double bh_ms;      /* Barnes-Hut build + traversal at the chosen leaf size */
} nb_tuning;
#define NB_TUNE_DIRECT 0x1
#define NB_TUNE_BH 0x2
int nb_autotune(int64_t n, int32_t what, int32_t reps, nb_tuning* out);
/* Batched ensembles: time every valid bodies_per_lane x j_split shape of the
 * persistent force pass on the batch's current clusters (reps passes each; the
 * particle state is not changed) and keep the fastest for later steps.
 * out->direct_ms = its force pass, out->n = bodies in the batch. */
int nb_batch_tune(nb_batch* b, int32_t reps, nb_tuning* out);
/* Barnes-Hut leaf size and range pairing tuned on the workload itself: times
 * leaf_size in {1,2,4,8,16,32}, each with NB_F_BH_PAIRS and with
 * NB_F_BH_SINGLE_RANGES, on the context's current bodies (the particle state
 * is not changed): one untimed round, then reps rounds of one force pass per
 * candidate, interleaved, each candidate keeping its fastest pass. The fastest
 * candidate serves the context's later steps; out->leaf_size,
 * out->bh_single_ranges and out->bh_ms (its fastest pass) receive the choice. */
int nb_tune_leaf(nb_ctx* ctx, int32_t reps, nb_tuning* out);
/* The defaults nb_create would use for n bodies (tuned or built in). */
int nb_get_tuning(int64_t n, nb_tuning* out);

/* Library / device description for logs: writes a NUL-terminated string. */
int nb_device_info(char* buf, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif /* NBODY_H */
