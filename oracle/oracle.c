/*
 * oracle.c — CPU restatement of the reference's N-body hot path.
 * TEST INFRASTRUCTURE ONLY (checker + timed CPU baseline); see oracle.h.
 *
 * Parity pins (tests/test_oracle_*.py): two-body and Lagrange-triangle closed
 * forms, momentum antisymmetry, translation/permutation equivariance, hand-derived
 * integrator steps, exhaustive expandBits, brute-force sort/tree construction,
 * theta=0 Barnes-Hut == direct sum, and distant-cluster monopole limits.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int clamp_threads(int t) { return t < 1 ? 1 : t; }

/* ------------------------------------------------------------------------ */
/* FP64 direct sum: closed form of ParticleSolverCPUSequential.cpp:26-45.    */
/* ------------------------------------------------------------------------ */
void or_direct_f64(int64_t n, const double* pos3, const double* mass, double eps2, double G,
                   int64_t nt, const int64_t* targets, double* acc3, int nthreads) {
  if (!targets) nt = n;
#pragma omp parallel for schedule(dynamic, 16) num_threads(clamp_threads(nthreads))
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t i = targets ? targets[t] : t;
    const double xi = pos3[3 * i], yi = pos3[3 * i + 1], zi = pos3[3 * i + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue; /* self excluded (ParticleSolverCPUSequential.cpp:34) */
      const double dx = pos3[3 * j] - xi, dy = pos3[3 * j + 1] - yi, dz = pos3[3 * j + 2] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double s = mass[j] / (r2 * sqrt(r2));
      ax += s * dx;
      ay += s * dy;
      az += s * dz;
    }
    acc3[3 * t] = G * ax;
    acc3[3 * t + 1] = G * ay;
    acc3[3 * t + 2] = G * az;
  }
}

/* ------------------------------------------------------------------------ */
/* v1 / v2 force pass in FP32 (ParticleSolverCPUSequential.cpp:26-45,        */
/* ParticleSolverCPUParallel.cpp:25-38).                                     */
/* ------------------------------------------------------------------------ */
void or_pp_forces(int64_t n, const float* pos4, const float* mass4, float eps2, float G,
                  float* force4, int64_t t0, int64_t t1, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(clamp_threads(nthreads))
  for (int64_t i = t0; i < t1; ++i) {
    const float px = pos4[4 * i], py = pos4[4 * i + 1], pz = pos4[4 * i + 2], pw = pos4[4 * i + 3];
    float tx = 0.f, ty = 0.f, tz = 0.f, tw = 0.f;
    for (int64_t j = 0; j < n; ++j) {
      if (i == j) continue;
      /* vector_i_j = positions[j] - particlePosition (vec4, w included) */
      const float vx = pos4[4 * j] - px, vy = pos4[4 * j + 1] - py;
      const float vz = pos4[4 * j + 2] - pz, vw = pos4[4 * j + 3] - pw;
      /* glm::length2(vec4) = dot(v, v) = (x*x + y*y) + (z*z + w*w) (glm 0.9.9.8) */
      const float l2 = (vx * vx + vy * vy) + (vz * vz + vw * vw);
      const float d = powf(l2 + eps2, 1.5f);
      const float s = (G * mass4[4 * j]) / d;
      tx += s * vx;
      ty += s * vy;
      tz += s * vz;
      tw += s * vw;
    }
    force4[4 * i] = tx;
    force4[4 * i + 1] = ty;
    force4[4 * i + 2] = tz;
    force4[4 * i + 3] = tw;
  }
}

/* Lagged KDK (ParticleSystem.h:34-51): vel += acc*dt/2; pos += vel*dt; acc = force;
 * vel += acc*dt/2 — all four vec4 components. */
void or_kdk_update(int64_t n, float* pos4, float* vel4, float* acc4, const float* force4,
                   float dt, int nthreads) {
  const float h = dt * 0.5f;
#pragma omp parallel for schedule(static) num_threads(clamp_threads(nthreads))
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 4; ++c) {
      const int64_t k = 4 * i + c;
      vel4[k] += acc4[k] * h;
      pos4[k] += vel4[k] * dt;
      acc4[k] = force4[k];
      vel4[k] += acc4[k] * h;
    }
  }
}

/* Semi-implicit Euler (updateParticles.glsl:44-55), xyz only. */
void or_euler_update(int64_t n, float* pos4, float* vel4, const float* force4, float dt) {
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      const int64_t k = 4 * i + c;
      vel4[k] = vel4[k] + force4[k] * dt;
      pos4[k] = pos4[k] + vel4[k] * dt;
    }
  }
}

void or_pp_step(int64_t n, float* pos4, float* vel4, float* acc4, const float* mass4,
                float* force4, float eps2, float G, float dt, int nthreads) {
  or_pp_forces(n, pos4, mass4, eps2, G, force4, 0, n, nthreads);
  or_kdk_update(n, pos4, vel4, acc4, force4, dt, nthreads);
}

void or_kdk_update_f64(int64_t n, double* pos3, double* vel3, double* acc3,
                       const double* force3, double dt) {
  const double h = dt * 0.5;
  for (int64_t k = 0; k < 3 * n; ++k) {
    vel3[k] += acc3[k] * h;
    pos3[k] += vel3[k] * dt;
    acc3[k] = force3[k];
    vel3[k] += acc3[k] * h;
  }
}

void or_euler_update_f64(int64_t n, double* pos3, double* vel3, double* acc3,
                         const double* force3, double dt) {
  for (int64_t k = 0; k < 3 * n; ++k) {
    acc3[k] = force3[k];
    vel3[k] += force3[k] * dt;
    pos3[k] += vel3[k] * dt;
  }
}

void or_energy_f64(int64_t n, const double* pos3, const double* vel3, const double* mass,
                   double eps2, double G, double* ekin, double* epot, double p[3], int nthreads) {
  double* part = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
#pragma omp parallel for schedule(dynamic, 16) num_threads(clamp_threads(nthreads))
  for (int64_t i = 0; i < n; ++i) {
    double phi = 0.0;
    for (int64_t j = i + 1; j < n; ++j) {
      const double dx = pos3[3 * j] - pos3[3 * i], dy = pos3[3 * j + 1] - pos3[3 * i + 1];
      const double dz = pos3[3 * j + 2] - pos3[3 * i + 2];
      phi += mass[j] / sqrt(dx * dx + dy * dy + dz * dz + eps2);
    }
    part[i] = mass[i] * phi;
  }
  double ek = 0.0, ep = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double vx = vel3[3 * i], vy = vel3[3 * i + 1], vz = vel3[3 * i + 2];
    ek += 0.5 * mass[i] * (vx * vx + vy * vy + vz * vz);
    ep += part[i];
    px += mass[i] * vx;
    py += mass[i] * vy;
    pz += mass[i] * vz;
  }
  free(part);
  *ekin = ek;
  *epot = -G * ep;
  p[0] = px;
  p[1] = py;
  p[2] = pz;
}

/* ------------------------------------------------------------------------ */
/* v5 sequential octree (Octree.cpp:9-350, Node.h:7-93) with int32 links.    */
/* ------------------------------------------------------------------------ */
typedef struct v5_node {
  float mass;          /* Node.mass.x; -1 = empty leaf (createEmptyNode) */
  int32_t first_child; /* Node.mass.y (float in the reference, D3) */
  int32_t next;        /* Node.mass.z */
  float com[3];
  float minb[3];
  float maxb[3];
} v5_node;

typedef struct v5_tree {
  v5_node* nodes;
  int32_t* parents;
  int64_t max_nodes, node_count, parent_count;
  int overflow;
} v5_tree;

/* Node::getQuadrantCenter (Node.h:91-98): min + extent/2 per axis. */
static void v5_center(const v5_node* nd, float c[3]) {
  for (int a = 0; a < 3; ++a) c[a] = nd->minb[a] + (nd->maxb[a] - nd->minb[a]) / 2.f;
}

/* Octree::getNextNode (Octree.cpp:148-153): j = (x<cx) | (y<cy)<<1 | (z<cz)<<2. */
static int32_t v5_next_node(const v5_tree* t, const float p[3], int32_t i) {
  float c[3];
  v5_center(&t->nodes[i], c);
  const int q = (p[0] < c[0]) | ((p[1] < c[1]) << 1) | ((p[2] < c[2]) << 2);
  return t->nodes[i].first_child + q;
}

/* Octree::subdivide (Octree.cpp:159-231). A set octant bit selects the lower half. */
static void v5_subdivide(v5_tree* t, int32_t i) {
  if (t->node_count + 8 > t->max_nodes) { /* D5: the reference overflows silently */
    t->overflow = 1;
    return;
  }
  v5_node* nd = &t->nodes[i];
  nd->first_child = (int32_t)t->node_count;
  t->parents[t->parent_count++] = i;
  t->node_count += 8;
  const float w = nd->maxb[0] - nd->minb[0], h = nd->maxb[1] - nd->minb[1];
  const float d = nd->maxb[2] - nd->minb[2];
  const int32_t first = nd->first_child;
  for (int j = 0; j < 8; ++j) {
    v5_node* c = &t->nodes[first + j];
    if (j & 4) { c->minb[2] = nd->minb[2]; c->maxb[2] = nd->minb[2] + (d / 2.f); }
    else       { c->minb[2] = nd->minb[2] + (d / 2.f); c->maxb[2] = nd->maxb[2]; }
    if (j & 2) { c->minb[1] = nd->minb[1]; c->maxb[1] = nd->minb[1] + (h / 2.f); }
    else       { c->minb[1] = nd->minb[1] + (h / 2.f); c->maxb[1] = nd->maxb[1]; }
    if (j & 1) { c->minb[0] = nd->minb[0]; c->maxb[0] = nd->minb[0] + (w / 2.f); }
    else       { c->minb[0] = nd->minb[0] + (w / 2.f); c->maxb[0] = nd->maxb[0]; }
    c->mass = -1.f; /* createEmptyNode */
    c->first_child = -1;
    if (j < 7) c->next = first + j + 1;
  }
  t->nodes[first + 7].next = nd->next;
}

static void v5_put(v5_tree* t, const float p[3], float m, int32_t i) {
  v5_node* nd = &t->nodes[i];
  nd->com[0] = p[0];
  nd->com[1] = p[1];
  nd->com[2] = p[2];
  nd->mass = m;
}

/* Octree::insert (Octree.cpp:81-139). */
static void v5_insert(v5_tree* t, const float p[3], float m) {
  int32_t i = 0;
  while (i < t->max_nodes && !t->overflow) {
    v5_node* nd = &t->nodes[i];
    const int leaf = nd->first_child < 0, occ = nd->mass > 0.f;
    if (leaf && !occ) { v5_put(t, p, m, i); return; }
    if (leaf && occ) {
      const float p2[3] = {nd->com[0], nd->com[1], nd->com[2]};
      const float m2 = nd->mass;
      v5_subdivide(t, i);
      if (t->overflow) return;
      int32_t c1 = v5_next_node(t, p, i), c2 = v5_next_node(t, p2, i);
      /* length2(pos2 - pos) on vec4 with w = 0: (x*x + y*y) + (z*z + 0) */
      const float dx = p2[0] - p[0], dy = p2[1] - p[1], dz = p2[2] - p[2];
      const float dist = (dx * dx + dy * dy) + (dz * dz + 0.f);
      while (c1 == c2) {
        if (dist < 1e-6) { v5_put(t, p, m2 + m, c1); return; } /* merge rule (:112-114) */
        v5_subdivide(t, c1);
        if (t->overflow) return;
        c1 = v5_next_node(t, p, c1);
        c2 = v5_next_node(t, p2, c2);
      }
      v5_put(t, p, m, c1);
      v5_put(t, p2, m2, c2);
      return;
    }
    i = v5_next_node(t, p, i);
  }
}

/* Octree::propagate (Octree.cpp:238-258): reverse creation order. */
static void v5_propagate(v5_tree* t) {
  for (int64_t k = t->parent_count - 1; k >= 0; --k) {
    v5_node* par = &t->nodes[t->parents[k]];
    float c[3] = {0.f, 0.f, 0.f}, m = 0.f;
    for (int j = 0; j < 8; ++j) {
      const v5_node* ch = &t->nodes[par->first_child + j];
      if (!(ch->mass > 0.f)) continue;
      c[0] += ch->com[0] * ch->mass;
      c[1] += ch->com[1] * ch->mass;
      c[2] += ch->com[2] * ch->mass;
      m += ch->mass;
    }
    par->com[0] = c[0] / m;
    par->com[1] = c[1] / m;
    par->com[2] = c[2] / m;
    par->mass = m;
  }
}

/* Octree::prune (Octree.cpp:264-294): creation order. */
static void v5_prune(v5_tree* t) {
  for (int64_t k = 0; k < t->parent_count; ++k) {
    v5_node* par = &t->nodes[t->parents[k]];
    if (!(par->mass > 0.f)) { par->first_child = par->next; continue; }
    int32_t first = -1, last = -1;
    for (int j = 0; j < 8; ++j) {
      const int32_t ci = par->first_child + j;
      if (t->nodes[ci].mass > 0.f) {
        if (first == -1) { first = last = ci; continue; }
        t->nodes[last].next = ci;
        last = ci;
      }
    }
    t->nodes[last].next = par->next;
    par->first_child = first;
  }
}

/* Octree::computeGravityForce (Octree.cpp:303-338). */
static void v5_force(const v5_tree* t, const float p[3], float eps2, float G, float theta2,
                     float out[3]) {
  float f[3] = {0.f, 0.f, 0.f};
  int32_t i = 0;
  while (i >= 0) {
    const v5_node* nd = &t->nodes[i];
    const float sx = nd->maxb[0] - nd->minb[0], sy = nd->maxb[1] - nd->minb[1];
    const float sz = nd->maxb[2] - nd->minb[2];
    const float s = fmaxf(fmaxf(sx, sy), sz);
    const float s2 = s * s;
    const float vx = nd->com[0] - p[0], vy = nd->com[1] - p[1], vz = nd->com[2] - p[2];
    const float d2 = (vx * vx + vy * vy) + (vz * vz + 0.f);
    if (nd->first_child < 0 || s2 < theta2 * d2) {
      if (nd->mass > 0.f && d2 > 0) {
        const float e = d2 + eps2;
        const float inv = 1.0f / powf(e, 1.5f);
        const float gm = G * nd->mass;
        f[0] += (vx * gm) * inv;
        f[1] += (vy * gm) * inv;
        f[2] += (vz * gm) * inv;
      }
      i = nd->next;
    } else {
      i = nd->first_child;
    }
  }
  out[0] = f[0];
  out[1] = f[1];
  out[2] = f[2];
}

int or_v5_forces(int64_t n, const float* pos4, const float* mass4, float eps2, float G,
                 float theta, float* force4, double pool_factor, int64_t* node_count,
                 int nthreads) {
  v5_tree t;
  memset(&t, 0, sizeof(t));
  /* Octree::Octree (Octree.cpp:9-16) */
  t.max_nodes = pool_factor > 0 ? (int64_t)(pool_factor * (double)n) : (n < 200 ? n * 200 : n * 8);
  if (t.max_nodes < 9) t.max_nodes = 9;
  t.nodes = (v5_node*)calloc((size_t)t.max_nodes, sizeof(v5_node));
  t.parents = (int32_t*)calloc((size_t)t.max_nodes, sizeof(int32_t));
  if (!t.nodes || !t.parents) { free(t.nodes); free(t.parents); return -3; }
  /* Octree::adjustBoundingBox + reset (Octree.cpp:23-72) */
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = pos4[4 * i + a];
      if (v < mn[a]) mn[a] = v;
      if (v > mx[a]) mx[a] = v;
    }
  v5_node* root = &t.nodes[0];
  for (int a = 0; a < 3; ++a) { root->minb[a] = mn[a]; root->maxb[a] = mx[a]; }
  root->first_child = -1;
  root->mass = -1.f;
  root->next = -1;
  t.node_count = 1;
  /* ParticleSolverBHutCPUSeq::updateParticlePositions (ParticleSolverBHutCPUSeq.cpp:16-36) */
  for (int64_t i = 0; i < n && !t.overflow; ++i) v5_insert(&t, &pos4[4 * i], mass4[4 * i]);
  if (node_count) *node_count = t.node_count;
  if (t.overflow) { free(t.nodes); free(t.parents); return -5; }
  v5_propagate(&t);
  v5_prune(&t);
  const float theta2 = theta * theta;
#pragma omp parallel for schedule(dynamic, 64) num_threads(clamp_threads(nthreads))
  for (int64_t i = 0; i < n; ++i) {
    float f[3];
    v5_force(&t, &pos4[4 * i], eps2, G, theta2, f);
    force4[4 * i] = f[0];
    force4[4 * i + 1] = f[1];
    force4[4 * i + 2] = f[2];
    force4[4 * i + 3] = 0.f;
  }
  free(t.nodes);
  free(t.parents);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* v6 parallel octree (ParallelOctreeCPU.cpp:7-505), int32 links.            */
/* ------------------------------------------------------------------------ */
static double wall_s(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

/* Subtree insertion of ParallelOctreeCPU::insert(pos, mass, subTreeId)
 * (ParallelOctreeCPU.cpp:359-414): the v5 loop with the node counter and parent
 * list of the task, and the merge rule d^2 < 1e-3 or 10 extra subdivisions. */
typedef struct v6_task {
  int64_t base, cap, count; /* node pool [base, base + cap), nodes used */
  int32_t* parents;
  int64_t nparents;
  int overflow;
} v6_task;

static void v6_subdivide(v5_node* nodes, v6_task* tk, int32_t i) {
  if (tk->count + 8 > tk->cap) { tk->overflow = 1; return; }
  const int32_t first = (int32_t)(tk->base + tk->count);
  v5_node* nd = &nodes[i];
  nd->first_child = first;
  tk->parents[tk->nparents++] = i;
  tk->count += 8;
  const float w = nd->maxb[0] - nd->minb[0], h = nd->maxb[1] - nd->minb[1];
  const float d = nd->maxb[2] - nd->minb[2];
  for (int j = 0; j < 8; ++j) {
    v5_node* c = &nodes[first + j];
    if (j & 4) { c->minb[2] = nd->minb[2]; c->maxb[2] = nd->minb[2] + (d / 2.f); }
    else       { c->minb[2] = nd->minb[2] + (d / 2.f); c->maxb[2] = nd->maxb[2]; }
    if (j & 2) { c->minb[1] = nd->minb[1]; c->maxb[1] = nd->minb[1] + (h / 2.f); }
    else       { c->minb[1] = nd->minb[1] + (h / 2.f); c->maxb[1] = nd->maxb[1]; }
    if (j & 1) { c->minb[0] = nd->minb[0]; c->maxb[0] = nd->minb[0] + (w / 2.f); }
    else       { c->minb[0] = nd->minb[0] + (w / 2.f); c->maxb[0] = nd->maxb[0]; }
    c->mass = -1.f;
    c->first_child = -1;
    if (j < 7) c->next = first + j + 1;
  }
  nodes[first + 7].next = nd->next;
}

static int32_t v6_next(const v5_node* nodes, const float p[3], int32_t i) {
  float c[3];
  v5_center(&nodes[i], c);
  return nodes[i].first_child + ((p[0] < c[0]) | ((p[1] < c[1]) << 1) | ((p[2] < c[2]) << 2));
}

static void v6_insert(v5_node* nodes, v6_task* tk, int32_t root, const float p[3], float m) {
  int32_t i = root;
  while (!tk->overflow) {
    v5_node* nd = &nodes[i];
    const int leaf = nd->first_child < 0, occ = nd->mass > 0.f;
    if (leaf && !occ) { nd->com[0] = p[0]; nd->com[1] = p[1]; nd->com[2] = p[2]; nd->mass = m; return; }
    if (leaf && occ) {
      const float p2[3] = {nd->com[0], nd->com[1], nd->com[2]};
      const float m2 = nd->mass;
      v6_subdivide(nodes, tk, i);
      if (tk->overflow) return;
      int32_t c1 = v6_next(nodes, p, i), c2 = v6_next(nodes, p2, i);
      const float dx = p2[0] - p[0], dy = p2[1] - p[1], dz = p2[2] - p[2];
      const float dist = (dx * dx + dy * dy) + (dz * dz + 0.f);
      int count = 0;
      while (c1 == c2) {
        if (dist < 1e-3 || count++ == 10) {
          v5_node* c = &nodes[c1];
          c->com[0] = p[0]; c->com[1] = p[1]; c->com[2] = p[2];
          c->mass = m2 + m;
          return;
        }
        v6_subdivide(nodes, tk, c1);
        if (tk->overflow) return;
        c1 = v6_next(nodes, p, c1);
        c2 = v6_next(nodes, p2, c2);
      }
      v5_node* a = &nodes[c1];
      a->com[0] = p[0]; a->com[1] = p[1]; a->com[2] = p[2]; a->mass = m;
      v5_node* b = &nodes[c2];
      b->com[0] = p2[0]; b->com[1] = p2[1]; b->com[2] = p2[2]; b->mass = m2;
      return;
    }
    i = v6_next(nodes, p, i);
  }
}

static void v6_propagate_list(v5_node* nodes, const int32_t* parents, int64_t k0, int64_t k1) {
  for (int64_t k = k1 - 1; k >= k0; --k) {
    v5_node* par = &nodes[parents[k]];
    float c[3] = {0.f, 0.f, 0.f}, m = 0.f;
    for (int j = 0; j < 8; ++j) {
      const v5_node* ch = &nodes[par->first_child + j];
      if (!(ch->mass > 0.f)) continue;
      c[0] += ch->com[0] * ch->mass;
      c[1] += ch->com[1] * ch->mass;
      c[2] += ch->com[2] * ch->mass;
      m += ch->mass;
    }
    par->com[0] = c[0] / m; par->com[1] = c[1] / m; par->com[2] = c[2] / m;
    par->mass = m;
  }
}

static void v6_prune_list(v5_node* nodes, const int32_t* parents, int64_t k0, int64_t k1) {
  for (int64_t k = k0; k < k1; ++k) {
    v5_node* par = &nodes[parents[k]];
    if (!(par->mass > 0.f)) { par->first_child = par->next; continue; }
    int32_t first = -1, last = -1;
    for (int j = 0; j < 8; ++j) {
      const int32_t ci = par->first_child + j;
      if (nodes[ci].mass > 0.f) {
        if (first == -1) { first = last = ci; continue; }
        nodes[last].next = ci;
        last = ci;
      }
    }
    nodes[last].next = par->next;
    par->first_child = first;
  }
}

int or_v6_forces(int64_t n, const float* pos4, const float* mass4, float eps2, float G,
                 float theta, float* force4, int64_t t0, int64_t t1, int nthreads,
                 double* build_s, double* force_s) {
  nthreads = clamp_threads(nthreads);
  const double tb = wall_s();
  /* ParallelOctreeCPU ctor (:7-56): top tree of depth <= 3 */
  int64_t top = 0, width = 1;
  int depth = -1;
  while (top < n && depth < 3) {
    if (top + width >= n) break;
    top += width;
    width *= 8;
    depth++;
  }
  if (depth < 0) { top = 1; depth = 0; width = 1; }
  const int64_t ntasks = (int64_t)1 << (3 * depth);
  /* bucket counts per task (deepest top cells), then per-task pools sized by need */
  int64_t* tcount = (int64_t*)calloc((size_t)ntasks, sizeof(int64_t));
  int64_t* tbase = (int64_t*)calloc((size_t)ntasks + 1, sizeof(int64_t));
  int32_t* troot = (int32_t*)malloc(sizeof(int32_t) * (size_t)ntasks);
  for (int64_t t = 0; t < ntasks; ++t) troot[t] = -1;
  /* the top tree is built first so cells can be found */
  const int64_t top_nodes = top;
  v5_node* topn = (v5_node*)calloc((size_t)top_nodes, sizeof(v5_node));
  float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float mn0 = FLT_MAX, mn1 = FLT_MAX, mn2 = FLT_MAX, mx0 = -FLT_MAX, mx1 = -FLT_MAX, mx2 = -FLT_MAX;
#pragma omp parallel for num_threads(nthreads) reduction(min : mn0, mn1, mn2) reduction(max : mx0, mx1, mx2)
  for (int64_t i = 0; i < n; ++i) {
    const float* p = &pos4[4 * i];
    if (p[0] < mn0) mn0 = p[0];
    if (p[1] < mn1) mn1 = p[1];
    if (p[2] < mn2) mn2 = p[2];
    if (p[0] > mx0) mx0 = p[0];
    if (p[1] > mx1) mx1 = p[1];
    if (p[2] > mx2) mx2 = p[2];
  }
  mn[0] = mn0; mn[1] = mn1; mn[2] = mn2; mx[0] = mx0; mx[1] = mx1; mx[2] = mx2;
  /* complete top tree: subdivide nodes 0, 1, 2, ... until top_nodes exist (:127-131) */
  v5_node* root = &topn[0];
  for (int a = 0; a < 3; ++a) { root->minb[a] = mn[a]; root->maxb[a] = mx[a]; }
  root->first_child = -1;
  root->mass = -1.f;
  root->next = -1;
  int32_t* top_parents = (int32_t*)malloc(sizeof(int32_t) * (size_t)(top_nodes + ntasks));
  int64_t ntop_parents = 0;
  {
    v6_task tk = {0, top_nodes, 1, top_parents, 0, 0};
    for (int32_t r = 0; tk.count < top_nodes; ++r) v6_subdivide(topn, &tk, r);
    ntop_parents = tk.nparents;
  }
  /* fill the tasks under one lock per deepest cell (:134-152) */
  int32_t* cell_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
#pragma omp parallel for num_threads(nthreads)
  for (int64_t j = 0; j < n; ++j) {
    int32_t i = 0;
    for (int d = 0; d < depth; ++d) i = v6_next(topn, &pos4[4 * j], i);
    cell_of[j] = i;
  }
  omp_lock_t* locks = (omp_lock_t*)malloc(sizeof(omp_lock_t) * (size_t)ntasks);
  for (int64_t t = 0; t < ntasks; ++t) omp_init_lock(&locks[t]);
  for (int64_t j = 0; j < n; ++j) tcount[cell_of[j] % ntasks]++;
  for (int64_t t = 0; t < ntasks; ++t) tbase[t + 1] = tbase[t] + tcount[t];
  int32_t* tparts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int64_t* tfill = (int64_t*)calloc((size_t)ntasks, sizeof(int64_t));
#pragma omp parallel for num_threads(nthreads)
  for (int64_t j = 0; j < n; ++j) {
    const int32_t i = cell_of[j];
    const int64_t t = i % ntasks;
    omp_set_lock(&locks[t]);
    tparts[tbase[t] + tfill[t]++] = (int32_t)j;
    troot[t] = i;
    omp_unset_lock(&locks[t]);
  }
  /* node pools: top tree first, then one pool per task (16 nodes per body + 64) */
  int64_t* pbase = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ntasks + 1));
  pbase[0] = top_nodes;
  for (int64_t t = 0; t < ntasks; ++t) pbase[t + 1] = pbase[t] + 16 * tcount[t] + 64;
  const int64_t total_nodes = pbase[ntasks];
  v5_node* nodes = (v5_node*)malloc(sizeof(v5_node) * (size_t)total_nodes);
  int32_t* sparents = (int32_t*)malloc(sizeof(int32_t) * (size_t)total_nodes);
  memcpy(nodes, topn, sizeof(v5_node) * (size_t)top_nodes);
  v6_task* tasks = (v6_task*)calloc((size_t)ntasks, sizeof(v6_task));
  int overflow = 0;
  /* one subtree per task, schedule(dynamic) (:158-189) */
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(| : overflow)
  for (int64_t t = 0; t < ntasks; ++t) {
    if (troot[t] < 0) continue;
    v6_task* tk = &tasks[t];
    tk->base = pbase[t];
    tk->cap = pbase[t + 1] - pbase[t];
    tk->count = 1;
    tk->parents = &sparents[pbase[t]];
    const int32_t r = (int32_t)pbase[t];
    v5_node* oct = &nodes[r];
    for (int a = 0; a < 3; ++a) { oct->minb[a] = nodes[troot[t]].minb[a]; oct->maxb[a] = nodes[troot[t]].maxb[a]; }
    oct->first_child = -1;
    oct->mass = -1.f;
    oct->next = nodes[troot[t]].next;
    nodes[troot[t]].first_child = r;
    for (int64_t k = 0; k < tcount[t]; ++k) {
      const int32_t pid = tparts[tbase[t] + k];
      v6_insert(nodes, tk, r, &pos4[4 * pid], mass4[4 * pid]);
    }
    v6_propagate_list(nodes, tk->parents, 0, tk->nparents);
    overflow |= tk->overflow;
  }
  /* top-level propagate (:262-291) and prune (:208-245) */
  for (int64_t c = top_nodes - ntasks; c < top_nodes; ++c)
    if (tcount[c % ntasks] > 0) top_parents[ntop_parents++] = (int32_t)c;
#pragma omp parallel for num_threads(nthreads)
  for (int64_t k = ntop_parents - 1; k >= top_nodes - ntasks; --k) {
    const int32_t pi = top_parents[k];
    const v5_node* ch = &nodes[nodes[pi].first_child];
    if (!(ch->mass > 0.f)) continue;
    nodes[pi].com[0] = ch->com[0]; nodes[pi].com[1] = ch->com[1]; nodes[pi].com[2] = ch->com[2];
    nodes[pi].mass = ch->mass;
  }
  v6_propagate_list(nodes, top_parents, 0, top_nodes - ntasks);
  v6_prune_list(nodes, top_parents, 0, top_nodes - ntasks);
#pragma omp parallel for num_threads(nthreads)
  for (int64_t c = top_nodes - ntasks; c < top_nodes; ++c)
    if (nodes[c].mass > 0.f) nodes[nodes[c].first_child].next = nodes[c].next;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
  for (int64_t t = 0; t < ntasks; ++t)
    if (tcount[t] > 0 && troot[t] >= 0) v6_prune_list(nodes, tasks[t].parents, 0, tasks[t].nparents);
  const double tf = wall_s();
  if (!overflow) {
    v5_tree view;
    view.nodes = nodes;
    const float theta2 = theta * theta;
    /* force loop, schedule(dynamic) (ParticleSolverBHutCPUParallel.cpp:28-31) */
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int64_t i = t0; i < t1; ++i) {
      float f[3];
      v5_force(&view, &pos4[4 * i], eps2, G, theta2, f);
      force4[4 * i] = f[0];
      force4[4 * i + 1] = f[1];
      force4[4 * i + 2] = f[2];
      force4[4 * i + 3] = 0.f;
    }
  }
  const double te = wall_s();
  if (build_s) *build_s = tf - tb;
  if (force_s) *force_s = te - tf;
  for (int64_t t = 0; t < ntasks; ++t) omp_destroy_lock(&locks[t]);
  free(locks); free(tcount); free(tbase); free(troot); free(topn); free(top_parents);
  free(cell_of); free(tparts); free(tfill); free(pbase); free(nodes); free(sparents); free(tasks);
  return overflow ? -5 : 0;
}

/* ------------------------------------------------------------------------ */
/* CPU build of the engine's Morton / sort / octree.                         */
/* ------------------------------------------------------------------------ */
void or_bbox(int64_t n, const float* pos4, float bmin[3], float bmax[3]) {
  for (int a = 0; a < 3; ++a) { bmin[a] = FLT_MAX; bmax[a] = -FLT_MAX; }
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = pos4[4 * i + a];
      if (v < bmin[a]) bmin[a] = v;
      if (v > bmax[a]) bmax[a] = v;
    }
}

/* morton.glsl:37-43 expandBits: 10 bits -> 30 bits, two zeros between bits. */
uint32_t or_expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

/* Quantize one coordinate: morton.glsl:49-51 and :68-75. */
static float axis_extent(float lo, float hi) {
  const float ext = hi - lo;
  return ext > 1e-6f ? ext : 1e-6f;              /* max(extent, 1e-6) */
}

static uint32_t quantize(float p, float lo, float ext) {
  float nrm = (p - lo) / ext;                    /* IEEE division */
  nrm = fminf(fmaxf(nrm, 0.f), 1.f);             /* clamp(n, 0, 1) */
  const float g = fminf(fmaxf(nrm * 1024.f, 0.f), 1023.f);
  return (uint32_t)g;                            /* uint() truncates */
}

void or_morton(int64_t n, const float* pos4, const float bmin[3], const float bmax[3], int cubic,
               uint32_t* codes) {
  float ext[3];
  for (int a = 0; a < 3; ++a) ext[a] = axis_extent(bmin[a], bmax[a]);
  if (cubic) {
    const float e = fmaxf(fmaxf(ext[0], ext[1]), ext[2]);
    ext[0] = ext[1] = ext[2] = e;
  }
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t qx = quantize(pos4[4 * i], bmin[0], ext[0]);
    const uint32_t qy = quantize(pos4[4 * i + 1], bmin[1], ext[1]);
    const uint32_t qz = quantize(pos4[4 * i + 2], bmin[2], ext[2]);
    codes[i] = or_expand_bits(qx) * 4u + or_expand_bits(qy) * 2u + or_expand_bits(qz);
  }
}

/* LSD counting sort, 8 bits per pass: stable, so ties keep index order. */
void or_sort_codes(int64_t n, const uint32_t* codes, uint32_t* perm, uint32_t* codes_sorted) {
  uint32_t* k0 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  uint32_t* v0 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  uint32_t* k1 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  uint32_t* v1 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) { k0[i] = codes[i]; v0[i] = (uint32_t)i; }
  for (int shift = 0; shift < 32; shift += 8) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; ++i) cnt[((k0[i] >> shift) & 255u) + 1]++;
    for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t d = (k0[i] >> shift) & 255u;
      const int64_t o = cnt[d]++;
      k1[o] = k0[i];
      v1[o] = v0[i];
    }
    uint32_t* t;
    t = k0; k0 = k1; k1 = t;
    t = v0; v0 = v1; v1 = t;
  }
  for (int64_t i = 0; i < n; ++i) {
    perm[i] = v0[i];
    if (codes_sorted) codes_sorted[i] = k0[i];
  }
  free(k0); free(v0); free(k1); free(v1);
}

int64_t or_tree_bound(int64_t n) {
  int64_t total = 0, width = 1;
  for (int l = 0; l <= 10; ++l) {
    total += width < n ? width : n;
    if (width < n) width *= 8;
  }
  return total;
}

int64_t or_build_tree(int64_t n, const uint32_t* codes, int leaf_size, int64_t max_nodes,
                      int32_t* first_child, int32_t* next, int32_t* start, int32_t* nbody,
                      int32_t* level) {
  if (n <= 0 || max_nodes < 1) return 0;
  first_child[0] = -1;
  next[0] = -1;
  start[0] = 0;
  nbody[0] = (int32_t)n;
  level[0] = 0;
  int64_t lo = 0, hi = 1, used = 1;
  for (int l = 0; l < 10; ++l) {
    const int shift = 27 - 3 * l; /* digit of level l+1 */
    for (int64_t k = lo; k < hi; ++k) {
      first_child[k] = -1;
      if (nbody[k] <= leaf_size) continue;
      const int64_t s = start[k], e = start[k] + nbody[k];
      first_child[k] = (int32_t)used;
      int64_t b = s;
      while (b < e) {
        const uint32_t d = (codes[b] >> shift) & 7u;
        int64_t c = b + 1;
        while (c < e && ((codes[c] >> shift) & 7u) == d) ++c;
        if (used >= max_nodes) return -1;
        start[used] = (int32_t)b;
        nbody[used] = (int32_t)(c - b);
        level[used] = l + 1;
        first_child[used] = -1;
        next[used] = c < e ? (int32_t)(used + 1) : next[k];
        ++used;
        b = c;
      }
    }
    lo = hi;
    hi = used;
  }
  return used;
}

void or_tree_com(int64_t nodes, const int32_t* start, const int32_t* nbody, int64_t n,
                 const float* pos4s, float* com4) {
  double* P = (double*)malloc(sizeof(double) * 4 * (size_t)(n + 1));
  P[0] = P[1] = P[2] = P[3] = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double m = pos4s[4 * i + 3];
    P[4 * (i + 1)] = P[4 * i] + m;
    P[4 * (i + 1) + 1] = P[4 * i + 1] + m * (double)pos4s[4 * i];
    P[4 * (i + 1) + 2] = P[4 * i + 2] + m * (double)pos4s[4 * i + 1];
    P[4 * (i + 1) + 3] = P[4 * i + 3] + m * (double)pos4s[4 * i + 2];
  }
  for (int64_t k = 0; k < nodes; ++k) {
    const int64_t s = start[k], e = start[k] + nbody[k];
    if (nbody[k] == 1) {
      for (int c = 0; c < 4; ++c) com4[4 * k + c] = pos4s[4 * s + c];
      continue;
    }
    const double m = P[4 * e] - P[4 * s];
    com4[4 * k + 3] = (float)m;
    for (int c = 0; c < 3; ++c)
      com4[4 * k + c] = m > 0 ? (float)((P[4 * e + 1 + c] - P[4 * s + 1 + c]) / m) : 0.f;
  }
  free(P);
}

static float root_side(const float bmin[3], const float bmax[3]) {
  float s = 0.f;
  for (int a = 0; a < 3; ++a) {
    float e = bmax[a] - bmin[a];
    e = e > 1e-6f ? e : 1e-6f;
    s = e > s ? e : s;
  }
  return s;
}

int64_t or_bh_forces(int64_t n, const float* pos4s, const uint32_t* perm, const float bmin[3],
                     const float bmax[3], int64_t nodes, const int32_t* first_child,
                     const int32_t* next, const int32_t* start, const int32_t* nbody,
                     const int32_t* level, const float* com4, float eps2, float G, float theta,
                     float* force4, int nthreads) {
  (void)nodes;
  float side2[11];
  const float s0 = root_side(bmin, bmax);
  for (int l = 0; l <= 10; ++l) {
    const float s = ldexpf(s0, -l);
    side2[l] = s * s;
  }
  const float theta2 = theta * theta;
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) num_threads(clamp_threads(nthreads))
  for (int64_t t = 0; t < n; ++t) {
    const float px = pos4s[4 * t], py = pos4s[4 * t + 1], pz = pos4s[4 * t + 2];
    float ax = 0.f, ay = 0.f, az = 0.f;
    int64_t cnt = 0;
    int32_t i = 0;
    while (i >= 0) {
      const float dx = com4[4 * i] - px, dy = com4[4 * i + 1] - py, dz = com4[4 * i + 2] - pz;
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (side2[level[i]] < theta2 * d2) {
        const float m = com4[4 * i + 3];
        if (m > 0.f && d2 > 0.f) {
          const float inv = 1.0f / sqrtf(d2 + eps2);
          const float s = m * inv * inv * inv;
          ax += dx * s;
          ay += dy * s;
          az += dz * s;
          ++cnt;
        }
        i = next[i];
      } else if (first_child[i] < 0) {
        const int64_t e = (int64_t)start[i] + nbody[i];
        for (int64_t b = start[i]; b < e; ++b) {
          const float bx = pos4s[4 * b] - px, by = pos4s[4 * b + 1] - py, bz = pos4s[4 * b + 2] - pz;
          const float b2 = bx * bx + by * by + bz * bz;
          if (b2 > 0.f) {
            const float inv = 1.0f / sqrtf(b2 + eps2);
            const float s = pos4s[4 * b + 3] * inv * inv * inv;
            ax += bx * s;
            ay += by * s;
            az += bz * s;
            ++cnt;
          }
        }
        i = next[i];
      } else {
        i = first_child[i];
      }
    }
    const int64_t o = perm[t];
    force4[4 * o] = G * ax;
    force4[4 * o + 1] = G * ay;
    force4[4 * o + 2] = G * az;
    force4[4 * o + 3] = 0.f;
    total += cnt;
  }
  return total;
}

int64_t or_bh_wave_ranges(int64_t n, int64_t nodes, const int32_t* first_child,
                          const int32_t* start, const int32_t* nbody, const int32_t* level,
                          int cap, int32_t* range_start, int32_t* range_count) {
  int64_t r = 0;
  if (n <= cap) {
    range_start[0] = 0;
    range_count[0] = (int32_t)n;
    return 1;
  }
  if (nodes > 0 && first_child[0] < 0) { /* the root is a leaf: chunks of it */
    for (int64_t o = 0; o < n; o += cap) {
      range_start[r] = (int32_t)o;
      range_count[r++] = (int32_t)(n - o < cap ? n - o : cap);
    }
    return r;
  }
  for (int64_t k = 0; k < nodes; ++k) {
    if (first_child[k] < 0 || nbody[k] <= cap) continue;
    const int64_t end = (int64_t)start[k] + nbody[k];
    int32_t run_s = 0, run_n = 0;
    for (int64_t c = first_child[k]; c < nodes && level[c] == level[k] + 1 && start[c] < end; ++c) {
      const int32_t cn = nbody[c];
      if (cn > cap) {
        if (run_n) { range_start[r] = run_s; range_count[r++] = run_n; }
        run_n = 0;
        if (first_child[c] < 0)  /* a leaf: chunks; an internal child packs its own children */
          for (int32_t o = 0; o < cn; o += cap) {
            range_start[r] = start[c] + o;
            range_count[r++] = cn - o < cap ? cn - o : cap;
          }
      } else {
        if (run_n + cn > cap) { range_start[r] = run_s; range_count[r++] = run_n; run_n = 0; }
        if (run_n == 0) run_s = start[c];
        run_n += cn;
      }
    }
    if (run_n) { range_start[r] = run_s; range_count[r++] = run_n; }
  }
  return r;
}

uint32_t or_compact_bits(uint32_t v) {
  /* inverse of or_expand_bits on 10-bit values: keep bits 0, 3, 6, ... */
  v &= 0x49249249u;
  v = (v ^ (v >> 2)) & 0xC30C30C3u;
  v = (v ^ (v >> 4)) & 0x0F00F00Fu;
  v = (v ^ (v >> 8)) & 0xFF0000FFu;
  v = (v ^ (v >> 16)) & 0x000003FFu;
  return v;
}

void or_tree_open_r2(int64_t nodes, const int32_t* start, const int32_t* level, const float* com4,
                     const uint32_t* codes_sorted, const float bmin[3], const float bmax[3],
                     int cubic, float theta, float* r2) {
  const float E = root_side(bmin, bmax);
  for (int64_t k = 0; k < nodes; ++k) {
    const float* c = &com4[4 * k];
    if (!(c[3] > 0.f)) {
      r2[k] = -1.f; /* massless: never opened, never listed */
      continue;
    }
    const int l = level[k];
    const uint32_t q = codes_sorted[start[k]] >> (30 - 3 * l); /* the cell's 3l-bit prefix */
    float d2 = 0.f;
    for (int a = 0; a < 3; ++a) {
      const float ext = cubic ? E : axis_extent(bmin[a], bmax[a]);
      const float h = ldexpf(ext, -l);
      const uint32_t idx = or_compact_bits(q >> (2 - a));
      const float ctr = bmin[a] + ((float)idx + 0.5f) * h;
      const float d = c[a] - ctr;
      d2 = a == 0 ? d * d : d2 + d * d;
    }
    const float rc = ldexpf(E, -l) / theta + sqrtf(d2);
    r2[k] = rc * rc;
  }
}

int64_t or_bh_forces_group(int64_t n, const float* pos4s, const uint32_t* perm,
                           const float bmin[3], const float bmax[3], int64_t nodes,
                           const int32_t* first_child, const int32_t* next, const int32_t* start,
                           const int32_t* nbody, const int32_t* level, const float* com4,
                           const uint32_t* codes_sorted, int cubic, float eps2, float G,
                           float theta, int group, float* force4, int nthreads) {
  if (group < 1 || group > 64) return -1;
  float* r2 = (float*)malloc(sizeof(float) * (size_t)(nodes ? nodes : 1));
  or_tree_open_r2(nodes, start, level, com4, codes_sorted, bmin, bmax, cubic, theta, r2);
  int32_t* rs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int32_t* rc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  const int64_t ngroups = or_bh_wave_ranges(n, nodes, first_child, start, nbody, level, group, rs, rc);
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total) num_threads(clamp_threads(nthreads))
  for (int64_t g = 0; g < ngroups; ++g) {
    const int64_t t0 = rs[g], t1 = (int64_t)rs[g] + rc[g];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t t = t0; t < t1; ++t)
      for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(lo[a], pos4s[4 * t + a]);
        hi[a] = fmaxf(hi[a], pos4s[4 * t + a]);
      }
    float acc[64][3];
    int64_t cnt = 0;
    memset(acc, 0, sizeof(acc));
    int32_t i = 0;
    while (i >= 0) {
      const float* c = &com4[4 * i];
      /* squared distance from the COM to the group's box */
      float e[3];
      for (int a = 0; a < 3; ++a) e[a] = fmaxf(fmaxf(lo[a] - c[a], c[a] - hi[a]), 0.f);
      const float dmin2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
      if (r2[i] < dmin2) {
        if (c[3] > 0.f) {
          for (int64_t t = t0; t < t1; ++t) {
            const float dx = c[0] - pos4s[4 * t], dy = c[1] - pos4s[4 * t + 1], dz = c[2] - pos4s[4 * t + 2];
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > 0.f) {
              const float inv = 1.0f / sqrtf(d2 + eps2);
              const float s = c[3] * inv * inv * inv;
              acc[t - t0][0] += dx * s;
              acc[t - t0][1] += dy * s;
              acc[t - t0][2] += dz * s;
              ++cnt;
            }
          }
        }
        i = next[i];
      } else if (first_child[i] < 0) {
        const int64_t e1 = (int64_t)start[i] + nbody[i];
        for (int64_t b = start[i]; b < e1; ++b)
          for (int64_t t = t0; t < t1; ++t) {
            const float bx = pos4s[4 * b] - pos4s[4 * t], by = pos4s[4 * b + 1] - pos4s[4 * t + 1];
            const float bz = pos4s[4 * b + 2] - pos4s[4 * t + 2];
            const float b2 = bx * bx + by * by + bz * bz;
            if (b2 > 0.f) {
              const float inv = 1.0f / sqrtf(b2 + eps2);
              const float s = pos4s[4 * b + 3] * inv * inv * inv;
              acc[t - t0][0] += bx * s;
              acc[t - t0][1] += by * s;
              acc[t - t0][2] += bz * s;
              ++cnt;
            }
          }
        i = next[i];
      } else {
        i = first_child[i];
      }
    }
    for (int64_t t = t0; t < t1; ++t) {
      const int64_t o = perm[t];
      force4[4 * o] = G * acc[t - t0][0];
      force4[4 * o + 1] = G * acc[t - t0][1];
      force4[4 * o + 2] = G * acc[t - t0][2];
      force4[4 * o + 3] = 0.f;
    }
    total += cnt;
  }
  free(rs);
  free(rc);
  free(r2);
  return total;
}

int64_t or_bh_pipeline_group(int64_t n, const float* pos4, float eps2, float G, float theta,
                             int leaf_size, int group, int cubic, float* force4, uint32_t* codes_out,
                             uint32_t* perm_out, int64_t* node_count, int nthreads) {
  float bmin[3], bmax[3];
  or_bbox(n, pos4, bmin, bmax);
  uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t* sorted = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
  or_morton(n, pos4, bmin, bmax, cubic, codes);
  or_sort_codes(n, codes, perm, sorted);
  float* pos4s = (float*)malloc(sizeof(float) * 4 * (size_t)n);
  for (int64_t k = 0; k < n; ++k)
    for (int c = 0; c < 4; ++c) pos4s[4 * k + c] = pos4[4 * (int64_t)perm[k] + c];
  const int64_t bound = or_tree_bound(n);
  int32_t* fc = (int32_t*)malloc(sizeof(int32_t) * (size_t)bound);
  int32_t* nx = (int32_t*)malloc(sizeof(int32_t) * (size_t)bound);
  int32_t* st = (int32_t*)malloc(sizeof(int32_t) * (size_t)bound);
  int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)bound);
  int32_t* lv = (int32_t*)malloc(sizeof(int32_t) * (size_t)bound);
  float* com = (float*)malloc(sizeof(float) * 4 * (size_t)bound);
  const int64_t nodes = or_build_tree(n, sorted, leaf_size, bound, fc, nx, st, nb, lv);
  or_tree_com(nodes, st, nb, n, pos4s, com);
  const int64_t inter =
      group > 0 ? or_bh_forces_group(n, pos4s, perm, bmin, bmax, nodes, fc, nx, st, nb, lv, com,
                                     sorted, cubic, eps2, G, theta, group, force4, nthreads)
                : or_bh_forces(n, pos4s, perm, bmin, bmax, nodes, fc, nx, st, nb, lv, com, eps2, G,
                               theta, force4, nthreads);
  if (codes_out) memcpy(codes_out, codes, sizeof(uint32_t) * (size_t)n);
  if (perm_out) memcpy(perm_out, perm, sizeof(uint32_t) * (size_t)n);
  if (node_count) *node_count = nodes;
  free(codes); free(sorted); free(perm); free(pos4s);
  free(fc); free(nx); free(st); free(nb); free(lv); free(com);
  return inter;
}

int64_t or_bh_pipeline(int64_t n, const float* pos4, float eps2, float G, float theta,
                       int leaf_size, float* force4, uint32_t* codes_out, uint32_t* perm_out,
                       int64_t* node_count, int nthreads) {
  return or_bh_pipeline_group(n, pos4, eps2, G, theta, leaf_size, 0, 0, force4, codes_out,
                              perm_out, node_count, nthreads);
}
