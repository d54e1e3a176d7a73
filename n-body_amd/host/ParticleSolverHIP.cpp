// ParticleSolverHIP.cpp — see ParticleSolverHIP.h.
#include "ParticleSolverHIP.h"

#include <stdexcept>
#include <vector>

namespace {
void check(int rc, const char* what) {
  if (rc != NB_OK) throw std::runtime_error(std::string(what) + ": " + nb_last_error());
}
}  // namespace

ParticleSolverHIPBase::~ParticleSolverHIPBase() {
  if (ctx_) nb_destroy(ctx_);
}

void ParticleSolverHIPBase::ensure(ParticleSystem* ps) {
  if (!ps) throw std::invalid_argument("updateParticlePositions: null particle system");
  if (ctx_ && cfg_.n != (int64_t)ps->size()) {
    nb_destroy(ctx_);
    ctx_ = nullptr;
  }
  if (!ctx_) {
    cfg_.n = ps->size();
    check(nb_create(&ctx_, &cfg_), "nb_create");
  }
  if (ps->device() != ctx_ || ps->hostDirty()) {
    const unsigned n = ps->size();
    std::vector<float> mass(n);
    for (unsigned i = 0; i < n; ++i) mass[i] = ps->getMasses()[i].x;
    check(nb_upload(ctx_, &ps->getPositions()[0].x, &ps->getVelocities()[0].x,
                    &ps->getAccelerations()[0].x, mass.data(), n),
          "nb_upload");
    ps->attachDevice(ctx_);
    ps->clearHostDirty();
  }
}

void ParticleSolverHIPBase::updateParticlePositions(ParticleSystem* ps) {
  updateParticlePositions(ps, 1);
}

void ParticleSolverHIPBase::updateParticlePositions(ParticleSystem* ps, int steps) {
  ensure(ps);
  check(nb_step(ctx_, steps), "nb_step");
}

static nb_config base_config(float dt, float eps2, int solver, int integrator) {
  nb_config c{};
  c.eps2 = eps2;
  c.dt = dt;
  c.G = 1.0f;  // ParticleSolverCPUSequential.cpp:11, ParticleSolverBHutGPU.cpp:13
  c.solver = solver;
  c.integrator = integrator;
  return c;
}

ParticleSolverHIP::ParticleSolverHIP(float stepSize, float squaredSoft, int bodiesPerLane,
                                     int jSplit, int integrator)
    : ParticleSolverHIPBase(base_config(stepSize, squaredSoft, NB_SOLVER_DIRECT, integrator)) {
  cfg_.bodies_per_lane = bodiesPerLane;
  cfg_.j_split = jSplit;
}

ParticleSolverBHutHIP::ParticleSolverBHutHIP(float stepSize, float squaredSoft, int n, float theta,
                                             int leafSize, int integrator)
    : ParticleSolverHIPBase(base_config(stepSize, squaredSoft, NB_SOLVER_BH, integrator)) {
  cfg_.n = n;
  cfg_.theta = theta;  // the reference hard-codes 0.5 (ParticleSolverBHutGPU.h:19)
  cfg_.leaf_size = leafSize;
}
