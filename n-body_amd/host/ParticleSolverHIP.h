// ParticleSolverHIP.h — the MI355X drop-ins behind the reference's solver
// contract:
//   ParticleSolverHIP      replaces ParticleSolverGPU with the tiled v4 shader
//                          (include/ParticleSolverGPU.h:11-12; main.cpp:106-117)
//   ParticleSolverBHutHIP  replaces ParticleSolverBHutGPU (v7)
//                          (include/ParticleSolverBHutGPU.h:11; main.cpp:141-152)
// Both drive libnbody through the C ABI of nbody.h. The device context is
// created on the first updateParticlePositions() call, when the system's size
// is known (the reference creates the system after the solver,
// ParticleSimulation.cpp:11-12). Errors are thrown as std::runtime_error with
// the engine's message (the reference's solvers return void).
#pragma once
#include <string>

#include "ParticleSolver.h"
#include "nbody.h"

class ParticleSolverHIPBase : public ParticleSolver {
 public:
  ~ParticleSolverHIPBase() override;
  void updateParticlePositions(ParticleSystem* particles) override;
  float getSquaredSoftening() override { return cfg_.eps2; }
  bool usesGPU() override { return true; }
  // Advance k steps in one call (captured into a hipGraph for k > 1).
  void updateParticlePositions(ParticleSystem* particles, int steps);
  nb_ctx* context() const { return ctx_; }

 protected:
  ParticleSolverHIPBase(const nb_config& cfg) : cfg_(cfg) {}
  void ensure(ParticleSystem* particles);
  nb_config cfg_;
  nb_ctx* ctx_ = nullptr;
};

class ParticleSolverHIP : public ParticleSolverHIPBase {
 public:
  // stepSize and squaredSoft as ParticleSolverGPU(float, float, ...); the shader
  // paths of the reference constructor have no counterpart (kernels are compiled
  // ahead of time). bodiesPerLane / jSplit = 0 use the tuned launch shape.
  ParticleSolverHIP(float stepSize, float squaredSoft, int bodiesPerLane = 0, int jSplit = 0,
                    int integrator = NB_INTEG_KDK_LAGGED);
};

class ParticleSolverBHutHIP : public ParticleSolverHIPBase {
 public:
  // As ParticleSolverBHutGPU(float timeStep, float squaredSoftening, int n, ...)
  // with the opening angle (unsquared) and bodies per leaf made explicit.
  ParticleSolverBHutHIP(float stepSize, float squaredSoft, int n, float theta = 0.5f,
                        int leafSize = 0, int integrator = NB_INTEG_KDK_LAGGED);
};
