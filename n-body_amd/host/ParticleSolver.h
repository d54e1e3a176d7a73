// ParticleSolver.h — the reference's solver contract, GL-free
// (include/ParticleSolver.h:8-15): one call advances the system one full step
// in place.
#pragma once
#include "ParticleSystem.h"

class ParticleSolver {
 public:
  virtual ~ParticleSolver() = default;
  ParticleSolver() = default;
  virtual void updateParticlePositions(ParticleSystem* particles) = 0;
  virtual float getSquaredSoftening() = 0;
  virtual bool usesGPU() = 0;
};
