// ParticleSystem.h — GL-free particle state with the reference's layout
// (include/ParticleSystem.h:80-86): five vec4 arrays per body (position,
// velocity, acceleration, mass in .x, force), w = 0.
//
// For GPU solvers the authoritative state lives in the solver's device context
// (nb_ctx, HBM-resident) and the host arrays are refreshed only by
// pullFromDevice(), as in the reference where GPU solvers never read back
// (ParticleSolverGPU.cpp:31-41). CPU-side edits are pushed with markHostDirty().
#pragma once
#include <cstddef>
#include <ostream>
#include <vector>

struct nb_ctx;

struct Vec4 {
  float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
};

// Initial-condition record (the reference's Particle, src/Particle/Particle.cpp:4-9).
struct Particle {
  Vec4 position, velocity, acceleration;
  float mass = 0.f;
};

class ParticleSystem {
 public:
  ParticleSystem(const std::vector<Particle>& particles, bool usesGPU);

  unsigned int size() const { return (unsigned int)positions_.size(); }
  Vec4* getPositions() { return positions_.data(); }
  Vec4* getVelocities() { return velocities_.data(); }
  Vec4* getAccelerations() { return accelerations_.data(); }
  Vec4* getMasses() { return masses_.data(); }
  Vec4* getForces() { return forces_.data(); }
  bool usesGPU() const { return uses_gpu_; }

  // Device attachment (used by the HIP solvers).
  nb_ctx* device() const { return device_; }
  void attachDevice(nb_ctx* ctx) { device_ = ctx; host_dirty_ = true; }
  bool hostDirty() const { return host_dirty_; }
  void markHostDirty() { host_dirty_ = true; }
  void clearHostDirty() { host_dirty_ = false; }
  // Copy the device state (positions, velocities, accelerations, last forces)
  // into the host arrays; returns 0 or an NB_E_* code.
  int pullFromDevice();

  // Full-precision text dump in the reference's format
  // (ParticleSystem.cpp:205-215 writes 6 significant digits; this writes 9,
  // which round-trips every float exactly).
  friend std::ostream& operator<<(std::ostream& os, const ParticleSystem& s);

 private:
  std::vector<Vec4> positions_, velocities_, accelerations_, masses_, forces_;
  bool uses_gpu_;
  nb_ctx* device_ = nullptr;
  bool host_dirty_ = true;
};
