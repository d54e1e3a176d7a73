// ParticleSystem.cpp — see ParticleSystem.h.
#include "ParticleSystem.h"

#include <iomanip>
#include <limits>

#include "nbody.h"

ParticleSystem::ParticleSystem(const std::vector<Particle>& particles, bool usesGPU)
    : uses_gpu_(usesGPU) {
  const size_t n = particles.size();
  positions_.resize(n);
  velocities_.resize(n);
  accelerations_.resize(n);
  masses_.resize(n);
  forces_.resize(n);
  for (size_t i = 0; i < n; ++i) {
    positions_[i] = particles[i].position;
    positions_[i].w = 0.f;
    velocities_[i] = particles[i].velocity;
    velocities_[i].w = 0.f;
    accelerations_[i] = particles[i].acceleration;
    accelerations_[i].w = 0.f;
    masses_[i] = Vec4{particles[i].mass, 0.f, 0.f, 0.f};  // ParticleSystem.cpp:18
  }
}

int ParticleSystem::pullFromDevice() {
  if (!device_) return NB_OK;
  int rc = nb_download(device_, &positions_[0].x, &velocities_[0].x, &accelerations_[0].x);
  if (rc != NB_OK) return rc;
  // the last force pass is the acceleration the integrator stored
  forces_ = accelerations_;
  return NB_OK;
}

std::ostream& operator<<(std::ostream& os, const ParticleSystem& s) {
  const auto old = os.precision();
  os << std::setprecision(std::numeric_limits<float>::max_digits10);
  os << "Particle System with " << s.size() << " particles:" << std::endl;
  for (unsigned int i = 0; i < s.size(); ++i) {
    const Vec4 &p = s.positions_[i], &v = s.velocities_[i], &a = s.accelerations_[i];
    os << "Particle ID: " << i << std::endl;
    os << "Position: (" << p.x << ", " << p.y << ", " << p.z << ")" << std::endl;
    os << "Velocity: (" << v.x << ", " << v.y << ", " << v.z << ")" << std::endl;
    os << "Acceleration: (" << a.x << ", " << a.y << ", " << a.z << ")" << std::endl;
    os << "Mass: " << s.masses_[i].x << std::endl;
  }
  os.precision(old);
  return os;
}
