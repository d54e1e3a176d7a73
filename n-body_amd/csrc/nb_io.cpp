// nb_io.cpp — the reference's particle text format (host only); see nbody_ic.h.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <limits>
#include <new>
#include <sstream>
#include <string>

#include "nbody.h"
#include "nbody_ic.h"

namespace {

// Parse "...(a, b, c)..." the way ParticleSystemFile.cpp:43-50 does: skip to
// '(', then read up to ',', ',' and ')' and convert each with std::stof.
bool parse_triplet(const std::string& line, float out[3]) {
  std::istringstream iss(line);
  iss.ignore(std::numeric_limits<std::streamsize>::max(), '(');
  std::string a, b, c;
  std::getline(iss, a, ',');
  std::getline(iss, b, ',');
  std::getline(iss, c, ')');
  try {
    out[0] = std::stof(a);
    out[1] = std::stof(b);
    out[2] = std::stof(c);
  } catch (...) {
    return false;
  }
  return true;
}

}  // namespace

namespace {

int64_t read_file(const char* path, int64_t cap, float* pos4, float* vel4, float* acc4,
                  float* mass) {
  if (!path) return NB_E_ARG;
  std::ifstream file(path);
  if (!file) return NB_E_ARG;
  std::string line;
  std::getline(file, line);
  const auto h = line.find("Particle System with ");
  if (h == std::string::npos) return NB_E_ARG;
  int64_t total = 0;
  try {
    total = std::stoll(line.substr(line.find("with ") + 5));
  } catch (...) {
    return NB_E_ARG;
  }
  float p[3] = {0, 0, 0}, v[3] = {0, 0, 0}, a[3] = {0, 0, 0};
  int64_t count = 0;
  while (count < total && std::getline(file, line)) {
    if (line.find("Particle ID") != std::string::npos) continue;
    if (line.find("Position:") != std::string::npos) {
      if (!parse_triplet(line, p)) return NB_E_ARG;
    } else if (line.find("Velocity:") != std::string::npos) {
      if (!parse_triplet(line, v)) return NB_E_ARG;
    } else if (line.find("Acceleration:") != std::string::npos) {
      if (!parse_triplet(line, a)) return NB_E_ARG;
    } else if (line.find("Mass:") != std::string::npos) {
      float m;
      try {
        m = std::stof(line.substr(line.find(':') + 1));
      } catch (...) {
        return NB_E_ARG;
      }
      if (count < cap) {
        for (int k = 0; k < 3; ++k) {
          if (pos4) pos4[4 * count + k] = p[k];
          if (vel4) vel4[4 * count + k] = v[k];
          if (acc4) acc4[4 * count + k] = a[k];
        }
        if (pos4) pos4[4 * count + 3] = 0.f;
        if (vel4) vel4[4 * count + 3] = 0.f;
        if (acc4) acc4[4 * count + 3] = 0.f;
        if (mass) mass[count] = m;
      }
      ++count;
    }
  }
  return count;
}

int write_file(const char* path, int64_t n, const float* pos4, const float* vel4,
               const float* acc4, const float* mass) {
  if (!path || n < 0 || !pos4 || !vel4 || !mass) return NB_E_ARG;
  std::ofstream os(path);
  if (!os) return NB_E_ARG;
  os << std::setprecision(std::numeric_limits<float>::max_digits10);
  os << "Particle System with " << n << " particles:\n";
  for (int64_t i = 0; i < n; ++i) {
    const float* p = pos4 + 4 * i;
    const float* v = vel4 + 4 * i;
    os << "Particle ID: " << i << "\n";
    os << "Position: (" << p[0] << ", " << p[1] << ", " << p[2] << ")\n";
    os << "Velocity: (" << v[0] << ", " << v[1] << ", " << v[2] << ")\n";
    if (acc4) {
      const float* a = acc4 + 4 * i;
      os << "Acceleration: (" << a[0] << ", " << a[1] << ", " << a[2] << ")\n";
    } else {
      os << "Acceleration: (0, 0, 0)\n";
    }
    os << "Mass: " << mass[i] << "\n";
  }
  return os.good() ? NB_OK : NB_E_ARG;
}

}  // namespace

// The C entry points: no C++ exception crosses them (a failed host allocation
// while reading or writing becomes NB_E_OOM).
extern "C" int64_t nb_ic_read_file(const char* path, int64_t cap, float* pos4, float* vel4,
                                   float* acc4, float* mass) {
  try {
    return read_file(path, cap, pos4, vel4, acc4, mass);
  } catch (const std::bad_alloc&) {
    return NB_E_OOM;
  } catch (...) {
    return NB_E_ARG;
  }
}

extern "C" int nb_ic_write_file(const char* path, int64_t n, const float* pos4, const float* vel4,
                                const float* acc4, const float* mass) {
  try {
    return write_file(path, n, pos4, vel4, acc4, mass);
  } catch (const std::bad_alloc&) {
    return NB_E_OOM;
  } catch (...) {
    return NB_E_ARG;
  }
}
