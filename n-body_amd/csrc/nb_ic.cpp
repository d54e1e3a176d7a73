// nb_ic.cpp — seeded initial conditions (host only); see include/nbody_ic.h.
#include "nbody_ic.h"

#include <cmath>
#include <cstdint>

namespace {

constexpr double kPi = 3.14159265358979323846;

// SplitMix64 (Steele, Lea & Flood 2014): a full-period 64-bit generator.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in (0, 1): 53 random bits, offset by half an ulp so 0 never appears
  double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  // standard normal deviate (Box & Muller 1958), one per call
  double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(2.0 * kPi * uniform()); }
};

// Velocity and mass of the cube and cube-surface initializers: each component
// uniform in [-L, L] with L = |(w, w, w)| / 20, and mass (w^3 / n) * 1.2 * U(0.8, 1.3).
void cube_velocity_mass(Rng& r, double w, int64_t n, double* vel3, float* mass) {
  const double len = std::sqrt(3.0) * w / 20.0;
  for (int k = 0; k < 3; ++k) vel3[k] = r.uniform(-len, len);
  *mass = (float)(w * w * w / (double)n * 1.2 * r.uniform(0.8, 1.3));
}

void isotropic(Rng& r, double len, double out[3]) {
  const double z = 2.0 * r.uniform() - 1.0;
  const double phi = 2.0 * kPi * r.uniform();
  const double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
  out[0] = len * s * std::cos(phi);
  out[1] = len * s * std::sin(phi);
  out[2] = len * z;
}

}  // namespace

extern "C" {

int nb_ic_ball(uint64_t seed, int64_t n, double radius, double mass_each, double* pos3,
               double* vel3, float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(radius > 0)) return -1;
  Rng r(seed);
  for (int64_t i = 0; i < n; ++i) {
    const double rad = radius * std::cbrt(r.uniform());
    isotropic(r, rad, &pos3[3 * i]);
    vel3[3 * i] = vel3[3 * i + 1] = vel3[3 * i + 2] = 0.0;
    mass[i] = (float)mass_each;
  }
  return 0;
}

int nb_ic_plummer(uint64_t seed, int64_t n, double a, double total_mass, double rmax_over_a,
                  double G, double* pos3, double* vel3, float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(a > 0) || !(total_mass > 0) || !(rmax_over_a > 0))
    return -1;
  Rng r(seed);
  const double m = total_mass / (double)n;
  const double rmax = rmax_over_a * a;
  double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    double rad;
    do {
      const double u = r.uniform();
      rad = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
    } while (!(rad <= rmax));
    isotropic(r, rad, &pos3[3 * i]);
    double q, g;
    do {  // g(q) = q^2 (1 - q^2)^3.5 <= 0.1 on [0, 1]
      q = r.uniform();
      g = 0.1 * r.uniform();
    } while (g > q * q * std::pow(1.0 - q * q, 3.5));
    const double vesc = std::sqrt(2.0 * G * total_mass / a) * std::pow(1.0 + rad * rad / (a * a), -0.25);
    isotropic(r, q * vesc, &vel3[3 * i]);
    mass[i] = (float)m;
    for (int k = 0; k < 3; ++k) {
      cp[k] += pos3[3 * i + k];
      cv[k] += vel3[3 * i + k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    cp[k] /= (double)n;
    cv[k] /= (double)n;
  }
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      pos3[3 * i + k] -= cp[k];
      vel3[3 * i + k] -= cv[k];
    }
  return 0;
}

int nb_ic_discs(uint64_t seed, int64_t n, double scale_length, double scale_height,
                double central_factor, double total_mass, double separation,
                double closing_speed, double eps2, double G, double* pos3, double* vel3,
                float* mass) {
  if (n < 4 || (n & 1) || !pos3 || !vel3 || !mass || !(scale_length > 0) ||
      !(scale_height > 0) || !(total_mass > 0))
    return -1;
  Rng r(seed);
  const int64_t half = n / 2;
  const double mb = total_mass / (double)n;
  const double mc = central_factor * mb;
  const double md = (double)(half - 1) * mb;  // disc mass excluding the central body
  for (int disc = 0; disc < 2; ++disc) {
    const double cx = disc == 0 ? -0.5 * separation : 0.5 * separation;
    const double bvx = disc == 0 ? 0.5 * closing_speed : -0.5 * closing_speed;
    const int64_t b0 = disc * half;
    pos3[3 * b0] = cx;
    pos3[3 * b0 + 1] = 0.0;
    pos3[3 * b0 + 2] = 0.0;
    vel3[3 * b0] = bvx;
    vel3[3 * b0 + 1] = 0.0;
    vel3[3 * b0 + 2] = 0.0;
    mass[b0] = (float)mc;
    for (int64_t k = 1; k < half; ++k) {
      const int64_t i = b0 + k;
      const double R = -scale_length * std::log(r.uniform() * r.uniform());
      const double phi = 2.0 * kPi * r.uniform();
      const double z = scale_height * std::atanh(2.0 * r.uniform() - 1.0);
      const double x = R * std::cos(phi), y = R * std::sin(phi);
      const double t = R / scale_length;
      const double menc = mc + md * (1.0 - (1.0 + t) * std::exp(-t));
      const double vc = std::sqrt(G * menc * R * R / std::pow(R * R + eps2, 1.5));
      pos3[3 * i] = cx + x;
      pos3[3 * i + 1] = y;
      pos3[3 * i + 2] = z;
      const double inv = R > 0 ? 1.0 / R : 0.0;
      vel3[3 * i] = bvx - vc * y * inv;
      vel3[3 * i + 1] = vc * x * inv;
      vel3[3 * i + 2] = 0.0;
      mass[i] = (float)mb;
    }
  }
  return 0;
}

int nb_ic_lagrange(double world_size, double* pos3, double* vel3, float* mass) {
  if (!pos3 || !vel3 || !mass || !(world_size > 0)) return -1;
  // ParticleSystemLagrange.cpp:14-48, single precision as in the reference.
  const float w = (float)world_size;
  const float m = 10.5f;
  const float side = std::sqrt(3.f * w * w) / 10.0f;  // glm::length(vec3(w)) / 10
  const float c[3] = {w / 2.0f, w / 2.0f, w / 2.0f};
  const float s3 = std::sqrt(3.0f);
  const float v[3][3] = {{c[0], c[1] + side / s3, c[2]},
                         {c[0] - side / 2.0f, c[1] - side / (2.0f * s3), c[2]},
                         {c[0] + side / 2.0f, c[1] - side / (2.0f * s3), c[2]}};
  float com[3];
  for (int k = 0; k < 3; ++k) com[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0f;
  const float total = m * 3.0f;
  for (int b = 0; b < 3; ++b) {
    const float d[3] = {v[b][0] - com[0], v[b][1] - com[1], v[b][2] - com[2]};
    const float dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const float speed = std::sqrt(1.f * total / dist);
    // normalize(cross(d, (0,0,1))) = (d.y, -d.x, 0) / |(d.y, -d.x)|
    const float cx = d[1], cy = -d[0];
    const float cl = std::sqrt(cx * cx + cy * cy);
    for (int k = 0; k < 3; ++k) pos3[3 * b + k] = v[b][k];
    vel3[3 * b] = cx / cl * speed * 0.5f;
    vel3[3 * b + 1] = cy / cl * speed * 0.5f;
    vel3[3 * b + 2] = 0.0;
    mass[b] = m;
  }
  return 0;
}

int nb_ic_cube(uint64_t seed, int64_t n, double world_size, double* pos3, double* vel3,
               float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(world_size > 0)) return -1;
  // ParticleSystemCubeInitializer.cpp:12-26 (the 3-D branch of line 23-24).
  Rng r(seed);
  const double w = world_size;
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) pos3[3 * i + k] = r.uniform(0.0, w);
    cube_velocity_mass(r, w, n, &vel3[3 * i], &mass[i]);
  }
  return 0;
}

int nb_ic_cube_surface(uint64_t seed, int64_t n, double world_size, double* pos3, double* vel3,
                       float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(world_size > 0)) return -1;
  // ParticleSystemCubeSurface.cpp:20-45: a face chosen uniformly, the other two
  // coordinates uniform over it, the cube [0, w]^3.
  Rng r(seed);
  const double w = world_size;
  for (int64_t i = 0; i < n; ++i) {
    const int face = (int)(6.0 * r.uniform());  // 0..5: -x, +x, -y, +y, -z, +z
    const int axis = face >> 1;
    for (int k = 0; k < 3; ++k)
      pos3[3 * i + k] = k == axis ? ((face & 1) ? w : 0.0) : r.uniform(0.0, w);
    cube_velocity_mass(r, w, n, &vel3[3 * i], &mass[i]);
  }
  return 0;
}

int nb_ic_sphere_surface(uint64_t seed, int64_t n, double world_size, int pattern, double* pos3,
                         double* vel3, float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(world_size > 0) || pattern < -1 || pattern > 2)
    return -1;
  // ParticleSystemSphere.cpp:28-130.
  Rng r(seed);
  const double w = world_size;
  const double diag = std::sqrt(3.0) * w;
  const double radius = diag / 3.0, c = 0.5 * w;
  const double base_mass = w * w * w / (double)n * 0.5;
  const double vbase = diag / 15.0;
  if (pattern < 0) pattern = (int)(3.0 * r.uniform());  // one pattern for the whole system
  const double an = std::sqrt(0.2 * 0.2 + 1.0 + 0.1 * 0.1);
  const double axis[3] = {0.2 / an, 1.0 / an, -0.1 / an};
  for (int64_t i = 0; i < n; ++i) {
    double d[3];
    isotropic(r, 1.0, d);  // uniform direction: the surface point is c + radius * d
    for (int k = 0; k < 3; ++k) pos3[3 * i + k] = c + radius * d[k];
    mass[i] = (float)(base_mass * r.uniform(0.7, 1.3));
    const double speed = vbase * r.uniform(0.5, 1.5);
    double* v = &vel3[3 * i];
    if (pattern == 0 || pattern == 1) {  // outward explosion / inward implosion
      const double s = pattern == 0 ? speed : -speed;
      for (int k = 0; k < 3; ++k) v[k] = s * d[k];
    } else {  // swirl about the fixed axis: direction axis x d
      const double t[3] = {axis[1] * d[2] - axis[2] * d[1], axis[2] * d[0] - axis[0] * d[2],
                           axis[0] * d[1] - axis[1] * d[0]};
      const double tl = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      if (tl * radius > 1e-6) {
        for (int k = 0; k < 3; ++k) v[k] = speed * t[k] / tl;
      } else {  // on the axis: a small random velocity (lines 106-115)
        double rd[3];
        isotropic(r, 0.1 * speed, rd);
        for (int k = 0; k < 3; ++k) v[k] = rd[k];
      }
    }
  }
  return 0;
}

int nb_ic_ball_world(uint64_t seed, int64_t n, double world_size, double* pos3, double* vel3,
                     float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(world_size > 0)) return -1;
  // ParticleSystemBall.cpp:19-89: a uniform ball by rejection from its bounding
  // cube, each body nudged by up to 5% of the radius per axis, small uniform
  // velocities, masses skewed towards the light end by u^2.
  Rng r(seed);
  const double w = world_size;
  const double diag = std::sqrt(3.0) * w;
  const double radius = diag / 3.0, c = 0.5 * w, nudge = 0.05 * radius;
  const double vscale = 0.001 * diag;
  const double avg = w * w * w / (double)n;
  for (int64_t i = 0; i < n; ++i) {
    double d[3], r2;
    do {
      for (int k = 0; k < 3; ++k) d[k] = r.uniform(-1.0, 1.0) * radius;
      r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    } while (r2 > radius * radius);
    for (int k = 0; k < 3; ++k) pos3[3 * i + k] = c + d[k] + r.uniform(-1.0, 1.0) * nudge;
    for (int k = 0; k < 3; ++k) vel3[3 * i + k] = r.uniform(-1.0, 1.0) * vscale;
    const double u = r.uniform();
    mass[i] = (float)(avg * (0.5 + 4.5 * u * u));
  }
  return 0;
}

int nb_ic_galaxy(uint64_t seed, int64_t n, double world_size, int style, double* pos3,
                 double* vel3, float* mass) {
  if (n < 1 || !pos3 || !vel3 || !mass || !(world_size > 0) || style < -1 || style > 1)
    return -1;
  Rng r(seed);
  const double w = world_size;
  const double c = 0.5 * w;
  const double vol = w * w * w;
  if (style < 0) style = r.uniform() < 0.5 ? 0 : 1;  // the 50/50 choice of line 21-22
  if (style == 0) {
    // ParticleSystemGalaxyInitializer.cpp:46-91: radius uniform in [0, w), angle
    // 2*pi*N(0, 50*pi), height U(0.1, w/5) above the centre plane, tangential speed
    // |(w, w, w)|/4 with a six-fold arm term inside w/2.
    const double len = std::sqrt(3.0) * w / 4.0;
    const double bulge = w / 10.0, arm = w / 2.0;
    for (int64_t i = 0; i < n; ++i) {
      const double rad = r.uniform(0.0, w);
      const double th = r.normal() * (3.1415 * 50.0) * 2.0 * kPi;
      const double h = r.uniform(0.1, w / 5.0);
      const double cs = std::cos(th), sn = std::sin(th);
      double* p = &pos3[3 * i];
      double* v = &vel3[3 * i];
      p[0] = c + rad * cs;
      p[1] = c + rad * sn;
      p[2] = c + h;
      mass[i] = (float)(vol / (double)n * (rad < bulge ? 0.7 : 0.5));
      v[0] = -len * sn;
      v[1] = len * cs;
      v[2] = 0.0;
      if (rad < arm) {
        const double ang = std::atan2(p[1] - c, p[0] - c);
        const double a = len * (rad / arm) * std::cos(6.0 * (ang - th));
        v[0] -= a * cs;
        v[1] -= a * sn;
      }
    }
    return 0;
  }
  // ParticleSystemGalaxyInitializer.cpp:92-205: structure drawn once per system,
  // then per body a centre-concentrated radius, a Gaussian layer, a capped
  // circular speed, a logarithmic-spiral radial perturbation and dispersion.
  const bool stable = r.uniform() < 0.7;
  const double bulge = w * r.uniform(0.05, 0.15);
  const double disc = w * r.uniform(0.3, 0.5);
  const double thick = disc * r.uniform(0.02, 0.10);
  const double arms = (double)(2 + (int)(5.0 * r.uniform()));  // 2..6
  const double tight = r.uniform(0.5, 2.0);
  const double vs_lo = stable ? 0.5 : 0.8, vs_hi = stable ? 0.9 : 1.5;
  const double as_lo = stable ? 0.05 : 0.1, as_hi = stable ? 0.2 : 0.5;
  const double disp_f = stable ? 0.05 : 0.15;
  const double vbase = 0.5 * disc * r.uniform(vs_lo, vs_hi);
  const double sigma = vbase * disp_f;
  const double base_mass = vol / (double)n;
  for (int64_t i = 0; i < n; ++i) {
    const double rad = disc * std::sqrt(r.uniform());
    const double th = r.uniform(0.0, 2.0 * kPi);
    const double h = r.normal() * thick / 3.0;
    const double cs = std::cos(th), sn = std::sin(th);
    double* p = &pos3[3 * i];
    double* v = &vel3[3 * i];
    p[0] = c + rad * cs;
    p[1] = c + rad * sn;
    p[2] = c + h;
    mass[i] = (float)(base_mass * (rad < bulge ? 1.5 : 1.0) * r.uniform(0.8, 1.2));
    const double vmag = vbase * std::fmin(rad / (bulge + 1e-6), 1.5);
    const double strength = r.uniform(as_lo, as_hi);
    const double phase = tight * std::log(std::fmax(rad / (bulge + 1e-6), 1.0));
    const double pert = strength * vmag * std::cos(arms * (th - phase));
    const double radial = rad > 1e-6 ? 1.0 : 0.0;  // (cos, sin) is the unit radial direction
    v[0] = -vmag * sn + radial * pert * cs + sigma * r.normal();
    v[1] = vmag * cs + radial * pert * sn + sigma * r.normal();
    v[2] = 0.5 * sigma * r.normal();
  }
  return 0;
}

int nb_ic_log_uniform_sizes(uint64_t seed, int32_t count, int32_t lo, int32_t hi, int32_t* out) {
  if (count < 0 || lo < 1 || hi < lo || !out) return -1;
  Rng r(seed);
  const double a = std::log((double)lo), b = std::log((double)hi);
  for (int32_t i = 0; i < count; ++i) {
    int64_t v = (int64_t)std::floor(std::exp(a + (b - a) * r.uniform()));
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    out[i] = (int32_t)v;
  }
  return 0;
}

int nb_ic_uniform(uint64_t seed, int32_t count, double lo, double hi, double* out) {
  if (count < 0 || !out) return -1;
  Rng r(seed);
  for (int32_t i = 0; i < count; ++i) out[i] = lo + (hi - lo) * r.uniform();
  return 0;
}

}  // extern "C"
