// sanitize_driver.cpp — runs the host-only C code (the seeded initial
// conditions, the particle-file reader and writer, and the CPU oracle) on small
// and edge-case inputs, so that a build with AddressSanitizer and
// UndefinedBehaviorSanitizer checks every access and every arithmetic step.
// Built and run by tests/test_sanitizers.py; exits non-zero on the first wrong
// status code (a sanitizer report aborts the process by itself).
//
// usage: sanitize_driver <scratch directory>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "nbody.h"
#include "nbody_ic.h"
#include "oracle.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

struct Bodies {
  std::vector<double> pos, vel;
  std::vector<float> mass;
  explicit Bodies(int64_t n) : pos(3 * n), vel(3 * n), mass(n) {}
};

// FP32 vec4 views of an FP64 system (w = 0; mass in .x of mass4)
void to_vec4(const Bodies& b, int64_t n, std::vector<float>& p4, std::vector<float>& v4,
             std::vector<float>& m4) {
  p4.assign(4 * n, 0.f);
  v4.assign(4 * n, 0.f);
  m4.assign(4 * n, 0.f);
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      p4[4 * i + k] = (float)b.pos[3 * i + k];
      v4[4 * i + k] = (float)b.vel[3 * i + k];
    }
    m4[4 * i] = b.mass[i];
  }
}

void initial_conditions() {
  for (int64_t n : {1, 2, 7, 300}) {
    Bodies b(n);
    expect(nb_ic_ball(1, n, 10.0, 1.0, b.pos.data(), b.vel.data(), b.mass.data()) == 0, "ball");
    expect(nb_ic_plummer(2, n, 1.0, 1.0, 10.0, 1.0, b.pos.data(), b.vel.data(), b.mass.data()) == 0,
           "plummer");
    expect(nb_ic_cube(3, n, 5.0, b.pos.data(), b.vel.data(), b.mass.data()) == 0, "cube");
    expect(nb_ic_cube_surface(4, n, 5.0, b.pos.data(), b.vel.data(), b.mass.data()) == 0,
           "cube surface");
    expect(nb_ic_ball_world(5, n, 5.0, b.pos.data(), b.vel.data(), b.mass.data()) == 0,
           "ball (world)");
    for (int pattern = -1; pattern <= 2; ++pattern)
      expect(nb_ic_sphere_surface(6, n, 5.0, pattern, b.pos.data(), b.vel.data(),
                                  b.mass.data()) == 0,
             "sphere surface");
    for (int style = -1; style <= 1; ++style)
      expect(nb_ic_galaxy(7, n, 5.0, style, b.pos.data(), b.vel.data(), b.mass.data()) == 0,
             "galaxy");
    if (n >= 4 && n % 2 == 0)
      expect(nb_ic_discs(8, n, 30.0, 2.0, 1e4, 1e4, 300.0, 0.5, 40.0, 1.0, b.pos.data(),
                         b.vel.data(), b.mass.data()) == 0,
             "discs");
    for (int64_t i = 0; i < 3 * n; ++i)
      expect(std::isfinite(b.pos[i]) && std::isfinite(b.vel[i]), "finite initial conditions");
  }
  Bodies b(300);
  double* p = b.pos.data();
  double* v = b.vel.data();
  float* m = b.mass.data();
  // rejected arguments
  expect(nb_ic_ball(1, 0, 10.0, 1.0, p, v, m) == NB_E_ARG, "ball n = 0");
  expect(nb_ic_cube(1, 10, -1.0, p, v, m) == NB_E_ARG, "cube negative world");
  expect(nb_ic_cube(1, 10, 5.0, nullptr, v, m) == NB_E_ARG, "cube null");
  expect(nb_ic_sphere_surface(1, 10, 5.0, 3, p, v, m) == NB_E_ARG, "sphere bad pattern");
  expect(nb_ic_galaxy(1, 10, 5.0, 2, p, v, m) == NB_E_ARG, "galaxy bad style");
  expect(nb_ic_discs(1, 7, 30.0, 2.0, 1e4, 1e4, 300.0, 0.5, 40.0, 1.0, p, v, m) == NB_E_ARG,
         "discs odd n");
  expect(nb_ic_ball_world(1, 10, 0.0, p, v, m) == NB_E_ARG, "ball (world) zero world");
  Bodies l(3);
  expect(nb_ic_lagrange(5.0, l.pos.data(), l.vel.data(), l.mass.data()) == 0, "lagrange");
  std::vector<int32_t> sizes(512);
  expect(nb_ic_log_uniform_sizes(4000, 512, 256, 32768, sizes.data()) == 0, "log-uniform sizes");
  for (int32_t s : sizes) expect(s >= 256 && s <= 32768, "log-uniform size in range");
  expect(nb_ic_log_uniform_sizes(1, 4, 10, 5, sizes.data()) == NB_E_ARG, "log-uniform hi < lo");
  std::vector<double> u(100);
  expect(nb_ic_uniform(1, 100, 1.0, 40.0, u.data()) == 0, "uniform");
  expect(nb_ic_uniform(1, -1, 1.0, 40.0, u.data()) == NB_E_ARG, "uniform negative count");
}

void particle_files(const std::string& dir) {
  const int64_t n = 50;
  std::vector<float> p(4 * n), v(4 * n), a(4 * n), m(n);
  for (int64_t i = 0; i < 4 * n; ++i) {
    p[i] = (i % 4 == 3) ? 0.f : 0.1f * (float)i - 3.3f;
    v[i] = (i % 4 == 3) ? 0.f : 1.0f / (float)(i + 1);
    a[i] = (i % 4 == 3) ? 0.f : -2.5e-3f * (float)i;
  }
  for (int64_t i = 0; i < n; ++i) m[i] = 1.0f + (float)i / 7.0f;
  const std::string path = dir + "/bodies.txt";
  expect(nb_ic_write_file(path.c_str(), n, p.data(), v.data(), a.data(), m.data()) == 0, "write");
  expect(nb_ic_read_file(path.c_str(), 0, nullptr, nullptr, nullptr, nullptr) == n, "count");
  std::vector<float> p2(4 * n), v2(4 * n), a2(4 * n), m2(n);
  expect(nb_ic_read_file(path.c_str(), n, p2.data(), v2.data(), a2.data(), m2.data()) == n, "read");
  expect(p2 == p && v2 == v && a2 == a && m2 == m, "write/read round trip is exact");
  // a cap below the count writes only that many bodies and still counts all
  std::vector<float> p3(4 * 10), v3(4 * 10), a3(4 * 10), m3(10);
  const int64_t got = nb_ic_read_file(path.c_str(), 10, p3.data(), v3.data(), a3.data(), m3.data());
  expect(got == n, "capped read counts every body");
  expect(std::vector<float>(p.begin(), p.begin() + 40) == p3 && std::vector<float>(m.begin(),
         m.begin() + 10) == m3, "capped read writes the first bodies");
  // fewer bodies than the header states: the complete ones are read
  const std::string cut = dir + "/cut.txt";
  if (FILE* f = std::fopen(cut.c_str(), "w")) {
    std::fprintf(f, "Particle System with 5 particles:\nParticle ID: 0\nPosition: (1, 2, 3)\n"
                    "Velocity: (0, 0, 0)\nAcceleration: (0, 0, 0)\nMass: 4\n");
    std::fclose(f);
  }
  expect(nb_ic_read_file(cut.c_str(), 5, p2.data(), v2.data(), a2.data(), m2.data()) == 1,
         "short file");
  // a triplet cut in the middle is malformed
  const std::string torn = dir + "/torn.txt";
  if (FILE* f = std::fopen(torn.c_str(), "w")) {
    std::fprintf(f, "Particle System with 5 particles:\nParticle ID: 0\nPosition: (4, 5");
    std::fclose(f);
  }
  expect(nb_ic_read_file(torn.c_str(), 5, p2.data(), v2.data(), a2.data(), m2.data()) == NB_E_ARG,
         "torn triplet");
  // no header, no file
  const std::string bad = dir + "/bad.txt";
  if (FILE* f = std::fopen(bad.c_str(), "w")) {
    std::fprintf(f, "not a particle file\n\x01\x02");
    std::fclose(f);
  }
  expect(nb_ic_read_file(bad.c_str(), 5, p2.data(), v2.data(), a2.data(), m2.data()) == NB_E_ARG,
         "file without header");
  expect(nb_ic_read_file((dir + "/missing.txt").c_str(), 0, nullptr, nullptr, nullptr, nullptr) ==
             NB_E_ARG,
         "missing file");
  expect(nb_ic_write_file((dir + "/no/such/dir/x.txt").c_str(), n, p.data(), v.data(), a.data(),
                          m.data()) != 0,
         "unwritable path");
}

void oracle() {
  const int64_t n = 600;
  Bodies b(n);
  expect(nb_ic_plummer(11, n, 1.0, 1.0, 10.0, 1.0, b.pos.data(), b.vel.data(), b.mass.data()) == 0,
         "oracle input");
  b.pos[3] = b.pos[0];  // a coincident pair
  b.pos[4] = b.pos[1];
  b.pos[5] = b.pos[2];
  std::vector<double> m64(b.mass.begin(), b.mass.end()), acc(3 * n);
  or_direct_f64(n, b.pos.data(), m64.data(), 0.01, 1.0, 0, nullptr, acc.data(), 2);
  const int64_t targets[3] = {0, 17, n - 1};
  or_direct_f64(n, b.pos.data(), m64.data(), 0.01, 1.0, 3, targets, acc.data(), 2);
  double ek = 0, ep = 0, mom[3];
  or_energy_f64(n, b.pos.data(), b.vel.data(), m64.data(), 0.01, 1.0, &ek, &ep, mom, 2);
  expect(std::isfinite(ek) && std::isfinite(ep), "energy");
  std::vector<double> f3(3 * n, 0.0), a3(3 * n, 0.0), pos = b.pos, vel = b.vel;
  or_kdk_update_f64(n, pos.data(), vel.data(), a3.data(), acc.data(), 1e-3);
  or_euler_update_f64(n, pos.data(), vel.data(), a3.data(), acc.data(), 1e-3);

  std::vector<float> p4, v4, m4, a4(4 * n, 0.f), f4(4 * n);
  to_vec4(b, n, p4, v4, m4);
  or_pp_forces(n, p4.data(), m4.data(), 0.01f, 1.0f, f4.data(), 0, n, 1);
  or_pp_forces(n, p4.data(), m4.data(), 0.01f, 1.0f, f4.data(), 100, 200, 2);
  or_kdk_update(n, p4.data(), v4.data(), a4.data(), f4.data(), 1e-3f, 2);
  or_euler_update(n, p4.data(), v4.data(), f4.data(), 1e-3f);
  or_pp_step(n, p4.data(), v4.data(), a4.data(), m4.data(), f4.data(), 0.01f, 1.0f, 1e-3f, 2);

  int64_t nodes = 0;
  expect(or_v5_forces(n, p4.data(), m4.data(), 0.01f, 1.0f, 0.5f, f4.data(), 0.0, &nodes, 2) == 0,
         "v5");
  expect(or_v5_forces(n, p4.data(), m4.data(), 0.01f, 1.0f, 0.5f, f4.data(), 0.01, &nodes, 2) == -5,
         "v5 pool overflow is reported");
  double bs = 0, fs = 0;
  expect(or_v6_forces(n, p4.data(), m4.data(), 0.01f, 1.0f, 0.5f, f4.data(), 0, n, 2, &bs, &fs) == 0,
         "v6");

  float bmin[3], bmax[3];
  or_bbox(n, p4.data(), bmin, bmax);
  std::vector<uint32_t> codes(n), perm(n), sorted(n);
  for (int cubic = 0; cubic <= 1; ++cubic) {
    or_morton(n, p4.data(), bmin, bmax, cubic, codes.data());
    or_sort_codes(n, codes.data(), perm.data(), sorted.data());
    std::vector<float> p4s(4 * n), com;
    for (int64_t k = 0; k < n; ++k)
      for (int c = 0; c < 4; ++c)
        p4s[4 * k + c] = c == 3 ? m4[4 * perm[k]] : p4[4 * perm[k] + c];
    const int64_t bound = or_tree_bound(n);
    for (int leaf : {1, 4, 16}) {
      std::vector<int32_t> fc(bound), nx(bound), st(bound), nb(bound), lv(bound);
      const int64_t cnt = or_build_tree(n, sorted.data(), leaf, bound, fc.data(), nx.data(),
                                        st.data(), nb.data(), lv.data());
      expect(cnt > 0 && cnt <= bound, "tree node count");
      com.assign(4 * cnt, 0.f);
      or_tree_com(cnt, st.data(), nb.data(), n, p4s.data(), com.data());
      or_bh_forces(n, p4s.data(), perm.data(), bmin, bmax, cnt, fc.data(), nx.data(), st.data(),
                   nb.data(), lv.data(), com.data(), 0.01f, 1.0f, 0.5f, f4.data(), 2);
      std::vector<int32_t> rs(n), rc(n);
      const int64_t ranges =
          or_bh_wave_ranges(n, cnt, fc.data(), st.data(), nb.data(), lv.data(), NB_BH_WAVE_BODIES,
                            rs.data(), rc.data());
      expect(ranges > 0 && ranges <= n, "wave ranges");
      std::vector<float> r2(cnt);
      or_tree_open_r2(cnt, st.data(), lv.data(), com.data(), sorted.data(), bmin, bmax, 1, 0.5f,
                      r2.data());
      or_bh_forces_group(n, p4s.data(), perm.data(), bmin, bmax, cnt, fc.data(), nx.data(),
                         st.data(), nb.data(), lv.data(), com.data(), sorted.data(), 1, 0.01f, 1.0f,
                         0.5f, NB_BH_WAVE_BODIES, f4.data(), 2);
    }
  }
  // the pipelines take (x, y, z, m) bodies
  std::vector<float> pm(p4);
  for (int64_t i = 0; i < n; ++i) pm[4 * i + 3] = m4[4 * i];
  int64_t cnt = 0;
  or_bh_pipeline(n, pm.data(), 0.01f, 1.0f, 0.5f, 8, f4.data(), codes.data(), perm.data(), &cnt, 2);
  or_bh_pipeline_group(n, pm.data(), 0.01f, 1.0f, 0.5f, 16, NB_BH_WAVE_BODIES, 1, f4.data(),
                       nullptr, nullptr, nullptr, 2);
  // one body, and every body at one point
  std::vector<float> one(4, 0.f), one_f(4);
  one[3] = 1.f;
  or_bh_pipeline(1, one.data(), 0.01f, 1.0f, 0.5f, 1, one_f.data(), nullptr, nullptr, nullptr, 1);
  // fewer bodies than the largest leaf: the root is a leaf, cut into several ranges
  std::vector<float> fifty(4 * 50), fifty_f(4 * 50);
  for (int64_t i = 0; i < 50; ++i)
    for (int c = 0; c < 4; ++c) fifty[4 * i + c] = c == 3 ? 1.f : pm[4 * i + c];
  or_bh_pipeline_group(50, fifty.data(), 0.01f, 1.0f, 0.5f, 64, NB_BH_WAVE_BODIES, 1, fifty_f.data(),
                       nullptr, nullptr, nullptr, 1);
  std::vector<float> same(4 * 64, 0.f);
  for (int64_t i = 0; i < 64; ++i) same[4 * i + 3] = 1.f;
  std::vector<float> same_f(4 * 64);
  or_bh_pipeline_group(64, same.data(), 0.01f, 1.0f, 0.5f, 1, NB_BH_WAVE_BODIES, 1, same_f.data(),
                       nullptr, nullptr, nullptr, 1);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
    return 2;
  }
  initial_conditions();
  particle_files(argv[1]);
  oracle();
  std::printf("sanitize_driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
