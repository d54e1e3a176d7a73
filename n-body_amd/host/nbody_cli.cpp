// nbody_cli.cpp — headless driver with the reference's flags
// (ArgumentsParser.cpp:49-106; version -> solver wiring of main.cpp:73-153),
// running the HIP solvers through the ParticleSolver contract instead of a
// render loop.
//
//   -v 3|4   direct sum (ParticleSolverHIP; the reference's v3/v4 GPU solvers)
//   -v 7     Barnes-Hut (ParticleSolverBHutHIP; the reference's v7)
//   -n N     number of bodies          -t dt     time step
//   -i 1..6   the reference's initializers in its world cube of edge 5
//            (main.cpp:29): 1 cube, 2 disc galaxy (default), 3 Lagrange triangle,
//            4 sphere surface, 5 ball, 6 cube surface
//   -i 8|9   Plummer sphere (a=50, M=1e4), two colliding discs
//            every random initializer is seeded (--seed)
//   -s eps2  squared softening         -f path   particle file (reference format)
//   --steps K  steps to run (default 100)        --theta T / --leaf L (v7)
//   --out path   write the final state in the reference format (full precision)
//   --snapshot path / --resume path   binary device snapshots
//   --energy     print FP64 energy and momentum before and after
//   --reorder    Morton-reorder the resident state once after upload (gpuSort)
//   --frames K --frame-out prefix   every K steps write prefix_NNNNN.xyzs: the
//                render consumer's float32 (x, y, z, |v|) per body (nb_render_view)
// The CPU versions (-v 1, 2, 5, 6) live in the test oracle, not in this engine.
// Unlike the reference (D12), -f does not stop the parsing of later flags.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ParticleSolverHIP.h"
#include "nbody.h"
#include "nbody_ic.h"

namespace {

[[noreturn]] void usage(const char* argv0, const char* why) {
  std::fprintf(stderr, "%s\nusage: %s [-v 4|7] [-n N] [-i 1..6|8|9] [-t dt] [-s eps2] [-f file] "
               "[--steps K] [--seed S] [--theta T] [--leaf L] [--out file] [--snapshot file] "
               "[--resume file] [--energy] [--reorder] [--frames K --frame-out prefix]\n", why, argv0);
  std::exit(EXIT_FAILURE);
}

constexpr double kWorld = 5.0;  // the reference's world cube edge (main.cpp:29)

std::vector<Particle> generate(int init, int64_t n, uint64_t seed, float eps2) {
  std::vector<double> p(3 * n), v(3 * n);
  std::vector<float> m(n);
  int rc = -1;
  switch (init) {
    case 3:
      n = 3;
      p.resize(9);
      v.resize(9);
      m.resize(3);
      rc = nb_ic_lagrange(kWorld, p.data(), v.data(), m.data());
      break;
    case 1: rc = nb_ic_cube(seed, n, kWorld, p.data(), v.data(), m.data()); break;
    case 2: rc = nb_ic_galaxy(seed, n, kWorld, -1, p.data(), v.data(), m.data()); break;
    case 4: rc = nb_ic_sphere_surface(seed, n, kWorld, -1, p.data(), v.data(), m.data()); break;
    case 5: rc = nb_ic_ball_world(seed, n, kWorld, p.data(), v.data(), m.data()); break;
    case 6: rc = nb_ic_cube_surface(seed, n, kWorld, p.data(), v.data(), m.data()); break;
    case 8: rc = nb_ic_plummer(seed, n, 50.0, 1e4, 10.0, 1.0, p.data(), v.data(), m.data()); break;
    case 9:
      rc = nb_ic_discs(seed, n, 30.0, 2.0, 1e4, 1e4, 300.0, 0.5 / 0.0035, eps2, 1.0, p.data(),
                       v.data(), m.data());
      break;
    default: break;
  }
  if (rc != 0) {
    std::fprintf(stderr, "initialization %d failed (n=%lld)\n", init, (long long)n);
    std::exit(EXIT_FAILURE);
  }
  std::vector<Particle> out(n);
  for (int64_t i = 0; i < n; ++i) {
    out[i].position = Vec4{(float)p[3 * i], (float)p[3 * i + 1], (float)p[3 * i + 2], 0.f};
    out[i].velocity = Vec4{(float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2], 0.f};
    out[i].mass = m[i];
  }
  return out;
}

std::vector<Particle> read_file(const std::string& path) {
  const int64_t n = nb_ic_read_file(path.c_str(), 0, nullptr, nullptr, nullptr, nullptr);
  if (n < 0) {
    std::fprintf(stderr, "Failed to open or parse file: %s\n", path.c_str());
    std::exit(EXIT_FAILURE);
  }
  std::vector<float> p(4 * n), v(4 * n), a(4 * n), m(n);
  nb_ic_read_file(path.c_str(), n, p.data(), v.data(), a.data(), m.data());
  std::vector<Particle> out(n);
  for (int64_t i = 0; i < n; ++i) {
    out[i].position = Vec4{p[4 * i], p[4 * i + 1], p[4 * i + 2], 0.f};
    out[i].velocity = Vec4{v[4 * i], v[4 * i + 1], v[4 * i + 2], 0.f};
    out[i].acceleration = Vec4{a[4 * i], a[4 * i + 1], a[4 * i + 2], 0.f};
    out[i].mass = m[i];
  }
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  int version = 4, init = 2;             // init GALAXY (ArgumentsParser.cpp:7)
  int64_t n = 100;                       // ArgumentsParser.cpp:6
  float dt = 0.001f, eps2 = 0.0935f;     // ArgumentsParser.cpp:8-9
  int steps = 100, leaf = 0;
  float theta = 0.5f;
  uint64_t seed = 1;
  bool energy = false, reorder = false;
  int frames = 0;
  std::string file, out, snap, resume, frame_out;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> const char* {
      if (i + 1 >= argc) usage(argv[0], ("missing value for " + a).c_str());
      return argv[++i];
    };
    if (a == "-v" || a == "-version") version = std::atoi(val());
    else if (a == "-n") n = std::atoll(val());
    else if (a == "-i" || a == "-init") init = std::atoi(val());
    else if (a == "-t" || a == "-time-step") dt = std::strtof(val(), nullptr);
    else if (a == "-s" || a == "-softening") eps2 = std::strtof(val(), nullptr);
    else if (a == "-f" || a == "-file") file = val();
    else if (a == "--steps") steps = std::atoi(val());
    else if (a == "--seed") seed = std::strtoull(val(), nullptr, 10);
    else if (a == "--theta") theta = std::strtof(val(), nullptr);
    else if (a == "--leaf") leaf = std::atoi(val());
    else if (a == "--out") out = val();
    else if (a == "--snapshot") snap = val();
    else if (a == "--resume") resume = val();
    else if (a == "--energy") energy = true;
    else if (a == "--reorder") reorder = true;
    else if (a == "--frames") frames = std::atoi(val());
    else if (a == "--frame-out") frame_out = val();
    else usage(argv[0], ("unknown argument " + a).c_str());
  }
  if (n <= 0) usage(argv[0], "Invalid number of particles");
  if (frames < 0 || (frames > 0) != !frame_out.empty())
    usage(argv[0], "--frames K and --frame-out prefix go together");
  if (version == 1 || version == 2 || version == 5 || version == 6)
    usage(argv[0], "CPU solvers (-v 1, 2, 5, 6) are not part of this engine");
  if (version != 3 && version != 4 && version != 7) usage(argv[0], "Invalid version");
  if (!(init >= 1 && init <= 6) && init != 8 && init != 9)
    usage(argv[0], "Invalid initialization type");

  std::vector<Particle> particles = file.empty() ? generate(init, n, seed, eps2) : read_file(file);
  ParticleSystem system(particles, true);
  ParticleSolverHIPBase* solver = nullptr;
  if (version == 7)
    solver = new ParticleSolverBHutHIP(dt, eps2, (int)system.size(), theta, leaf);
  else
    solver = new ParticleSolverHIP(dt, eps2);
  char dev[256];
  nb_device_info(dev, sizeof(dev));
  std::printf("device: %s\nversion %d, %u bodies, dt %g, eps2 %g, %d steps\n", dev, version,
              system.size(), dt, eps2, steps);
  int rc = 0;
  try {
    solver->updateParticlePositions(&system, 0);  // create the context and upload
    nb_ctx* ctx = solver->context();
    if (!resume.empty() && (rc = nb_load_snapshot(ctx, resume.c_str())) != NB_OK)
      throw std::runtime_error(nb_last_error());
    if (reorder && (rc = nb_reorder(ctx)) != NB_OK) throw std::runtime_error(nb_last_error());
    double ek0 = 0, ep0 = 0, p0[3];
    if (energy && nb_energy(ctx, &ek0, &ep0, p0) == NB_OK)
      std::printf("E0 = %.17g (kin %.17g, pot %.17g)\n", ek0 + ep0, ek0, ep0);
    std::vector<float> view(frames > 0 ? 4 * (size_t)system.size() : 0);
    int frame_no = 0;
    auto write_frame = [&]() {
      if ((rc = nb_render_view(ctx, view.data())) != NB_OK) throw std::runtime_error(nb_last_error());
      char name[32];
      std::snprintf(name, sizeof(name), "_%05d.xyzs", frame_no++);
      FILE* f = std::fopen((frame_out + name).c_str(), "wb");
      if (!f || std::fwrite(view.data(), sizeof(float), view.size(), f) != view.size()) {
        if (f) std::fclose(f);
        throw std::runtime_error("cannot write frame " + frame_out + name);
      }
      std::fclose(f);
    };
    const auto t0 = std::chrono::steady_clock::now();
    if (frames > 0) {
      write_frame();
      for (int done = 0; done < steps; done += frames) {
        solver->updateParticlePositions(&system, std::min(frames, steps - done));
        write_frame();
      }
    } else {
      solver->updateParticlePositions(&system, steps);
    }
    if ((rc = nb_sync(ctx)) != NB_OK) throw std::runtime_error(nb_last_error());
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%d steps in %.3f s (%.3f ms/step)\n", steps, s, 1e3 * s / (steps ? steps : 1));
    if (version != 7)
      std::printf("%.4e interactions/s\n",
                  (double)system.size() * (system.size() - 1) * steps / (s > 0 ? s : 1));
    if (energy) {
      double ek, ep, p[3];
      if (nb_energy(ctx, &ek, &ep, p) == NB_OK)
        std::printf("E1 = %.17g (kin %.17g, pot %.17g), drift %.3e, p = (%.3e, %.3e, %.3e)\n",
                    ek + ep, ek, ep, (ek + ep - ek0 - ep0) / std::abs(ek0 + ep0), p[0], p[1], p[2]);
    }
    if (!snap.empty() && (rc = nb_save_snapshot(ctx, snap.c_str())) != NB_OK)
      throw std::runtime_error(nb_last_error());
    if (!out.empty()) {
      if ((rc = system.pullFromDevice()) != NB_OK) throw std::runtime_error(nb_last_error());
      std::vector<float> m(system.size());
      for (unsigned i = 0; i < system.size(); ++i) m[i] = system.getMasses()[i].x;
      rc = nb_ic_write_file(out.c_str(), system.size(), &system.getPositions()[0].x,
                            &system.getVelocities()[0].x, &system.getAccelerations()[0].x, m.data());
      if (rc != NB_OK) throw std::runtime_error("cannot write " + out);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    delete solver;
    return EXIT_FAILURE;
  }
  delete solver;
  return EXIT_SUCCESS;
}
