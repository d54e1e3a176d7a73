/*
 * nbody_ic.h — seeded initial conditions (host-only, part of libnbody).
 *
 * Counterpart of the reference's ParticleSystemInitializer::generateParticles
 * (include/ParticleSystemInitializer.h and the classes under
 * src/ParticleSystemInitializer/), whose
 * random initializers are unseeded (ParticleSystemBall.cpp:15-16,
 * ParticleSystemGalaxyInitializer.cpp:17-18). Here every generator is a pure
 * function of (seed, parameters): a SplitMix64 stream gives uniform doubles in
 * (0, 1), so CPU checks and GPU runs start from identical bytes.
 *
 * Outputs: pos3/vel3 are double[3n] (x, y, z interleaved), mass is float[n].
 * All return 0 on success or NB_E_ARG (-1); the particle-file functions return
 * NB_E_OOM (-3) when a host allocation fails.
 */
#ifndef NBODY_IC_H
#define NBODY_IC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Uniform ball: r = radius * u^(1/3), isotropic direction, zero velocity. */
int nb_ic_ball(uint64_t seed, int64_t n, double radius, double mass_each, double* pos3,
               double* vel3, float* mass);

/* Plummer sphere (Aarseth, Henon & Wielen 1974): scale radius a, total mass M,
 * radii a / sqrt(u^(-2/3) - 1) resampled above rmax_over_a * a, speeds q * v_esc(r)
 * with q drawn by rejection from q^2 (1 - q^2)^(7/2), isotropic directions; the
 * centre of mass and its velocity are moved to the origin. */
int nb_ic_plummer(uint64_t seed, int64_t n, double a, double total_mass, double rmax_over_a,
                  double G, double* pos3, double* vel3, float* mass);

/* Two exponential discs of n/2 bodies each (n even). Per disc: body 0 is a central
 * point mass of central_factor * m_b; the others have R ~ Gamma(2, scale_length)
 * (i.e. -scale_length * ln(u1 u2)), z = scale_height * atanh(2u - 1) (sech^2
 * layer), circular speed sqrt(G M(<R) R^2 / (R^2 + eps2)^(3/2)) with M(<R) the
 * central mass plus the enclosed exponential-disc mass. m_b = total_mass / n.
 * Disc centres at x = -separation/2 and +separation/2 with bulk velocities
 * +closing_speed/2 and -closing_speed/2 along x. */
int nb_ic_discs(uint64_t seed, int64_t n, double scale_length, double scale_height,
                double central_factor, double total_mass, double separation,
                double closing_speed, double eps2, double G, double* pos3, double* vel3,
                float* mass);

/* The reference's deterministic Lagrange triangle, 3 bodies
 * (ParticleSystemLagrange.cpp:8-50), computed in single precision as there. */
int nb_ic_lagrange(double world_size, double* pos3, double* vel3, float* mass);

/* The reference's visual initializers, seeded; world_size w is the edge of the
 * reference's world cube (main.cpp:29 uses 5), and every mass scale follows
 * the reference's (w^3 / n) rule.
 *
 * Cube (-i 1, ParticleSystemCubeInitializer.cpp:8-29): positions uniform in
 * [0, w]^3, velocity components uniform in [-L, L] with L = sqrt(3) w / 20,
 * masses (w^3/n) * 1.2 * U(0.8, 1.3). */
int nb_ic_cube(uint64_t seed, int64_t n, double world_size, double* pos3, double* vel3,
               float* mass);
/* Cube surface (-i 6, ParticleSystemCubeSurface.cpp:8-49): a face of [0, w]^3
 * chosen uniformly, uniform over the face; velocity and mass as the cube. */
int nb_ic_cube_surface(uint64_t seed, int64_t n, double world_size, double* pos3, double* vel3,
                       float* mass);
/* Sphere surface (-i 4, ParticleSystemSphere.cpp:20-133): uniform on the sphere
 * of radius sqrt(3) w / 3 about (w/2, w/2, w/2), masses (w^3/n) * 0.5 * U(0.7, 1.3),
 * speeds sqrt(3) w / 15 * U(0.5, 1.5). pattern 0: radially outward, 1: inward,
 * 2: along axis x r with axis = normalize(0.2, 1, -0.1); -1 draws the pattern from
 * the seed, as the reference draws it once per system. */
int nb_ic_sphere_surface(uint64_t seed, int64_t n, double world_size, int pattern, double* pos3,
                         double* vel3, float* mass);
/* Ball (-i 5, ParticleSystemBall.cpp:10-92): uniform in the ball of radius
 * R = sqrt(3) w / 3 about (w/2, w/2, w/2) (rejection from its bounding cube),
 * then each coordinate nudged by U(-1, 1) * 0.05 R; velocity components
 * U(-1, 1) * 0.001 sqrt(3) w; masses (w^3/n) * (0.5 + 4.5 u^2). */
int nb_ic_ball_world(uint64_t seed, int64_t n, double world_size, double* pos3, double* vel3,
                     float* mass);
/* Disc galaxy (-i 2, the reference's default, ParticleSystemGalaxyInitializer.cpp:
 * 12-210), centred at (w/2, w/2, w/2) in the z = w/2 plane. style 0 (lines 46-91):
 * radius U(0, w), angle 2 pi N(0, 50 pi), height U(0.1, w/5), tangential speed
 * sqrt(3) w / 4 plus a six-fold arm term inside w/2, masses (w^3/n) * (0.7 inside
 * w/10, else 0.5). style 1 (lines 92-205): bulge, disc radius, thickness, arm
 * count 2..6, arm tightness and a stable/energetic velocity range drawn once;
 * per body radius disc * sqrt(u), Gaussian height, circular speed capped at 1.5x,
 * a logarithmic-spiral radial perturbation and Gaussian dispersion. -1 draws the
 * style from the seed (the reference's 50/50 choice). */
int nb_ic_galaxy(uint64_t seed, int64_t n, double world_size, int style, double* pos3,
                 double* vel3, float* mass);

/* Reference particle-file format (ParticleSystemFile.cpp:10-90):
 *   "Particle System with N particles:" then per body "Particle ID: i",
 *   "Position: (x, y, z)", "Velocity: (...)", "Acceleration: (...)", "Mass: m".
 * A body is emitted at each "Mass:" line with the most recent position,
 * velocity and acceleration; reading stops after N bodies. Returns the number of
 * bodies read (writing at most cap of them; pass cap = 0 to count), or NB_E_ARG
 * when the file cannot be opened or has no header (the reference exits). */
int64_t nb_ic_read_file(const char* path, int64_t cap, float* pos4, float* vel4, float* acc4,
                        float* mass);
/* Writes the same format with 9 significant digits (every float round-trips;
 * the reference's writer uses the stream default of 6, ParticleSystem.cpp:205-215). */
int nb_ic_write_file(const char* path, int64_t n, const float* pos4, const float* vel4,
                     const float* acc4, const float* mass);

/* Log-uniform integers floor(exp(U(ln lo, ln hi))) in [lo, hi]. */
int nb_ic_log_uniform_sizes(uint64_t seed, int32_t count, int32_t lo, int32_t hi, int32_t* out);
/* Uniform doubles in [lo, hi). */
int nb_ic_uniform(uint64_t seed, int32_t count, double lo, double hi, double* out);

#ifdef __cplusplus
}
#endif
#endif
