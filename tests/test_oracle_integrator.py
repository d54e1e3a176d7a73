"""Pins the oracle's integrators (ParticleSystem.h:34-51 lagged KDK;
updateParticles.glsl:44-55 semi-implicit Euler) and the energy diagnostic."""
import math

import numpy as np

from helpers import mass4, vec4


def test_kdk_single_body_hand_values(ora):
    # One body, no others: F = 0. With a stored acceleration a0 (file input,
    # ParticleSystemFile.cpp:62-71): v = v0 + a0 dt/2, x = x0 + v dt, a = 0.
    pos = np.array([[1.0, 2.0, 3.0, 0.0]], np.float32)
    vel = np.array([[0.5, 0.0, -1.0, 0.0]], np.float32)
    acc = np.array([[4.0, -2.0, 0.0, 0.0]], np.float32)
    ora.pp_step(pos, vel, acc, mass4(np.ones(1)), 1.0, 0.5)
    np.testing.assert_array_equal(vel[0, :3], [1.5, -0.5, -1.0])
    np.testing.assert_array_equal(pos[0, :3], [1.75, 1.75, 2.5])
    np.testing.assert_array_equal(acc[0, :3], [0.0, 0.0, 0.0])


def test_kdk_two_body_uses_pre_drift_force(ora):
    # Two bodies at rest, 2 apart, unit masses, eps2 = 0, dt = 0.1:
    # F at the pre-drift positions is +-1/4; a0 = 0 so the first drift is zero.
    pos = np.array([[-1, 0, 0, 0], [1, 0, 0, 0]], np.float32)
    vel = np.zeros((2, 4), np.float32)
    acc = np.zeros((2, 4), np.float32)
    ora.pp_step(pos, vel, acc, mass4(np.ones(2)), 0.0, 0.1)
    np.testing.assert_array_equal(pos[:, 0], [-1, 1])
    np.testing.assert_allclose(acc[:, 0], [0.25, -0.25], rtol=1e-7)
    np.testing.assert_allclose(vel[:, 0], [0.0125, -0.0125], rtol=1e-6)
    # second step drifts with v + a dt/2 = 0.025
    ora.pp_step(pos, vel, acc, mass4(np.ones(2)), 0.0, 0.1)
    np.testing.assert_allclose(pos[:, 0], [-1 + 0.0025, 1 - 0.0025], rtol=1e-6)


def test_euler_hand_values(ora):
    pos = np.array([[1.0, 0.0, 0.0, 0.0]], np.float32)
    vel = np.array([[0.0, 2.0, 0.0, 0.0]], np.float32)
    f = np.array([[-1.0, 0.0, 4.0, 0.0]], np.float32)
    ora.euler_update(pos, vel, f, 0.5)
    np.testing.assert_array_equal(vel[0, :3], [-0.5, 2.0, 2.0])
    np.testing.assert_array_equal(pos[0, :3], [0.75, 1.0, 1.0])


def test_euler_f64_hand_values(ora):
    # updateParticles.glsl:44-55: v += F dt first, then x += v dt with the new v;
    # the stored acceleration becomes F
    pos = np.array([[1.0, 0.0, 0.0]])
    vel = np.array([[0.0, 2.0, 0.0]])
    acc = np.array([[9.0, 9.0, 9.0]])
    f = np.array([[-1.0, 0.0, 4.0]])
    ora.euler_update_f64(pos, vel, acc, f, 0.5)
    np.testing.assert_array_equal(vel[0], [-0.5, 2.0, 2.0])
    np.testing.assert_array_equal(pos[0], [0.75, 1.0, 1.0])
    np.testing.assert_array_equal(acc[0], [-1.0, 0.0, 4.0])


def test_euler_f64_constant_force_closed_form(ora):
    # Under a constant force the semi-implicit update gives v_k = v0 + k F dt and
    # x_k = x0 + k v0 dt + F dt^2 k (k + 1) / 2 (explicit Euler would give k (k - 1) / 2)
    x0, v0, F, dt, k = np.array([[3.0, -1.0, 0.5]]), np.array([[0.25, 0.0, -2.0]]), \
        np.array([[1.0, -3.0, 0.125]]), 0.125, 12
    pos, vel, acc = x0.copy(), v0.copy(), np.zeros((1, 3))
    for _ in range(k):
        ora.euler_update_f64(pos, vel, acc, F, dt)
    np.testing.assert_array_equal(vel, v0 + k * F * dt)
    np.testing.assert_array_equal(pos, x0 + k * v0 * dt + F * dt * dt * k * (k + 1) / 2)


def test_euler_f64_matches_f32_on_exact_inputs(ora):
    # inputs with few significant bits and dt = 1/8: every product and sum is
    # exact in FP32 and FP64 alike, so the two restatements agree bit for bit
    rng = np.random.default_rng(3)
    n = 200
    p, v, f = (rng.integers(-1000, 1000, (n, 3)) / 16.0 for _ in range(3))
    p4, v4, f4 = vec4(p), vec4(v), vec4(f)
    ora.euler_update(p4, v4, f4, 0.125)
    p64, v64, a64 = p.copy(), v.copy(), np.zeros((n, 3))
    ora.euler_update_f64(p64, v64, a64, f, 0.125)
    np.testing.assert_array_equal(p4[:, :3].astype(np.float64), p64)
    np.testing.assert_array_equal(v4[:, :3].astype(np.float64), v64)
    np.testing.assert_array_equal(a64, f)


def test_kdk_f64_matches_f32(ora):
    rng = np.random.default_rng(1)
    n = 100
    p, v, a, f = (rng.normal(size=(n, 3)) for _ in range(4))
    p4, v4, a4, f4 = vec4(p), vec4(v), vec4(a), vec4(f)
    ora.kdk_update(p4, v4, a4, f4, 0.01)
    p64, v64, a64 = (x.astype(np.float32).astype(np.float64) for x in (p, v, a))
    ora.kdk_update_f64(p64, v64, a64, f.astype(np.float32).astype(np.float64), 0.01)
    np.testing.assert_allclose(p4[:, :3], p64, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(v4[:, :3], v64, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(a4[:, :3], f.astype(np.float32))


def _orbit_drift(ora, dt, t_end=1.0):
    r, m = 1.0, 1.0
    vc = math.sqrt(m / (4 * r))
    pos = np.array([[-r, 0, 0], [r, 0, 0]], np.float64)
    vel = np.array([[0, -vc, 0], [0, vc, 0]], np.float64)
    acc = ora.direct_f64(pos, np.ones(2), 0.0)
    ek0, ep0, _ = ora.energy_f64(pos, vel, np.ones(2), 0.0)
    assert abs(ep0 - (-1.0 / (2 * r))) < 1e-15
    for _ in range(int(round(t_end / dt))):
        f = ora.direct_f64(pos, np.ones(2), 0.0)
        ora.kdk_update_f64(pos, vel, acc, f, dt)
    ek, ep, p = ora.energy_f64(pos, vel, np.ones(2), 0.0)
    assert np.all(np.abs(p) < 1e-12)  # momentum is conserved exactly up to rounding
    return ((ek + ep) - (ek0 + ep0)) / abs(ek0 + ep0)


def test_lagged_kdk_is_first_order(ora):
    # x_{n+1} = x_n + v_n dt + a_{n-1} dt^2/2, v_{n+1} = v_n + (a_{n-1} + a_n) dt/2:
    # velocity Verlet with accelerations one step late, so the energy error is O(dt)
    # (a textbook KDK would give O(dt^2) and a halving ratio of ~4).
    d1 = _orbit_drift(ora, 1e-3)
    d2 = _orbit_drift(ora, 5e-4)
    assert abs(d1) < 1e-3
    assert 1.8 < d1 / d2 < 2.2


def test_energy_closed_form(ora):
    pos = np.array([[0, 0, 0], [3, 4, 0]], np.float64)
    vel = np.array([[1, 0, 0], [0, -2, 0]], np.float64)
    m = np.array([2.0, 3.0])
    ek, ep, p = ora.energy_f64(pos, vel, m, 11.0, G=2.0)
    assert ek == 0.5 * 2 * 1 + 0.5 * 3 * 4
    assert abs(ep - (-2.0 * 6.0 / 6.0)) < 1e-15  # sqrt(25 + 11) = 6
    np.testing.assert_array_equal(p, [2.0, -6.0, 0.0])
