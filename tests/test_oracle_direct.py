"""Pins the oracle's direct sum (or_direct_f64, or_pp_forces) with closed forms
and invariants (reference: ParticleSolverCPUSequential.cpp:26-45)."""
import math

import numpy as np
import pytest

from helpers import mass4, rel_err, seeded_ball, vec4


@pytest.mark.parametrize("eps2", [0.0, 0.0935, 40.0])
def test_two_body_closed_form(ora, eps2):
    d, m1, m2, G = 3.0, 2.0, 5.0, 1.3
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    m = np.array([m1, m2])
    acc = ora.direct_f64(pos, m, eps2, G)
    k = G * d / (d * d + eps2) ** 1.5
    np.testing.assert_allclose(acc[0], [m2 * k, 0, 0], rtol=1e-14)
    np.testing.assert_allclose(acc[1], [-m1 * k, 0, 0], rtol=1e-14)
    f32 = ora.pp_forces(vec4(pos), mass4(m), eps2, G)
    np.testing.assert_allclose(f32[:, :3], acc, rtol=2e-6, atol=0)
    assert np.all(f32[:, 3] == 0)


@pytest.mark.parametrize("eps2", [0.0, 0.0935])
def test_lagrange_triangle_closed_form(ora, eps2):
    # ParticleSystemLagrange.cpp:8-50: equal masses 10.5 on an equilateral
    # triangle of side |world|/10; each body accelerates toward the centroid with
    # |a| = sqrt(3) G m L / (L^2 + eps2)^(3/2).
    w = 5.0
    L = math.sqrt(3 * w * w) / 10
    c = np.array([w / 2] * 3)
    v = np.array([[0, L / math.sqrt(3), 0], [-L / 2, -L / (2 * math.sqrt(3)), 0],
                  [L / 2, -L / (2 * math.sqrt(3)), 0]]) + c
    m = np.full(3, 10.5)
    acc = ora.direct_f64(v, m, eps2)
    mag = math.sqrt(3) * 10.5 * L / (L * L + eps2) ** 1.5
    for i in range(3):
        u = (c - v[i]) / np.linalg.norm(c - v[i])
        np.testing.assert_allclose(acc[i], mag * u, rtol=1e-12, atol=1e-12)


def test_momentum_and_equivariance(ora):
    pos = seeded_ball(500, seed=3)
    m = np.random.default_rng(4).uniform(0.5, 2.0, 500)
    acc = ora.direct_f64(pos, m, 40.0)
    # sum_i m_i a_i = 0 by antisymmetry of the pair term
    tot = np.abs(m[:, None] * acc).sum(0)
    assert np.all(np.abs((m[:, None] * acc).sum(0)) < 1e-12 * tot)
    # translation invariance
    acc_t = ora.direct_f64(pos + np.array([123.0, -45.0, 6.0]), m, 40.0)
    assert rel_err(acc_t, acc).max() < 1e-9
    # permutation equivariance
    p = np.random.default_rng(5).permutation(500)
    acc_p = ora.direct_f64(pos[p], m[p], 40.0)
    assert rel_err(acc_p, acc[p]).max() < 1e-12
    # mirror: x -> -x flips the x component
    mir = pos * np.array([-1.0, 1.0, 1.0])
    acc_m = ora.direct_f64(mir, m, 40.0)
    np.testing.assert_allclose(acc_m, acc * np.array([-1.0, 1.0, 1.0]), rtol=1e-12, atol=1e-15)


def test_targets_and_threads(ora):
    pos = seeded_ball(300, seed=6)
    m = np.ones(300)
    full = ora.direct_f64(pos, m, 1.0, nthreads=1)
    sub = ora.direct_f64(pos, m, 1.0, targets=np.array([5, 0, 299]), nthreads=4)
    np.testing.assert_array_equal(sub, full[[5, 0, 299]])
    assert np.array_equal(ora.direct_f64(pos, m, 1.0, nthreads=8), full)


def test_v1_matches_fp64(ora):
    pos = seeded_ball(1024, seed=7)
    m = np.random.default_rng(8).uniform(0.5, 1.5, 1024)
    p4 = vec4(pos)
    ref = ora.direct_f64(p4[:, :3].astype(np.float64), m.astype(np.float32).astype(np.float64), 40.0)
    f1 = ora.pp_forces(p4, mass4(m), 40.0, nthreads=1)
    f2 = ora.pp_forces(p4, mass4(m), 40.0, nthreads=8)
    assert np.array_equal(f1, f2)  # v2's static schedule does not change per-body arithmetic
    e = rel_err(f1, ref)
    assert np.median(e) < 2e-6 and e.max() < 1e-4


def test_v1_skips_self_without_softening(ora):
    pos = seeded_ball(64, seed=9)
    f = ora.pp_forces(vec4(pos), mass4(np.ones(64)), 0.0)
    assert np.all(np.isfinite(f))


def test_v1_partial_targets(ora):
    pos = seeded_ball(200, seed=10)
    mm = mass4(np.ones(200))
    full = ora.pp_forces(vec4(pos), mm, 4.0)
    part = ora.pp_forces(vec4(pos), mm, 4.0, t0=50, t1=60)
    np.testing.assert_array_equal(part[50:60], full[50:60])
    assert np.all(part[:50] == 0) and np.all(part[60:] == 0)
