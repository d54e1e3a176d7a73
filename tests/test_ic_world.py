"""The reference's visual initializers (-i 1, 2, 4, 5, 6), seeded: geometry, velocity
and mass rules of ParticleSystemCubeInitializer.cpp:8-29,
ParticleSystemGalaxyInitializer.cpp:12-210, ParticleSystemSphere.cpp:20-133,
ParticleSystemBall.cpp:10-92 and ParticleSystemCubeSurface.cpp:8-49 at the
reference's world edge w = 5 (main.cpp:29). Host-only generators: no GPU is needed.

The reference draws from an unseeded std::mt19937, so no byte of its output can
be reproduced; each test checks a property the reference's construction fixes
(a bound, a surface, a direction, a moment of the distribution), with
statistical checks held to at least five standard deviations."""
import numpy as np
import pytest

W = 5.0
N = 200_000


def zscore(sample_mean, expected, var, n):
    return abs(sample_mean - expected) / np.sqrt(var / n)


@pytest.mark.parametrize("name", ["ic_cube", "ic_cube_surface", "ic_sphere_surface", "ic_galaxy",
                                  "ic_ball_world"])
def test_seeded_and_deterministic(nb, name):
    f = getattr(nb, name)
    a, b, c = f(11, 4000), f(11, 4000), f(12, 4000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    for x in a:
        assert np.all(np.isfinite(x))


@pytest.mark.parametrize("name", ["ic_cube", "ic_cube_surface", "ic_sphere_surface", "ic_galaxy",
                                  "ic_ball_world"])
def test_arguments_checked(nb, name):
    with pytest.raises(nb.NBodyError):
        getattr(nb, name)(1, 10, -1.0)


def check_cube_velocity_and_mass(vel, m, n):
    L = np.sqrt(3.0) * W / 20.0
    assert np.abs(vel).max() <= L
    for k in range(3):
        assert zscore(vel[:, k].mean(), 0.0, L * L / 3.0, n) < 5
        assert abs(vel[:, k].var() / (L * L / 3.0) - 1.0) < 0.02  # uniform on [-L, L]
    base = W ** 3 / n * 1.2
    m64 = m.astype(np.float64)
    assert m64.min() >= base * 0.8 * (1 - 1e-6) and m64.max() <= base * 1.3 * (1 + 1e-6)
    assert zscore(m64.mean(), base * 1.05, (base * 0.5) ** 2 / 12.0, n) < 5


def test_cube(nb):
    pos, vel, m = nb.ic_cube(3, N)
    assert pos.min() >= 0.0 and pos.max() <= W
    for k in range(3):
        assert zscore(pos[:, k].mean(), W / 2, W * W / 12, N) < 5
        assert abs(pos[:, k].var() / (W * W / 12) - 1.0) < 0.02
    # the three coordinates are independent
    cov = np.cov(pos.T) / (W * W / 12)
    assert np.abs(cov - np.eye(3)).max() < 0.02
    check_cube_velocity_and_mass(vel, m, N)


def test_cube_surface(nb):
    pos, vel, m = nb.ic_cube_surface(4, N)
    on = (pos == 0.0) | (pos == W)
    # every body lies on exactly one face (an edge has probability zero)
    assert np.all(on.sum(1) == 1)
    face = np.argmax(on, axis=1) * 2 + (pos[on] == W)
    counts = np.bincount(face, minlength=6)
    p = 1 / 6
    assert np.all(np.abs(counts - N * p) < 5 * np.sqrt(N * p * (1 - p)))
    # over its face a body is uniform
    free = pos[~on].reshape(N, 2)
    assert free.min() >= 0.0 and free.max() <= W
    assert zscore(free.mean(), W / 2, W * W / 12, 2 * N) < 5
    check_cube_velocity_and_mass(vel, m, N)


@pytest.mark.parametrize("pattern", [0, 1, 2])
def test_sphere_surface(nb, pattern):
    pos, vel, m = nb.ic_sphere_surface(5, N, W, pattern)
    c = np.full(3, W / 2)
    R = np.sqrt(3.0) * W / 3.0
    d = pos - c
    r = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(r, R, rtol=1e-12)
    u = d / r[:, None]
    # uniform on the sphere: each coordinate of the direction has mean 0 and variance 1/3
    for k in range(3):
        assert zscore(u[:, k].mean(), 0.0, 1 / 3, N) < 5
        assert abs((u[:, k] ** 2).mean() - 1 / 3) < 0.01
    vb = np.sqrt(3.0) * W / 15.0
    speed = np.linalg.norm(vel, axis=1)
    radial = (vel * u).sum(1)
    if pattern == 0:
        np.testing.assert_allclose(radial, speed, rtol=1e-12)
    elif pattern == 1:
        np.testing.assert_allclose(radial, -speed, rtol=1e-12)
    else:
        axis = np.array([0.2, 1.0, -0.1]) / np.sqrt(1.05)
        swirl = speed > 0.5 * vb * 0.99  # all but bodies on the axis (measure zero)
        assert swirl.all()
        np.testing.assert_allclose(radial, 0.0, atol=1e-12 * vb)
        np.testing.assert_allclose(vel @ axis, 0.0, atol=1e-12 * vb)
        # sense of rotation: v along axis x r
        t = np.cross(axis, u)
        assert np.all((vel * t).sum(1) > 0)
    assert speed.min() >= 0.5 * vb * (1 - 1e-12) and speed.max() <= 1.5 * vb * (1 + 1e-12)
    assert zscore(speed.mean(), vb, vb * vb / 12, N) < 5
    base = W ** 3 / N * 0.5
    m64 = m.astype(np.float64)
    assert m64.min() >= 0.7 * base * (1 - 1e-6) and m64.max() <= 1.3 * base * (1 + 1e-6)


def test_sphere_surface_pattern_from_seed(nb):
    seen = set()
    for seed in range(30):
        pos, vel, _ = nb.ic_sphere_surface(seed, 64)
        u = pos - W / 2
        u /= np.linalg.norm(u, axis=1)[:, None]
        cosang = (vel * u).sum(1) / np.linalg.norm(vel, axis=1)
        if np.allclose(cosang, 1):
            seen.add(0)
        elif np.allclose(cosang, -1):
            seen.add(1)
        else:
            assert np.allclose(cosang, 0, atol=1e-9)
            seen.add(2)
    assert seen == {0, 1, 2}


def test_galaxy_style0(nb):
    pos, vel, m = nb.ic_galaxy(6, N, W, 0)
    c = W / 2
    dx, dy = pos[:, 0] - c, pos[:, 1] - c
    rad = np.hypot(dx, dy)
    assert rad.max() < W
    # radius uniform on [0, w): mean w/2
    assert zscore(rad.mean(), W / 2, W * W / 12, N) < 5
    h = pos[:, 2] - c
    assert h.min() >= 0.1 and h.max() <= W / 5
    base = W ** 3 / N
    inner = rad < W / 10
    m64 = m.astype(np.float64)
    np.testing.assert_allclose(m64[inner], 0.7 * base, rtol=1e-6)
    np.testing.assert_allclose(m64[~inner], 0.5 * base, rtol=1e-6)
    L = np.sqrt(3.0) * W / 4.0
    assert np.all(vel[:, 2] == 0)
    tx, ty = -dy / rad, dx / rad
    vt = vel[:, 0] * tx + vel[:, 1] * ty
    vr = (vel[:, 0] * dx + vel[:, 1] * dy) / rad
    # tangential speed L everywhere; inside w/2 the arm term points inward with
    # magnitude L * rad / (w/2) (its angle term is cos(6 * 2 pi k) = 1)
    np.testing.assert_allclose(vt, L, rtol=1e-9)
    arm = rad < W / 2
    np.testing.assert_allclose(vr[arm], -L * rad[arm] / (W / 2), rtol=1e-6, atol=1e-9 * L)
    np.testing.assert_allclose(vr[~arm], 0.0, atol=1e-9 * L)


def test_galaxy_style1(nb):
    pos, vel, m = nb.ic_galaxy(7, N, W, 1)
    c = W / 2
    dx, dy = pos[:, 0] - c, pos[:, 1] - c
    rad = np.hypot(dx, dy)
    disc = rad.max()
    assert 0.3 * W * 0.99 < disc <= 0.5 * W
    # radius = disc * sqrt(u): rad^2 uniform on [0, disc^2]
    q = (rad / disc) ** 2
    assert zscore(q.mean(), 0.5, 1 / 12, N) < 5
    # Gaussian layer about the centre plane, thickness at most 10% of the disc
    h = pos[:, 2] - c
    assert zscore(h.mean(), 0.0, h.var(), N) < 5
    assert h.std() <= disc * 0.10 / 3 * 1.05
    base = W ** 3 / N
    f = m.astype(np.float64) / base
    bulge = (f > 1.2 * (1 + 1e-6))  # 1.5 * U(0.8, 1.2) lies in [1.2, 1.8]
    assert np.all((f >= 0.8 * (1 - 1e-6)) & (f <= 1.8 * (1 + 1e-6)))
    # the bulge is a disc of radius 0.05 w .. 0.15 w: heavy bodies lie inside it
    assert rad[bulge].max() <= 0.15 * W
    # counter-clockwise rotation on average, with a small vertical dispersion
    vt = (-vel[:, 0] * dy + vel[:, 1] * dx) / rad
    assert vt.mean() > 0
    assert abs(vel[:, 2].mean()) < 5 * vel[:, 2].std() / np.sqrt(N)


def test_galaxy_style_from_seed(nb):
    seen = set()
    for seed in range(20):
        pos, _, _ = nb.ic_galaxy(seed, 256)
        h = pos[:, 2] - W / 2
        seen.add(0 if h.min() >= 0.1 else 1)
    assert seen == {0, 1}


def test_ball(nb):
    # ParticleSystemBall.cpp:19-89: uniform in the ball of radius R = |w|/3 about
    # the world centre, each coordinate nudged by U(-1, 1) * 0.05 R
    n = 1_000_000
    pos, vel, m = nb.ic_ball_world(8, n)
    R = np.sqrt(3.0) * W / 3.0
    d = 0.05 * R
    rel = pos - W / 2
    assert np.abs(rel).max() <= R + d
    r2 = (rel * rel).sum(1)
    assert np.sqrt(r2.max()) <= R + np.sqrt(3.0) * d
    for k in range(3):
        assert zscore(rel[:, k].mean(), 0.0, R * R / 5 + d * d / 3, n) < 5
    # E|x|^2 = 3R^2/5 for the uniform ball, plus d^2 from the nudge (3 axes of
    # variance d^2/3); a missing nudge is about 9 standard deviations away, a
    # nudge of 5% of w instead of R about 20
    var_r2 = R ** 4 * (3.0 / 7.0 - 9.0 / 25.0)
    assert zscore(r2.mean(), 0.6 * R * R + d * d, var_r2, n) < 5
    assert (r2 > R * R).any()  # the nudge pushes some bodies past R
    s = 0.001 * np.sqrt(3.0) * W
    assert np.abs(vel).max() <= s
    for k in range(3):
        assert zscore(vel[:, k].mean(), 0.0, s * s / 3, n) < 5
        assert abs(vel[:, k].var() / (s * s / 3) - 1.0) < 0.01
    # masses (w^3/n) (0.5 + 4.5 u^2): in [0.5, 5] x the mean mass, average 2x
    avg = W ** 3 / n
    m64 = m.astype(np.float64)
    assert m64.min() >= 0.5 * avg * (1 - 1e-6) and m64.max() <= 5.0 * avg * (1 + 1e-6)
    assert zscore(m64.mean(), 2.0 * avg, (4.5 * avg) ** 2 * 4.0 / 45.0, n) < 5
