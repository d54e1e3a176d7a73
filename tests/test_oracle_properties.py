"""Randomised property tests of the CPU oracle (hypothesis): systems of 2..64
bodies with random positions, masses, length scale and softening. Every
property is fixed by the mathematics of the pair sum, not by the oracle's
own code:

- momentum balance, sum_i m_i a_i = 0 (Newton's third law);
- permutation equivariance and the x -> -x mirror;
- homogeneity: scaling positions by 2^k and eps2 by 4^k scales accelerations by
  4^-k exactly in binary floating point, and doubling every mass doubles them;
- the FP32 v1 restatement and the Barnes-Hut pipeline with theta -> 0 reduce to
  the FP64 direct sum;
- the octree conserves mass, and its leaves partition the bodies;
- the Morton sort returns a permutation with non-decreasing 30-bit codes."""
import numpy as np
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from helpers import rel_err

SETTINGS = settings(max_examples=40, deadline=None,
                    suppress_health_check=[HealthCheck.too_slow])


@st.composite
def systems(draw, nmin=2, nmax=64):
    n = draw(st.integers(nmin, nmax))
    rng = np.random.default_rng(draw(st.integers(0, 2 ** 32 - 1)))
    scale = draw(st.sampled_from([1e-2, 1.0, 1e3]))
    pos = rng.normal(size=(n, 3)) * scale
    m = rng.uniform(0.1, 10.0, n)
    eps2 = draw(st.sampled_from([1e-4, 1.0, 40.0])) * scale * scale
    return pos, m, eps2


@SETTINGS
@given(systems())
def test_direct_momentum_permutation_mirror(ora, s):
    pos, m, eps2 = s
    acc = ora.direct_f64(pos, m, eps2, nthreads=1)
    ma = m[:, None] * acc
    assert np.all(np.abs(ma.sum(0)) <= 1e-10 * np.abs(ma).sum(0) + 1e-300)
    p = np.random.default_rng(len(m)).permutation(len(m))
    acc_p = ora.direct_f64(pos[p], m[p], eps2, nthreads=1)
    assert rel_err(acc_p, acc[p]).max() < 1e-12
    flip = np.array([-1.0, 1.0, 1.0])
    acc_m = ora.direct_f64(pos * flip, m, eps2, nthreads=1)
    np.testing.assert_array_equal(acc_m, acc * flip)


@SETTINGS
@given(systems(), st.integers(-4, 4))
def test_direct_homogeneity(ora, s, k):
    pos, m, eps2 = s
    acc = ora.direct_f64(pos, m, eps2, nthreads=1)
    lam = 2.0 ** k
    acc_s = ora.direct_f64(pos * lam, m, eps2 * lam * lam, nthreads=1)
    np.testing.assert_array_equal(acc_s, acc / (lam * lam))
    np.testing.assert_array_equal(ora.direct_f64(pos, 2.0 * m, eps2, nthreads=1), 2.0 * acc)


def vec4(pos, m):
    p4 = np.zeros((len(m), 4), np.float32)
    p4[:, :3] = pos
    m4 = np.zeros((len(m), 4), np.float32)
    m4[:, 0] = m
    return p4, m4


@SETTINGS
@given(systems())
def test_fp32_v1_is_the_direct_sum(ora, s):
    pos, m, eps2 = s
    p4, m4 = vec4(pos, m)
    f = ora.pp_forces(p4, m4, eps2)
    ref = ora.direct_f64(p4[:, :3].astype(np.float64), m4[:, 0].astype(np.float64), eps2,
                         nthreads=1)
    assert np.median(rel_err(f, ref)) < 1e-5


@SETTINGS
@given(systems(nmin=1, nmax=64), st.sampled_from([1, 4, 16]))
def test_tree_partition_mass_and_sort(ora, s, leaf):
    pos, m, _ = s
    p4, _ = vec4(pos, m)
    p4[:, 3] = m
    n = len(m)
    mn, mx = ora.bbox(p4)
    codes = ora.morton(p4, mn, mx, cubic=True)
    assert codes.max() < 2 ** 30
    perm, srt = ora.sort_codes(codes)
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.all(np.diff(srt.astype(np.int64)) >= 0)
    np.testing.assert_array_equal(srt, codes[perm])
    tree = ora.build_tree(srt, leaf)
    leaves = tree["first_child"] < 0
    cover = np.zeros(n, np.int64)
    for s0, c in zip(tree["start"][leaves], tree["count"][leaves]):
        cover[s0:s0 + c] += 1
    assert np.all(cover == 1)  # every body in exactly one leaf
    com = ora.tree_com(tree, p4[perm])
    np.testing.assert_allclose(com[0, 3], m.astype(np.float32).astype(np.float64).sum(),
                               rtol=1e-6)


@SETTINGS
@given(systems(nmin=2, nmax=48))
def test_barnes_hut_theta_to_zero_is_the_direct_sum(ora, s):
    pos, m, eps2 = s
    p4, _ = vec4(pos, m)
    p4[:, 3] = m
    f, *_ = ora.bh_pipeline(p4, eps2, theta=1e-6, leaf_size=4, nthreads=1, group=32, cubic=True)
    ref = ora.direct_f64(p4[:, :3].astype(np.float64), p4[:, 3].astype(np.float64), eps2,
                         nthreads=1)
    assert np.percentile(rel_err(f, ref), 99) < 1e-4
