"""GPU parity of the direct-sum engine (drop-in for the reference's v4,
ParticleSolverGPU.cpp:31-41) against the CPU oracle, through the C ABI.

Tolerances: forces within 1e-5 relative (p99) of the FP32 v1 restatement and
a median <= 1e-5 against the FP64 oracle; positions after 100 steps within
1e-5 of the system size of the FP32 v1 run."""
import math

import numpy as np
import pytest

from helpers import mass4, rel_err

pytestmark = pytest.mark.gpu

EPS2, DT = 40.0, 0.0035


def c1_inputs(nb, n=4096, seed=101):
    pos, vel, m = nb.ic_ball(seed, n, 100.0, 1.0)
    return pos.astype(np.float32), vel.astype(np.float32), m


def test_c1_forces_vs_v1_and_fp64(nb, ora):
    p32, v32, m = c1_inputs(nb)
    with nb.Simulation(4096, EPS2, DT) as sim:
        sim.upload(nb.to_vec4(p32), nb.to_vec4(v32), m)
        f = sim.forces()
        f64 = sim.forces_f64()
    v1 = ora.pp_forces(nb.to_vec4(p32), mass4(m), EPS2)
    ref = ora.direct_f64(p32.astype(np.float64), m.astype(np.float64), EPS2)
    e1 = rel_err(f, v1)
    e64 = rel_err(f64, ref)
    print(f"C1 forces: vs v1 p50={np.median(e1):.2e} p99={np.percentile(e1, 99):.2e}; "
          f"vs FP64 p50={np.median(e64):.2e} p99={np.percentile(e64, 99):.2e}")
    assert np.percentile(e1, 99) < 1e-5
    assert np.median(e64) < 1e-6 and np.percentile(e64, 99) < 1e-5
    assert np.all(f[:, 3] == 0)


def test_c1_100_steps_vs_v1(nb, ora):
    p32, v32, m = c1_inputs(nb)
    with nb.Simulation(4096, EPS2, DT) as sim:
        sim.upload(nb.to_vec4(p32), nb.to_vec4(v32), m)
        e0 = sim.energy()
        sim.step(100)
        pos, vel, acc = sim.download()
        pos64, vel64, _ = sim.download_f64()
        e1 = sim.energy()
    # reference v1: FP32 state, lagged KDK
    rp, rv, ra = nb.to_vec4(p32), nb.to_vec4(v32), np.zeros((4096, 4), np.float32)
    for _ in range(100):
        ora.pp_step(rp, rv, ra, mass4(m), EPS2, DT, nthreads=8)
    # FP64 oracle with the same integrator
    op, ov = p32.astype(np.float64), v32.astype(np.float64)
    oa = np.zeros_like(op)
    for _ in range(100):
        # the engine evaluates forces at the FP32 view of its FP64 positions
        f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2)
        ora.kdk_update_f64(op, ov, oa, f, DT)
    size = 100.0
    d_v1 = np.abs(pos[:, :3] - rp[:, :3]).max() / size
    d_64 = np.abs(pos64 - op).max() / size
    print(f"C1 100 steps: max|x - x_v1|/R = {d_v1:.2e}, max|x - x_fp64|/R = {d_64:.2e}")
    assert d_v1 < 1e-5
    assert d_64 < 1e-9
    assert rel_err(vel[:, :3], rv[:, :3]).max() < 1e-3
    assert rel_err(vel64, ov).max() < 1e-6
    ek0, ep0, _ = e0
    ek1, ep1, p1 = e1
    # nb_energy against the oracle's FP64 energy of the same downloaded state
    ref1 = ora.energy_f64(pos64, vel64, m.astype(np.float64), EPS2)
    assert abs(ek1 - ref1[0]) < 1e-12 * abs(ref1[0])
    assert abs(ep1 - ref1[1]) < 1e-12 * abs(ref1[1])
    np.testing.assert_allclose(p1, ref1[2], atol=1e-9 * np.abs(vel64).sum())
    drift = abs((ek1 + ep1) - (ek0 + ep0)) / abs(ek0 + ep0)
    print(f"C1 energy drift over 100 steps: {drift:.2e}")


def test_c2_size_fp64_subsample(nb, ora):
    n = 1 << 18
    pos, vel, m = nb.ic_plummer(202, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    tgt = np.random.default_rng(3).choice(n, 4096, replace=False)
    p32 = pos.astype(np.float32).astype(np.float64)  # the force pass sees FP32 positions
    ref = ora.direct_f64(p32, m.astype(np.float64), EPS2, targets=tgt)
    e = rel_err(f[tgt], ref)
    print(f"N=2^18 Plummer vs FP64: p50={np.median(e):.2e} p99={np.percentile(e, 99):.2e}")
    assert np.median(e) <= 1e-5
    assert np.percentile(e, 99) < 1e-4


@pytest.mark.parametrize("bpl,js", [(1, 1), (1, 4), (1, 8), (1, 16), (2, 2), (2, 8), (4, 1),
                                    (4, 4)])
def test_launch_shapes_agree(nb, ora, bpl, js):
    n = 3000
    pos, vel, m = nb.ic_plummer(5, n, 20.0, 100.0)
    with nb.Simulation(n, 4.0, DT, bodies_per_lane=bpl, j_split=js) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
        f2 = sim.forces_f64()
    assert np.array_equal(f, f2)  # deterministic: no atomics, fixed merge order
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), 4.0)
    assert np.percentile(rel_err(f, ref), 99) < 1e-5


@pytest.mark.parametrize("n", [1, 2, 3, 255, 257, 1000])
def test_ragged_sizes(nb, ora, n):
    pos, vel, m = nb.ic_ball(n + 11, n, 10.0, 1.0)
    with nb.Simulation(n, 0.5, DT) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), 0.5)
    if n == 1:
        assert np.all(f == 0)
    else:
        assert rel_err(f, ref).max() < 1e-5


def test_unsoftened_pair_guard(nb, ora):
    # eps2 = 0 selects the guarded kernel: the self pair must not produce NaN.
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0], [0, 3.0, 0]])
    with nb.Simulation(3, 0.0, DT) as sim:
        sim.upload(pos, np.zeros((3, 3)), np.array([1, 2, 3], np.float32))
        f = sim.forces_f64()
    ref = ora.direct_f64(pos, np.array([1.0, 2, 3]), 0.0)
    assert np.all(np.isfinite(f))
    assert rel_err(f, ref).max() < 1e-6


def test_lagrange_triangle(nb):
    pos, vel, m = nb.ic_lagrange(5.0)
    with nb.Simulation(3, 0.0935, 0.001) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    c = pos.astype(np.float32).astype(np.float64).mean(0)
    L = np.linalg.norm(pos[0] - pos[1])
    mag = math.sqrt(3) * 10.5 * L / (L * L + 0.0935) ** 1.5
    for i in range(3):
        u = (c - pos[i]) / np.linalg.norm(c - pos[i])
        np.testing.assert_allclose(f[i], mag * u, rtol=2e-5, atol=1e-5 * mag)


def test_euler_integrator(nb, ora):
    n = 512
    pos, vel, m = nb.ic_plummer(9, n, 10.0, 50.0)
    with nb.Simulation(n, 1.0, 0.002, integrator=nb.INTEG_EULER) as sim:
        sim.upload(pos, vel, m)
        sim.step(20)
        p, v, a = sim.download_f64()
    op, ov, oa = pos.copy(), vel.copy(), np.zeros_like(pos)
    for _ in range(20):
        f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), 1.0)
        ora.euler_update_f64(op, ov, oa, f, 0.002)
    assert np.abs(p - op).max() < 1e-6
    assert rel_err(v, ov).max() < 1e-4


@pytest.mark.parametrize("cases", [
    ((0, [10]), (1, [1] * 10), (0, [3, 3, 4])),
    # small systems replay eight steps per graph: whole graphs, remainders run
    # eagerly, and any split of the steps gives the same bits
    ((0, [26]), (1, [26]), (0, [17, 1, 8]), (0, [7, 9, 2, 8])),
])
def test_graph_replay_matches_eager(nb, cases):
    n = 2048
    pos, vel, m = nb.ic_plummer(10, n, 10.0, 100.0)
    out = []
    for eager, chunks in cases:
        with nb.Simulation(n, 1.0, 0.001, flags=nb.F_NO_GRAPH if eager else 0) as sim:
            sim.upload(pos, vel, m)
            for k in chunks:
                sim.step(k)
            out.append(sim.download_f64())
    for o in out[1:]:
        for x, y in zip(out[0], o):
            assert np.array_equal(x, y)


def test_nonfinite_state_is_reported(nb):
    pos = np.array([[0.0, 0, 0], [np.nan, 0, 0]])
    with nb.Simulation(2, 1.0, 0.01) as sim:
        sim.upload(pos, np.zeros((2, 3)), np.ones(2, np.float32))
        sim.step(1)
        with pytest.raises(nb.NBodyError) as e:
            sim.sync()
    assert e.value.code == nb.NB_E_NONFINITE


def test_timing_counters(nb):
    n = 4096
    pos, vel, m = nb.ic_ball(1, n)
    with nb.Simulation(n, EPS2, DT) as sim:
        sim.upload(pos, vel, m)
        sim.set_timing(True)
        sim.step(3)
        t = sim.timing()
    assert t["force_calls"] == 3 and t["step_calls"] == 3
    assert 0 < t["force_ms"] <= t["step_ms"]


def test_autotune_installs_defaults(nb, ora):
    before = nb.get_tuning(20000)
    t = nb.autotune(20000, nb.TUNE_DIRECT | nb.TUNE_BH, reps=2)
    assert (t["bodies_per_lane"], t["j_split"]) in {(b, k) for b in (1, 2, 4) for k in (1, 2, 4)} | \
        {(1, 8), (1, 16), (2, 8)}
    assert t["leaf_size"] in (1, 2, 4, 8, 16, 32)
    assert t["direct_ms"] > 0 and t["bh_ms"] > 0
    after = nb.get_tuning(30000)  # same power-of-two class
    assert (after["bodies_per_lane"], after["j_split"], after["leaf_size"]) == \
        (t["bodies_per_lane"], t["j_split"], t["leaf_size"])
    assert before["leaf_size"] == 8  # built-in default before tuning
    pos, vel, m = nb.ic_plummer(3, 20000, 20.0, 100.0)
    with nb.Simulation(20000, 4.0, 0.001) as sim:  # uses the tuned shape
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    tgt = np.arange(0, 20000, 97)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), 4.0,
                         targets=tgt)
    assert np.percentile(rel_err(f[tgt], ref), 99) < 1e-5



def test_c2_size_symmetries(nb):
    """Size-independent properties at the full N = 2^18: the x -> -x mirror is
    bitwise (every operation of the pass is odd in dx), a permutation of the
    input permutes the forces up to the summation order, the momentum balance
    sum_i m_i a_i = 0 holds to FP32 pair rounding, and doubling every mass
    doubles every force bit for bit."""
    n = 1 << 18
    pos, vel, m = nb.ic_plummer(203, n, 50.0, 1e4)

    def forces(p, mm):
        with nb.Simulation(n, EPS2, DT) as sim:
            sim.upload(p, vel, mm)
            return sim.forces_f64()

    f = forces(pos, m)
    flip = np.array([-1.0, 1.0, 1.0])
    np.testing.assert_array_equal(forces(pos * flip, m), f * flip)
    perm = np.random.default_rng(4).permutation(n)
    e = rel_err(forces(pos[perm], m[perm]), f[perm])
    print(f"N=2^18 permuted input: p99={np.percentile(e, 99):.2e} max={e.max():.2e}")
    assert np.percentile(e, 99) < 1e-6 and e.max() < 1e-5
    ma = m.astype(np.float64)[:, None] * f
    assert np.all(np.abs(ma.sum(0)) < 1e-6 * np.abs(ma).sum(0))
    np.testing.assert_array_equal(forces(pos, 2.0 * m), 2.0 * f)  # every mass term scales exactly
