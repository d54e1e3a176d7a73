"""GPU checks of nb_reorder (Morton reordering of the resident state, the
reference's dormant ParticleSystem::gpuSort, ParticleSystem.cpp:106-171):
the caller's body order is kept at every copy, Barnes-Hut forces are unchanged
bit for bit, the direct sum stays within its tolerance, and the slots end up
in Morton order."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

EPS2, DT = 40.0, 0.0035


@pytest.mark.parametrize("solver", [0, 1])
def test_copies_keep_the_callers_order(nb, solver, tmp_path):
    n = 5000
    pos, vel, m = nb.ic_plummer(41, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT, solver=solver) as sim:
        sim.upload(pos, vel, m)
        sim.reorder()
        order = sim.order()
        assert sorted(order) == list(range(n)) and not np.array_equal(order, np.arange(n))
        p, v, a = sim.download_f64()
        assert np.array_equal(p, pos) and np.array_equal(v, vel) and not a.any()
        snap = str(tmp_path / "s.nbsnap")
        sim.save_snapshot(snap)
        # a second reorder of an already sorted state changes nothing
        sim.reorder()
        assert np.array_equal(sim.order(), order)
        # an upload returns the slots to the caller's order
        sim.upload(pos, vel, m)
        assert np.array_equal(sim.order(), np.arange(n))
    with nb.Simulation(n, EPS2, DT, solver=solver) as sim:
        sim.load_snapshot(snap)
        p, v, _ = sim.download_f64()
        assert np.array_equal(p, pos) and np.array_equal(v, vel)


def test_bh_forces_unchanged_and_slots_sorted(nb):
    n = 20000
    pos, vel, m = nb.ic_plummer(42, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:
        sim.upload(pos, vel, m)
        f0 = sim.forces_f64()
        sim.reorder()
        f1 = sim.forces_f64()
        codes, perm = sim.debug_morton()
    assert np.array_equal(f0, f1)
    # after the reorder the slots hold the bodies in (code, index) order
    assert np.all(np.diff(codes.astype(np.int64)) >= 0)
    assert np.array_equal(perm, np.arange(n))


def test_direct_forces_within_tolerance(nb, ora):
    n = 8192
    pos, vel, m = nb.ic_plummer(43, n, 50.0, 1e4)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2)
    with nb.Simulation(n, EPS2, DT) as sim:
        sim.upload(pos, vel, m)
        f0 = sim.forces_f64()
        sim.reorder()
        f1 = sim.forces_f64()
    e0, e1 = rel_err(f0, ref), rel_err(f1, ref)
    print(f"direct vs FP64: p50 {np.median(e0):.2e} -> {np.median(e1):.2e} after reorder")
    assert np.median(e1) < 1e-6 and np.percentile(e1, 99) < 1e-5
    assert np.percentile(rel_err(f1, f0), 99) < 1e-6


def test_bh_trajectory_matches_unreordered(nb):
    n = 20000
    pos, vel, m = nb.ic_ball(44, n, 100.0, 1.0)
    out = []
    for reorder in (False, True):
        with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:
            sim.upload(pos, vel, m)
            if reorder:
                sim.reorder()
            sim.step(10)
            p, v, _ = sim.download_f64()
            out.append((p, v))
    scale = np.abs(out[0][0]).max()
    assert np.abs(out[0][0] - out[1][0]).max() < 1e-9 * scale


def test_render_view(nb):
    n = 3000
    pos, vel, m = nb.ic_plummer(45, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:
        sim.upload(pos, vel, m)
        sim.step(3)
        sim.reorder()  # the view stays in the caller's order
        view = sim.render_view()
        p, v, _ = sim.download_f64()
    assert np.array_equal(view[:, :3], p.astype(np.float32))
    speed = np.sqrt((v * v).sum(axis=1)).astype(np.float32)
    np.testing.assert_allclose(view[:, 3], speed, rtol=1e-6)
