"""GPU edge cases of both solvers against the FP64 oracle: massless bodies
(test particles that feel the field and exert none), a system far from the
origin, and coincident bodies."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

EPS2, DT = 40.0, 0.0035


def oracle_forces(ora, pos, m, eps2=EPS2):
    return ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), eps2)


@pytest.mark.parametrize("solver", [0, 1])
def test_massless_bodies(nb, ora, solver):
    n = 6000
    pos, vel, m = nb.ic_plummer(61, n, 50.0, 1e4)
    m = m.copy()
    m[::3] = 0.0  # every third body is a test particle
    with nb.Simulation(n, EPS2, DT, solver=solver) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    ref = oracle_forces(ora, pos, m)
    e = rel_err(f, ref)
    print(f"solver {solver}: massless bodies, vs FP64 p50={np.median(e):.2e} max={e.max():.2e}")
    # test particles feel the field like everyone else
    assert np.all(np.linalg.norm(f[::3], axis=1) > 0)
    if solver == 0:
        assert e.max() < 1e-5
    else:
        assert np.median(e) < 5e-3
    # and exert none: removing them leaves the others' forces unchanged
    keep = np.ones(n, bool)
    keep[::3] = False
    ref_rest = oracle_forces(ora, pos[keep], m[keep])
    assert np.percentile(rel_err(f[keep], ref_rest), 99) < (1e-5 if solver == 0 else 1e-2)


def test_direct_far_from_origin(nb, ora):
    # a 10-unit cluster 3e4 units out: FP32 positions keep ~2e-3 absolute
    # resolution, and differences of nearby FP32 coordinates are exact, so the
    # forces on the FP32 inputs match the FP64 oracle on the same inputs
    n = 3000
    pos, vel, m = nb.ic_plummer(62, n, 2.0, 10.0, rmax_over_a=5.0)
    pos = pos + np.array([3.0e4, -2.0e4, 1.0e4])
    with nb.Simulation(n, 0.25, DT) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    ref = oracle_forces(ora, pos, m, 0.25)
    e = rel_err(f, ref)
    print(f"far from origin: p50={np.median(e):.2e} max={e.max():.2e}")
    assert e.max() < 1e-5


def test_direct_coincident_bodies(nb, ora):
    # pairs at the same point contribute exactly nothing to each other
    n = 1000
    pos, vel, m = nb.ic_ball(63, n, 20.0, 1.0)
    pos[1::2] = pos[0::2]
    with nb.Simulation(n, 0.5, DT) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
    ref = oracle_forces(ora, pos, m, 0.5)
    assert np.all(np.isfinite(f))
    assert rel_err(f, ref).max() < 1e-5
    # the two bodies of a pair feel the same force
    np.testing.assert_array_equal(f[0::2], f[1::2])
