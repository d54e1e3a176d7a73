"""Contexts driven from several host threads at once: graph capture on one
thread while others create, upload, copy out and destroy contexts (the serving
loop's pattern) gives the same bits as running everything on one thread."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS2, DT = 40.0, 0.0035


def bh_run(nb, seed, n, steps):
    pos, vel, m = nb.ic_plummer(seed, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:  # first step: graph capture
        sim.upload(pos, vel, m)
        sim.step(steps)
        return sim.download_f64()[0]


def copy_churn(nb, seed, n, rounds):
    out = []
    pos, vel, m = nb.ic_ball(seed, n, 100.0, 1.0)
    for _ in range(rounds):
        with nb.Simulation(n, EPS2, DT, flags=nb.F_NO_GRAPH) as sim:
            sim.upload(pos, vel, m)
            sim.step(1)
            out.append(sim.download_f64()[0])
    return out


def test_capture_while_other_threads_copy(nb):
    ref_bh = [bh_run(nb, 50 + k, 30000, 3) for k in range(4)]
    ref_churn = copy_churn(nb, 60, 20000, 1)[0]
    results, errors = {}, []

    def worker(name, fn):
        try:
            results[name] = fn()
        except Exception as e:  # reported below
            errors.append((name, e))

    threads = [threading.Thread(target=worker, args=(f"bh{k}", lambda k=k: bh_run(nb, 50 + k, 30000, 3)))
               for k in range(4)]
    threads += [threading.Thread(target=worker, args=(f"churn{k}", lambda: copy_churn(nb, 60, 20000, 6)))
                for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    for k in range(4):
        assert np.array_equal(results[f"bh{k}"], ref_bh[k])
    for k in range(2):
        assert all(np.array_equal(x, ref_churn) for x in results[f"churn{k}"])
