"""GPU checks of the serving loop (nb_server_*): results are identical alone and
in a mix, direct-sum jobs match the FP64 oracle, Barnes-Hut jobs match a
standalone context bit for bit, and admission respects the HBM budget."""
import time

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

JOBS = [  # (n, steps, ic, seed)
    (1500, 7, 0, 1), (3000, 3, 1, 2), (9000, 5, 0, 3), (700, 9, 1, 4),
    (20000, 4, 0, 5), (5000, 6, 1, 6), (12000, 2, 0, 7), (300, 1, 0, 8),
]


def run_server(nb, jobs, **kw):
    with nb.Server(bh_threshold=8192, direct_pool=1 << 18, **kw) as s:
        ids = [s.submit(n, st, ic, seed) for n, st, ic, seed in jobs]
        t0 = time.monotonic()
        while time.monotonic() - t0 < 60:
            running, _ = s.tick()
            if all(s.status(i)["state"] in (nb.JOB_DONE, nb.JOB_FAILED) for i in ids):
                break
        out = [s.result(i) for i in ids]
        st = [s.status(i) for i in ids]
    return out, st


def test_alone_equals_mix(nb):
    mix, st = run_server(nb, JOBS)
    assert all(x["state"] == nb.JOB_DONE for x in st)
    assert all(x["steps_done"] == x["steps"] for x in st)
    for k in (0, 2, 4):  # two direct-sum jobs and one Barnes-Hut job
        alone, _ = run_server(nb, [JOBS[k]], max_jobs=1)
        for a, b in zip(alone[0], mix[k]):
            assert np.array_equal(a, b)


def test_jobs_match_standalone(nb, ora):
    mix, st = run_server(nb, JOBS[:5])
    for k, (n, steps, ic, seed) in enumerate(JOBS[:5]):
        if ic == 0:
            pos, vel, m = nb.ic_plummer(seed, n, 50.0, 1e4)
        else:
            pos, vel, m = nb.ic_ball(seed, n, 100.0, 1.0)
        if n >= 8192:
            assert st[k]["solver"] == nb.SOLVER_BH
            with nb.Simulation(n, 40.0, 0.0035, solver=nb.SOLVER_BH) as sim:
                sim.upload(pos, vel, m)
                sim.step(steps)
                p, v, _ = sim.download_f64()
            assert np.array_equal(p, mix[k][0]) and np.array_equal(v, mix[k][1])
        else:
            op, ov, oa = pos.copy(), vel.copy(), np.zeros_like(pos)
            for _ in range(steps):
                f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), 40.0)
                ora.kdk_update_f64(op, ov, oa, f, 0.0035)
            assert np.abs(mix[k][0] - op).max() < 1e-9 * np.abs(op).max()


def test_budget_limits_concurrency(nb):
    # a budget that holds the direct pool plus one Barnes-Hut context of 20000
    # bodies (about 12 MB): the Barnes-Hut jobs run one after another and all finish
    jobs = [(20000, 2, 0, 11), (20000, 2, 0, 12), (20000, 2, 0, 13)]
    pool = 1 << 16
    with nb.Server(bh_threshold=8192, direct_pool=pool, hbm_budget=pool * 116 + (18 << 20)) as s:
        ids = [s.submit(*j) for j in jobs]
        max_running = 0
        t0 = time.monotonic()
        while time.monotonic() - t0 < 60:
            running, _ = s.tick()
            max_running = max(max_running, running)
            if all(s.status(i)["state"] == nb.JOB_DONE for i in ids):
                break
        assert all(s.status(i)["state"] == nb.JOB_DONE for i in ids)
    assert max_running == 1


def test_tick_budget_limits_concurrency(nb):
    # direct-sum jobs of 20000 bodies are estimated at 0.127 ms a step each: a
    # 0.2 ms tick budget runs them one after another (no two run intervals
    # overlap), a 1 ms budget selects all three at once (their 300-step run
    # intervals overlap)
    jobs = [(20000, 300, 1, 21), (20000, 300, 1, 22), (20000, 300, 1, 23)]
    for budget, serial in ((0.2, True), (1.0, False)):
        with nb.Server(direct_pool=1 << 17, tick_budget_ms=budget) as s:
            ids = [s.submit(*j) for j in jobs]
            t0 = time.monotonic()
            while time.monotonic() - t0 < 60:
                s.tick()
                if all(s.status(i)["state"] == nb.JOB_DONE for i in ids):
                    break
            st = [s.status(i) for i in ids]
        assert all(x["state"] == nb.JOB_DONE for x in st)
        spans = sorted((x["t_admit"], x["t_done"]) for x in st)
        overlaps = [spans[k + 1][0] < spans[k][1] for k in range(2)]
        assert overlaps == [not serial] * 2, (budget, spans)


def test_job_that_never_fits_is_refused(nb):
    with nb.Server(direct_pool=1 << 12, bh_threshold=8192) as s:
        with pytest.raises(nb.NBodyError):
            s.submit(5000, 2, 0, 1)  # 5120 padded bodies > the 4096-body pool


def test_poll_returns_every_retired_job_once(nb):
    jobs = [(2000, 3, 0, 31), (12000, 2, 1, 32), (500, 1, 0, 33), (9000, 0, 1, 34)]
    with nb.Server(bh_threshold=8192, direct_pool=1 << 16) as s:
        ids = [s.submit(*j) for j in jobs]
        seen = []
        t0 = time.monotonic()
        while time.monotonic() - t0 < 60 and len(seen) < len(ids):
            s.tick()
            seen += s.poll()
        assert sorted(seen) == sorted(ids) and len(set(seen)) == len(seen)
        assert s.poll() == []
        assert all(s.status(i)["state"] == nb.JOB_DONE for i in ids)
