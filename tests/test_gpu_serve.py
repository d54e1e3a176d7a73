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
    args = dict(bh_threshold=8192, direct_pool=1 << 18)
    args.update(kw)
    with nb.Server(**args) as s:
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


def test_large_bh_job_matches_standalone(nb):
    # a job of 2^20 bodies walks its ranges in pairs (NB_BH_PAIR_MIN_BODIES)
    # with its work ordered by the previous step's times; served beside a
    # small job on a pooled context it equals a standalone context bit for bit
    jobs = [(1 << 20, 3, 0, 11), (3000, 3, 1, 2)]
    mix, st = run_server(nb, jobs)
    assert st[0]["solver"] == nb.SOLVER_BH and st[0]["state"] == nb.JOB_DONE
    n, steps, _, seed = jobs[0]
    pos, vel, m = nb.ic_plummer(seed, n, 50.0, 1e4)
    with nb.Simulation(n, 40.0, 0.0035, solver=nb.SOLVER_BH) as sim:
        sim.upload(pos, vel, m)
        sim.step(steps)
        p, v, _ = sim.download_f64()
    assert np.array_equal(p, mix[0][0]) and np.array_equal(v, mix[0][1])


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
    # a budget that holds the direct pool plus one Barnes-Hut context of the
    # 32768-body capacity class (about 15 MB): the Barnes-Hut jobs of 20000
    # bodies run one after another (no two run intervals overlap; the second and
    # third reuse the first one's pooled context) and all finish
    jobs = [(20000, 6, 0, 11), (20000, 6, 0, 12), (20000, 6, 0, 13)]
    pool = 1 << 16
    with nb.Server(bh_threshold=8192, direct_pool=pool, hbm_budget=pool * 116 + (18 << 20)) as s:
        ids = [s.submit(*j) for j in jobs]
        t0 = time.monotonic()
        while time.monotonic() - t0 < 60:
            s.tick()
            if all(s.status(i)["state"] == nb.JOB_DONE for i in ids):
                break
        st = [s.status(i) for i in ids]
    assert all(x["state"] == nb.JOB_DONE for x in st)
    spans = sorted((x["t_admit"], x["t_done"]) for x in st)
    assert all(spans[k + 1][0] >= spans[k][1] for k in range(2)), spans


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


def test_quotas_and_policy_keep_results(nb):
    """Extra steps per tick (the tick budget left over) and shortest-work-first
    admission change when steps run, not what they compute: every job's result
    equals its result alone."""
    jobs = [(1200, 40, 0, 41), (3000, 25, 1, 42), (10000, 6, 0, 43), (800, 60, 1, 44), (16000, 3, 0, 45)]
    mix, st = run_server(nb, jobs, policy=nb.SERVE_SHORTEST, lookahead=64)
    assert all(x["state"] == nb.JOB_DONE and x["steps_done"] == x["steps"] for x in st)
    for k in range(len(jobs)):
        alone, _ = run_server(nb, [jobs[k]], max_jobs=1, tick_budget_ms=0.01)
        for a, b in zip(alone[0], mix[k]):
            assert np.array_equal(a, b), k


def test_pooled_context_serves_the_next_job(nb):
    """A Barnes-Hut context goes back to the pool when its job ends and is
    rebound to the next job of its capacity class; the results equal fresh
    standalone contexts bit for bit."""
    jobs = [(12000, 3, 0, 51), (9500, 4, 1, 52), (16000, 2, 0, 53)]  # one class: 16384
    with nb.Server(bh_threshold=8192, direct_pool=1 << 16, max_jobs=1) as s:
        ids = [s.submit(*j) for j in jobs]
        t0 = time.monotonic()
        while time.monotonic() - t0 < 60:
            s.tick()
            if all(s.status(i)["state"] == nb.JOB_DONE for i in ids):
                break
        res = [s.result(i) for i in ids]
    for (n, steps, ic, seed), (p, v) in zip(jobs, res):
        pos, vel, m = (nb.ic_plummer(seed, n, 50.0, 1e4) if ic == 0 else nb.ic_ball(seed, n, 100.0, 1.0))
        with nb.Simulation(n, 40.0, 0.0035, solver=nb.SOLVER_BH) as sim:
            sim.upload(pos, vel, m)
            sim.step(steps)
            rp, rv, _ = sim.download_f64()
        assert np.array_equal(rp, p) and np.array_equal(rv, v)


def test_loop_statistics(nb):
    """nb_server_get_stats: ticks and steps are counted (the steps of finished
    jobs, no more), finished Barnes-Hut contexts wait in the pool, and the
    loop's phases add up to no more than the wall time spent in ticks."""
    jobs = [(1500, 12, 0, 61), (9000, 5, 1, 62), (3000, 7, 0, 63), (12000, 4, 0, 64)]
    with nb.Server(bh_threshold=8192, direct_pool=1 << 16) as s:
        ids = [s.submit(*j) for j in jobs]
        t0 = time.monotonic()
        wall = 0.0
        while time.monotonic() - t0 < 60:
            t1 = time.perf_counter()
            s.tick()
            wall += time.perf_counter() - t1
            if all(s.status(i)["state"] == nb.JOB_DONE for i in ids):
                break
        st = s.stats()
        assert all(s.status(i)["state"] == nb.JOB_DONE for i in ids)
    assert st["ticks"] >= 1
    assert st["steps"] == sum(j[1] for j in jobs)
    assert st["pooled_contexts"] >= 1  # the Barnes-Hut jobs' contexts went back to the pool
    phases = [st[k] for k in ("select_ms", "idle_wait_ms", "start_ms", "enqueue_ms",
                              "device_wait_ms", "retire_ms")]
    assert min(phases) >= 0.0 and sum(phases) <= 1e3 * wall * 1.01 + 1.0


def poisson_stream(seed, rate, duration):
    """The serving workload's stream: Poisson arrivals, N log-uniform in
    [1024, 2^20], steps uniform in [10, 500], Plummer or ball with p = 1/2."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return jobs
        n = int(np.floor(np.exp(rng.uniform(np.log(1024), np.log(1 << 20)))))
        jobs.append((t, n, int(rng.integers(10, 501)), int(rng.integers(0, 2)), int(rng.integers(0, 2 ** 62))))


def test_c5_stream_at_full_shape(nb, ora):
    """One second of the serving stream at its stated shape (50 jobs/s, N up to
    2^20, direct below 65536, Barnes-Hut above, up to 500 steps), fed in real
    time, then run to completion: every job finishes, the smallest direct-sum
    jobs match the FP64 oracle, and sampled jobs (the smallest direct one, a
    mid-size direct one and the smallest Barnes-Hut one) equal their result alone."""
    stream = poisson_stream(2026, 50.0, 1.0)
    print(f"C5: {len(stream)} jobs, N {min(j[1] for j in stream)}..{max(j[1] for j in stream)}", flush=True)
    with nb.Server(lookahead=4096) as s:
        ids, k, t0 = {}, 0, time.monotonic()
        while time.monotonic() - t0 < 240:
            now = time.monotonic() - t0
            while k < len(stream) and stream[k][0] <= now:
                _, n, st, ic, seed = stream[k]
                ids[s.submit(n, st, ic, seed)] = stream[k]
                k += 1
            s.tick()
            if k == len(stream) and all(s.status(i)["state"] in (nb.JOB_DONE, nb.JOB_FAILED) for i in ids):
                break
        states = {i: s.status(i) for i in ids}
        results = {i: s.result(i) for i in ids if states[i]["state"] == nb.JOB_DONE}
    print(f"C5: all submitted after {now:.1f} s; finished {len(results)} of {len(ids)}", flush=True)
    assert k == len(stream) and len(results) == len(ids)
    assert any(st["solver"] == nb.SOLVER_BH for st in states.values())
    direct = sorted((j[1], i) for i, j in ids.items() if j[1] < 65536)
    bh = sorted((j[1], i) for i, j in ids.items() if j[1] >= 65536)
    for n, i in direct[:2]:  # FP64 oracle trajectories of the two smallest direct jobs
        _, _, st, ic, seed = ids[i]
        pos, vel, m = nb.ic_plummer(seed, n, 50.0, 1e4) if ic == 0 else nb.ic_ball(seed, n, 100.0, 1.0)
        op, ov, oa = pos.copy(), vel.copy(), np.zeros_like(pos)
        for _ in range(st):
            f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), 40.0)
            ora.kdk_update_f64(op, ov, oa, f, 0.0035)
        assert np.abs(results[i][0] - op).max() < 1e-7 * np.abs(op).max(), (n, st)
    sample = [direct[0][1], direct[len(direct) // 2][1], bh[0][1]]
    for i in sample:
        _, n, st, ic, seed = ids[i]
        print(f"C5: job N={n} steps={st} alone", flush=True)
        alone, _ = run_server(nb, [(n, st, ic, seed)], max_jobs=1, bh_threshold=65536, direct_pool=1 << 17)
        for a, b in zip(alone[0], results[i]):
            assert np.array_equal(a, b), n
