#!/usr/bin/env python3
"""Benchmark of the MI355X N-body engine (one GPU).

Headline workload (one "step" = one direct-sum force pass + FP64 lagged-KDK
integration over all bodies): N = 2^18 Plummer sphere (a = 50, total mass 1e4,
truncated at 10 a, seeded), eps2 = 40, dt = 0.0035. `value` is interactions/s =
N (N - 1) K / wall time of K steps, inputs resident in HBM.

Also reported on the same line:
  roofline      the direct-sum kernel: 20 flop per interaction / its mean
                duration from HIP events on the engine's stream, against the
                MI355X FP32 vector peak (157.3 TFLOP/s);
  cpu_baseline  the oracle's v2 restatement (OpenMP, the reference's flags) on a
                bounded sample of the same workload on this host;
  bh_4m         Barnes-Hut step time on two colliding exponential discs,
                N = 2^22, theta = 0.5 (skip with --no-bh), leaf size tuned on
                the discs themselves (nb_tune_leaf), with the force error on
                4096 seeded targets and the FP64 energy drift over the run
                (both outside the timed steps); its roofline object prices
                the traversal kernel at 20 flop per interaction performed on
                the first timed step, over the mean traversal launch time, and
                roofline_per_body_count at the interactions the same opening
                test needs body by body (group_excess is their ratio);
                ms_per_step_after_reorder times the same steps once more after
                one Morton reorder of the resident state (nb_reorder: layout
                only, same work, bitwise the same forces);
  batch_c4      512 independent Plummer clusters (N_c log-uniform in
                [256, 32768]) in one persistent launch per step (--no-batch),
                with the same roofline object for the batched force kernel and
                the v2 restatement timed cluster by cluster as its CPU baseline;
  c1            the reference's CPU-runnable case (N = 4096, 100 steps) on the
                GPU, and on the host's cores with the v1 and v2 restatements;
  serve         the 120-s Poisson job stream (50 jobs/s) through the continuous-
                batching server (look-ahead 16 and 4096) and one job at a time,
                the first S seconds each (--serve S, default 20, labelled a
                reduced window below 120; 0 skips it): jobs and body-steps
                finished per second and p50/p99 job latency.
Force errors are measured against nb_direct_f64, the FP64 direct sum on the
device (the GPU tests pin it to the CPU oracle).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "n-body_amd"))

METRIC = "interactions/s + % FP32 peak; BH 4M step ms; batched-job p50/p99 latency"
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_INTERACTION = 20
EPS2, DT = 40.0, 0.0035


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> str:
    """CPU model of this host, for the baseline records."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def load_traffic(name: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE),
    committed under profiles/ as a one-line JSON summary; None when absent."""
    p = ROOT / "profiles" / f"{name}.traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(pos, mass, budget_s: float):
    """v2 (ParticleSolverCPUParallel.cpp:12-40, OpenMP static) force pass over a
    bounded set of targets against all N sources, on this host's cores."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as ora

    n = pos.shape[0]
    p4 = np.zeros((n, 4), np.float32)
    p4[:, :3] = pos
    m4 = np.zeros((n, 4), np.float32)
    m4[:, 0] = mass
    th = host_threads()
    lib = ora.load(fast=True)
    out = np.zeros((n, 4), np.float32)
    t_cal = 32 * th
    t0 = time.perf_counter()
    lib.or_pp_forces(n, ora._p(p4), ora._p(m4), EPS2, 1.0, ora._p(out), 0, t_cal, th)
    dt_cal = time.perf_counter() - t0
    rate = t_cal * (n - 1) / max(dt_cal, 1e-9)
    targets = int(min(n, max(t_cal, rate * budget_s / (n - 1))))
    targets = max(th, targets // th * th)
    t0 = time.perf_counter()
    lib.or_pp_forces(n, ora._p(p4), ora._p(m4), EPS2, 1.0, ora._p(out), 0, targets, th)
    el = time.perf_counter() - t0
    return {
        "value": targets * (n - 1) / el,
        "unit": "interactions/s",
        "cores": th,
        "kind": "port",
        "cpu": host_cpu(),
        "sample": f"v2 force pass (OpenMP static, -O3 -ffast-math) over {targets} targets x {n} "
                  f"sources of the same Plummer input, {el:.1f} s",
    }


def cpu_baseline_bh(pos, mass, budget_s: float):
    """v6 (ParallelOctreeCPU.cpp:7-505, ParticleSolverBHutCPUParallel.cpp:18-41)
    Barnes-Hut force pass on this host's cores: the whole tree build, then the
    force loop over a bounded set of targets, extrapolated to one full step."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as ora

    n = pos.shape[0]
    p4 = np.zeros((n, 4), np.float32)
    p4[:, :3] = pos
    m4 = np.zeros((n, 4), np.float32)
    m4[:, 0] = mass
    th = host_threads()
    cal = 256 * th
    rc, _, build_s, force_s = ora.v6_forces(p4, m4, EPS2, 0.5, t0=0, t1=cal, nthreads=th, fast=True)
    if rc != 0:
        return None
    targets = int(min(n, max(cal, cal * budget_s / max(force_s, 1e-9))))
    rc, _, build_s, force_s = ora.v6_forces(p4, m4, EPS2, 0.5, t0=0, t1=targets, nthreads=th,
                                            fast=True)
    if rc != 0:
        return None
    step_s = build_s + force_s * n / targets
    return {
        "value": 1e3 * step_s,
        "unit": "ms/step",
        "cores": th,
        "kind": "port",
        "cpu": host_cpu(),
        "sample": f"v6 restatement (-O3 -ffast-math, OpenMP): full tree build {build_s:.2f} s, force "
                  f"loop over targets [0, {targets}) of the same discs in {force_s:.1f} s, "
                  f"extrapolated to {n} targets",
    }


def force_errors(sim, force, targets):
    """Relative force error on `targets` against the FP64 direct sum over all
    sources (nb_direct_f64: on the device, from the same FP32 positions the
    force pass reads; pinned against the CPU oracle by the GPU tests).
    Outside every timed region."""
    ref = sim.direct_f64(targets)
    f = force[np.asarray(targets)]
    e = np.linalg.norm(f - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return {"targets": len(targets), "p50": float(np.median(e)), "p99": float(np.percentile(e, 99)),
            "max": float(e.max())}


def valu_roofline(kernel: str, interactions: float, ms: float):
    """FP32 vector-ALU roofline of a force kernel: 20 flop per counted interaction
    over its mean launch time (HIP events on its own stream)."""
    achieved = 20.0 * interactions / (ms / 1e3) / 1e12
    return {"kernel": kernel, "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS}


def bench_bh(nb, steps: int, warmup: int, cpu_budget: float = 0.0, accuracy: bool = True):
    n = 1 << 22
    t0 = time.perf_counter()
    pos, vel, m = nb.ic_discs(2024, n)
    log(f"bh: discs generated ({time.perf_counter() - t0:.1f} s)")
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:
        sim.upload(pos, vel, m)
        tuned = sim.tune_leaf(3)  # on these discs, not on a stand-in system
        walk = "single ranges" if tuned["bh_single_ranges"] else "pairs of ranges"
        log(f"bh: leaf size {tuned['leaf_size']} with {walk} tuned on the discs "
            f"({tuned['bh_ms']:.2f} ms force pass)")
        inter = sim.count_interactions()
        inter_pb = sim.count_interactions_per_body()
        log(f"bh: {inter:.4e} interactions performed, {inter_pb:.4e} needed body by body "
            f"(ratio {inter / inter_pb:.3f})")
        e0 = sum(sim.energy()[:2]) if accuracy else None  # FP64 direct potential, untimed
        sim.step(warmup)
        sim.sync()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        wall = time.perf_counter() - t0
        t = sim.timing()
        acc = None
        if accuracy:  # forces at the final state, 4096 seeded targets
            sim.set_timing(False)
            f = sim.forces_f64()
            tg = np.random.default_rng(7).choice(n, 4096, replace=False)
            fe = force_errors(sim, f, tg)
            e1 = sum(sim.energy()[:2])
            acc = {"force_error_vs_fp64": fe, "energy_drift": abs(e1 - e0) / abs(e0),
                   "energy_steps": warmup + steps}
            log(f"bh: force error vs FP64 p50 {fe['p50']:.2e} p99 {fe['p99']:.2e}; "
                f"energy drift {acc['energy_drift']:.2e} over {warmup + steps} steps")
        # the same system after one Morton reorder of the resident state: the
        # step does the same work, with coalesced gathers and force writes
        sim.set_timing(False)
        sim.reorder()
        sim.step(warmup)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        reordered_ms = 1e3 * (time.perf_counter() - t0) / steps
        log(f"bh: {reordered_ms:.2f} ms/step after one Morton reorder")
    ms = 1e3 * wall / steps
    trav_ms = t["force_ms"] / max(1, t["force_calls"])
    cpu = None
    if cpu_budget > 0:
        log("bh: cpu baseline (oracle v6 restatement)")
        cpu = cpu_baseline_bh(pos.astype(np.float32), m, cpu_budget)
    return {
        "workload": "Barnes-Hut N=4194304, two exponential discs (Rd=30, h=2, 300 apart, "
                    "closing 0.5/step), theta=0.5, eps2=40, dt=0.0035",
        "leaf_size": tuned["leaf_size"],
        "single_ranges": bool(tuned["bh_single_ranges"]),
        "steps": steps,
        "ms_per_step": ms,
        "ms_per_step_after_reorder": reordered_ms,
        "build_ms": t["build_ms"] / max(1, t["build_calls"]),
        "traverse_ms": trav_ms,
        "interactions_first_step": inter,
        "interactions_per_body_test": inter_pb,
        "group_excess": inter / inter_pb,
        "interactions_per_s": inter / (trav_ms / 1e3),
        "roofline": valu_roofline("k_traverse_list", inter, trav_ms),
        "roofline_per_body_count": valu_roofline("k_traverse_list", inter_pb, trav_ms),
        "accuracy": acc,
        "cpu_baseline": cpu,
    }


def c4_clusters(nb, count=512, seed=4000):
    """512 Plummer clusters, N_c log-uniform in [256, 32768], eps2 ~ U[1, 40],
    dt ~ U[0.001, 0.005]; each cluster a = 10, m = 0.1 per body."""
    sizes = nb.log_uniform_sizes(seed, count, 256, 32768)
    eps2 = nb.uniform(seed + 1, count, 1.0, 40.0)
    dts = nb.uniform(seed + 2, count, 0.001, 0.005)
    out = []
    for c in range(count):
        pos, vel, m = nb.ic_plummer(seed * 1000 + c, int(sizes[c]), 10.0, 0.1 * int(sizes[c]))
        out.append((pos, vel, m, float(eps2[c]), float(dts[c])))
    return out


def cpu_baseline_batch(clusters, budget_s: float):
    """v2 (ParticleSolverCPUParallel.cpp:12-40) force pass of the ensemble on this
    host's cores, one cluster after another as a CPU would run them: seeded
    clusters are timed in a random order until the budget is spent, and the
    rate over their interactions is extrapolated to the whole ensemble."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as ora

    th = host_threads()
    lib = ora.load(fast=True)
    order = np.random.default_rng(17).permutation(len(clusters))
    done_inter, el, used = 0.0, 0.0, 0
    for c in order:
        pos, _, m, eps2, _ = clusters[c]
        k = len(m)
        p4 = np.zeros((k, 4), np.float32)
        p4[:, :3] = pos
        m4 = np.zeros((k, 4), np.float32)
        m4[:, 0] = m
        out = np.zeros((k, 4), np.float32)
        t0 = time.perf_counter()
        lib.or_pp_forces(k, ora._p(p4), ora._p(m4), eps2, 1.0, ora._p(out), 0, k, th)
        el += time.perf_counter() - t0
        done_inter += float(k) * (k - 1)
        used += 1
        if el > budget_s:
            break
    total = sum(float(len(c[2])) * (len(c[2]) - 1) for c in clusters)
    rate = done_inter / el
    return {
        "value": rate,
        "unit": "interactions/s",
        "ms_per_step": 1e3 * total / rate,
        "cores": th,
        "kind": "port",
        "cpu": host_cpu(),
        "sample": f"v2 force pass (OpenMP static over each cluster's bodies, -O3 -ffast-math) of "
                  f"{used} of the {len(clusters)} clusters in seeded random order "
                  f"({done_inter:.3e} of {total:.3e} interactions per step, {el:.1f} s), "
                  f"extrapolated to the whole ensemble",
    }


def bench_batch(nb, steps: int, warmup: int, energy: bool, tune: bool = True, cpu_budget: float = 0.0):
    t0 = time.perf_counter()
    cl = c4_clusters(nb)
    log(f"batch: 512 clusters generated ({time.perf_counter() - t0:.1f} s)")
    total = sum(len(c[2]) for c in cl)
    with nb.Batch(sum((len(c[2]) + 255) // 256 * 256 for c in cl), len(cl)) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        shape = b.tune(2) if tune else None
        if shape:
            log(f"batch: tuned shape {shape['bodies_per_lane']}x{shape['j_split']} "
                f"({shape['direct_ms']:.2f} ms force pass)")
        e0 = [b.energy(i) for i in ids] if energy else None
        _, bodies, inter = b.info()
        b.step(warmup)
        b.sync()
        b.set_timing(True)
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        wall = time.perf_counter() - t0
        t = b.timing()
        res = {
            "workload": "512 Plummer clusters, N_c log-uniform [256, 32768] (seeded), eps2~U[1,40], "
                        "dt~U[0.001,0.005], one persistent force launch + one integrate launch per step",
            "bodies": total,
            "interactions_per_step": inter,
            "steps": steps,
            "ms_per_step": 1e3 * wall / steps,
            "force_ms": t["force_ms"] / max(1, t["force_calls"]),
            "interactions_per_s": inter * steps / wall,
            "roofline": valu_roofline("k_batch_force", inter, t["force_ms"] / max(1, t["force_calls"])),
            "shape": {"bodies_per_lane": shape["bodies_per_lane"], "j_split": shape["j_split"],
                      "tuned": True} if shape else "built-in 1x4",
        }
        if energy:
            e1 = [b.energy(i) for i in ids]
            drift = np.array([abs((a[0] + a[1]) - (c[0] + c[1])) / abs(c[0] + c[1])
                              for a, c in zip(e1, e0)])
            res["energy_drift_p50"] = float(np.median(drift))
            res["energy_drift_max"] = float(drift.max())
            res["energy_steps"] = steps + warmup
    res["cpu_baseline"] = cpu_baseline_batch(cl, cpu_budget) if cpu_budget > 0 else None
    return res


def bench_c1(nb, cpu: bool):
    """The reference's own CPU-runnable case: N = 4096 uniform ball (R = 100, m = 1,
    v = 0), eps2 = 40, dt = 0.0035, 100 steps of the lagged KDK. The GPU runs all
    100 steps; the CPU restatements of v1 (sequential) and v2 (OpenMP) run them too."""
    n, steps = 4096, 100
    pos, vel, m = nb.ic_ball(1, n, 100.0, 1.0)
    with nb.Simulation(n, EPS2, DT) as sim:
        sim.upload(pos, vel, m)
        sim.step(2)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        wall = time.perf_counter() - t0
    out = {"workload": "direct sum N=4096 uniform ball R=100, m=1, eps2=40, dt=0.0035, 100 steps",
           "ms_per_step": 1e3 * wall / steps, "interactions_per_s": n * (n - 1) * steps / wall}
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as ora

        p4 = np.zeros((n, 4), np.float32)
        p4[:, :3] = pos
        m4 = np.zeros((n, 4), np.float32)
        m4[:, 0] = m
        th = host_threads()
        for name, threads, k in (("v1", 1, 10), ("v2", th, steps)):
            p, v, a = p4.copy(), np.zeros_like(p4), np.zeros_like(p4)
            t0 = time.perf_counter()
            for _ in range(k):
                ora.pp_step(p, v, a, m4, EPS2, DT, nthreads=threads, fast=True)
            el = time.perf_counter() - t0
            out[f"cpu_{name}"] = {"ms_per_step": 1e3 * el / k, "interactions_per_s": n * (n - 1) * k / el,
                                  "cores": threads, "kind": "port", "cpu": host_cpu(),
                                  "sample": f"{name} restatement, {k} of the 100 steps"}
    return out


def poisson_jobs(seed: int, rate: float, duration: float):
    """The serving stream: Poisson arrivals; N log-uniform in [1024, 1048576];
    steps uniform in [10, 500]; Plummer or uniform ball with p = 1/2; eps2 = 40."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        n = int(np.floor(np.exp(rng.uniform(np.log(1024), np.log(1 << 20)))))
        jobs.append({"t": t, "n": n, "steps": int(rng.integers(10, 501)), "ic": int(rng.integers(0, 2)),
                     "seed": int(rng.integers(0, 2 ** 62))})
    return jobs


def job_ic(nb, j):
    """The job's seeded initial conditions, as the server makes them."""
    if j["ic"] == nb.IC_PLUMMER:
        return nb.ic_plummer(j["seed"], j["n"], 50.0, 1e4)
    return nb.ic_ball(j["seed"], j["n"], 100.0, 1.0)


def run_stream(nb, jobs, limit_s: float, **server_kw):
    """Batched mode: feed the job stream to nb_server in real time for limit_s
    seconds. Returns finished-job records, body-steps done (finished jobs and the
    steps of unfinished ones), and the submitted / queued / running counts."""
    done, body_steps = [], 0
    with nb.Server(**server_kw) as s:
        ids, k = {}, 0
        t_start = time.perf_counter()
        while True:
            now = time.perf_counter() - t_start
            while k < len(jobs) and jobs[k]["t"] <= now:
                j = jobs[k]
                ids[s.submit(j["n"], j["steps"], j["ic"], j["seed"])] = j
                k += 1
            if now > limit_s:
                break
            running, finished = s.tick()
            if finished:
                for jid in s.poll():
                    st = s.status(jid)
                    if st["state"] == nb.JOB_DONE:
                        done.append({"n": ids[jid]["n"], "steps": st["steps"],
                                     "latency": st["t_done"] - st["t_submit"]})
                        body_steps += ids[jid]["n"] * st["steps"]
                    s.forget(jid)
                    del ids[jid]
            elif running == 0 and k < len(jobs):
                time.sleep(min(0.01, max(0.0, jobs[k]["t"] - now)))
        queued = running = 0
        for jid, j in ids.items():
            st = s.status(jid)
            queued += st["state"] == nb.JOB_QUEUED
            running += st["state"] == nb.JOB_RUNNING
            body_steps += j["n"] * st["steps_done"]
        log(f"serve: loop {s.stats()}")
    return done, body_steps, k, queued, running


def run_one_at_a_time(nb, jobs, limit_s: float, ic_threads: int):
    """The baseline the server is compared with: the same stream run one job at
    a time, each job alone on the device start to finish (nb_create, upload,
    its steps replayed in chunks of 32, FP64 download, nb_destroy), with the
    initial conditions of arrived jobs prepared ahead on ic_threads threads."""
    from concurrent.futures import ThreadPoolExecutor

    done, body_steps = [], 0
    futs, k, k_prep = {}, 0, 0
    with ThreadPoolExecutor(ic_threads) as ex:
        t_start = time.perf_counter()

        def feed():
            now = time.perf_counter() - t_start
            nonlocal k_prep
            while k_prep < len(jobs) and k_prep < k + 2 * ic_threads and jobs[k_prep]["t"] <= now:
                futs[k_prep] = ex.submit(job_ic, nb, jobs[k_prep])
                k_prep += 1
            return now

        while k < len(jobs):
            now = feed()
            if now > limit_s:
                break
            if k not in futs:
                time.sleep(min(0.01, max(0.0, jobs[k]["t"] - now)))
                continue
            j = jobs[k]
            pos, vel, m = futs.pop(k).result()
            solver = nb.SOLVER_BH if j["n"] >= 65536 else nb.SOLVER_DIRECT
            left = j["steps"]
            with nb.Simulation(j["n"], EPS2, DT, solver=solver) as sim:
                sim.upload(pos, vel, m)
                while left > 0 and time.perf_counter() - t_start <= limit_s:
                    c = min(32, left)
                    sim.step(c)
                    sim.sync()
                    left -= c
                    body_steps += j["n"] * c
                    feed()
                if left == 0:
                    sim.download_f64()
                    done.append({"n": j["n"], "steps": j["steps"],
                                 "latency": time.perf_counter() - t_start - j["t"]})
            k += 1
        for f in futs.values():
            f.cancel()
    queued = sum(1 for j in jobs[k:] if j["t"] <= limit_s)
    return done, body_steps, min(len(jobs), k_prep), queued, int(k < len(jobs))


def job_cost_model(nb):
    """Device seconds of one job run alone, from timed runs on this GPU:
    cost = steps * step(N) + setup(N), with step(N) and setup(N) (context
    creation, upload, download, teardown) interpolated in log N between timed
    sizes of each solver."""
    pts = {}
    for solver, sizes, k in ((nb.SOLVER_DIRECT, (1024, 4096, 16384, 65535), 32),
                             (nb.SOLVER_BH, (65536, 262144, 1048576), 16)):
        rows = []
        for n in sizes:
            pos, vel, m = nb.ic_plummer(99, n, 50.0, 1e4)
            t0 = time.perf_counter()
            with nb.Simulation(n, EPS2, DT, solver=solver) as sim:
                sim.upload(pos, vel, m)
                sim.step(2)
                sim.sync()
                t1 = time.perf_counter()
                sim.step(k)
                sim.sync()
                t2 = time.perf_counter()
                sim.download_f64()
            t3 = time.perf_counter()
            step = (t2 - t1) / k
            rows.append((n, step, (t1 - t0) - 2 * step + (t3 - t2)))
        pts[solver] = rows

    def cost(n, steps):
        rows = pts[nb.SOLVER_DIRECT if n < 65536 else nb.SOLVER_BH]
        x = np.log([r[0] for r in rows])
        step = np.exp(np.interp(np.log(n), x, np.log([r[1] for r in rows])))
        setup = np.interp(np.log(n), x, [r[2] for r in rows])
        return steps * step + setup

    return cost, {str(k): [[int(r[0]), r[1] * 1e3, r[2] * 1e3] for r in v] for k, v in pts.items()}


def scheduling_bound(jobs, window: float, cost):
    """Jobs per second finished within `window` by (a) first-in-first-out one
    job at a time and (b) preemptive shortest-remaining-work-first, the policy
    that maximises the number of finished jobs, both with every job costing its
    one-at-a-time device time (no overlap between jobs)."""
    import heapq

    arr = [j for j in jobs if j["t"] <= window]
    c = [cost(j["n"], j["steps"]) for j in arr]
    t, fifo = 0.0, 0
    for j, cj in zip(arr, c):
        t = max(t, j["t"]) + cj
        if t > window:
            break
        fifo += 1
    h, now, i, done = [], 0.0, 0, 0
    while now < window:
        while i < len(arr) and arr[i]["t"] <= now:
            heapq.heappush(h, c[i])
            i += 1
        if not h:
            if i >= len(arr):
                break
            now = arr[i]["t"]
            continue
        nxt = min(arr[i]["t"] if i < len(arr) else window, window)
        r = heapq.heappop(h)
        if now + r <= nxt:
            now += r
            done += 1
        else:
            heapq.heappush(h, r - (nxt - now))
            now = nxt
    return fifo / window, done / window


def bench_serve(nb, rate: float, duration: float, limit: float, lookaheads=(16, 4096)):
    jobs = poisson_jobs(5050, rate, duration)
    out = {"workload": f"Poisson stream {rate:g} jobs/s for {duration:g} s (seeded), N log-uniform "
                       f"[1024, 2^20], direct below 65536 else Barnes-Hut theta=0.5, steps U[10,500], "
                       f"Plummer or ball, eps2=40, dt=0.0035",
           "window_s": limit,
           "window": "the full stream" if limit >= duration else
                     f"reduced: the first {limit:g} s of the {duration:g}-s stream, per mode",
           "offered_jobs": len(jobs),
           "offered_body_steps": int(sum(j["n"] * j["steps"] for j in jobs if j["t"] <= limit))}
    ic_threads = 8
    modes = [("one_at_a_time", None, 0)] + [(f"batched_lookahead{la}", la, nb.SERVE_ARRIVAL)
                                            for la in lookaheads]
    modes.append((f"batched_shortest_lookahead{lookaheads[-1]}", lookaheads[-1], nb.SERVE_SHORTEST))
    for mode, la, policy in modes:
        log(f"serve: {mode}")
        if la is None:
            done, bs, submitted, queued, running = run_one_at_a_time(nb, jobs, limit, ic_threads)
        else:
            done, bs, submitted, queued, running = run_stream(nb, jobs, limit, lookahead=la,
                                                              ic_threads=ic_threads, policy=policy)
        lat = np.array([d["latency"] for d in done]) if done else np.array([np.nan])
        out[mode] = {"submitted": submitted, "completed": len(done), "running": running,
                     "queued": queued, "body_steps_per_s": bs / limit,
                     "completed_per_s": len(done) / limit,
                     "latency_p50_s": float(np.percentile(lat, 50)),
                     "latency_p99_s": float(np.percentile(lat, 99))}
        log(f"serve: {mode}: {out[mode]}")
    o = out["one_at_a_time"]
    for mode, la, _ in modes[1:]:
        b = out[mode]
        b["ratio_body_steps"] = b["body_steps_per_s"] / max(o["body_steps_per_s"], 1e-9)
        b["ratio_jobs"] = b["completed_per_s"] / max(o["completed_per_s"], 1e-9)
    log("serve: job costs alone (capacity bound)")
    cost, table = job_cost_model(nb)
    fifo, srpt = scheduling_bound(jobs, limit, cost)
    out["capacity_bound"] = {
        "what": "jobs/s any scheduler can finish in the window if every job costs its one-at-a-time "
                "device time: shortest-remaining-work-first (the maximum) and FIFO (the baseline's "
                "order), from job costs timed alone on this GPU",
        "srpt_jobs_per_s": srpt, "fifo_jobs_per_s": fifo, "ratio": srpt / max(fifo, 1e-9),
        "step_ms_and_setup_ms_by_n": table}
    out["capacity_bound"]["srpt_vs_measured_one_at_a_time"] = srpt / max(o["completed_per_s"], 1e-9)
    best = max(modes[1:], key=lambda m: out[m[0]]["completed_per_s"])[0]
    out["best_mode"] = {"mode": best, "ratio_jobs": out[best]["ratio_jobs"],
                        "fraction_of_srpt_bound": out[best]["completed_per_s"] / max(srpt, 1e-9)}
    log(f"serve: capacity bound {srpt:.1f} jobs/s (FIFO model {fifo:.1f}, ratio {srpt / max(fifo, 1e-9):.2f}; "
        f"measured one-at-a-time {o['completed_per_s']:.1f}); best {best}: "
        f"{out[best]['completed_per_s']:.1f} jobs/s, {out['best_mode']['fraction_of_srpt_bound']:.2f} of the bound")
    return out


def copy_probe(nbytes: int):
    """One device-to-device copy of nbytes (read nbytes + write nbytes): the
    reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: 3 copies of {nbytes} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 18)
    ap.add_argument("--bodies-per-lane", type=int, default=0)
    ap.add_argument("--j-split", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bh", action="store_true")
    ap.add_argument("--bh-steps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-energy", action="store_true",
                    help="kept for old command lines: the per-cluster FP64 energy drift is always reported")
    ap.add_argument("--no-batch-energy", action="store_true",
                    help="skip the per-cluster FP64 energies (outside the timed steps)")
    ap.add_argument("--serve", type=float, default=20.0,
                    help="seconds of the serving stream per mode (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (built-in shapes)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64 force-error and energy-drift report (outside the timed region)")
    ap.add_argument("--only", choices=["direct", "bh", "batch"], default=None,
                    help="run one workload alone (for profiling); prints its sub-object")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(the direct sum would shard its targets over GPUs with one all-gather of "
              "the sources per step); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe > 0:
        copy_probe(int(args.copy_probe * (1 << 20)))
        return

    import nbody_amd as nb

    log(nb.device_info())
    if args.only == "bh":
        print(json.dumps(bench_bh(nb, args.bh_steps, 2, 0.0, not args.no_accuracy)), flush=True)
        return
    if args.only == "batch":
        print(json.dumps(bench_batch(nb, args.batch_steps, 2, not args.no_batch_energy, not args.no_tune)),
              flush=True)
        return
    if args.only == "direct":
        args.no_bh = args.no_batch = args.no_cpu = True
        args.serve = 0.0
    n = args.n
    tuning = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tuning = nb.autotune(n, nb.TUNE_DIRECT, 3)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tuning}")
    pos, vel, m = nb.ic_plummer(2018, n, 50.0, 1e4)
    accuracy = None
    with nb.Simulation(n, EPS2, DT, bodies_per_lane=args.bodies_per_lane,
                       j_split=args.j_split) as sim:
        sim.upload(pos, vel, m)
        e0 = sum(sim.energy()[:2]) if not args.no_accuracy else None
        sim.step(args.warmup)
        sim.sync()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.step(args.steps)
        sim.sync()
        wall = time.perf_counter() - t0
        t = sim.timing()
        if not args.no_accuracy:  # outside the timed region
            sim.set_timing(False)
            e1 = sum(sim.energy()[:2])
            f = sim.forces_f64()
            tg = np.random.default_rng(5).choice(n, 4096, replace=False)
            accuracy = {"force_error_vs_fp64": force_errors(sim, f, tg),
                        "energy_drift": abs(e1 - e0) / abs(e0),
                        "energy_steps": args.warmup + args.steps}
            log(f"direct: force error vs FP64 p50 {accuracy['force_error_vs_fp64']['p50']:.2e} "
                f"p99 {accuracy['force_error_vs_fp64']['p99']:.2e}; energy drift "
                f"{accuracy['energy_drift']:.2e} over {accuracy['energy_steps']} steps")
    inter = float(n) * (n - 1)
    value = inter * args.steps / wall
    force_s = t["force_ms"] / t["force_calls"] / 1e3
    achieved = FLOP_PER_INTERACTION * inter / force_s / 1e12
    log(f"direct: {1e3 * wall / args.steps:.2f} ms/step, force kernel {1e3 * force_s:.2f} ms, "
        f"{value:.3e} interactions/s, {achieved:.1f} TFLOP/s")
    traffic = load_traffic("direct_n262144")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded Plummer sphere (a=50, M=1e4, r<=10a), FP32 force, FP64 integration",
        "config": {
            "workload": f"direct-sum N={n} Plummer, eps2=40, dt=0.0035, lagged KDK",
            "n": n,
            "eps2": EPS2,
            "dt": DT,
            "force_ms": 1e3 * force_s,
            "pct_fp32_peak": 100.0 * achieved / FP32_PEAK_TFLOPS,
            "shape": nb.get_tuning(n),
            "autotune": tuning,
            "accuracy": accuracy,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
        },
        "cpu_baseline": None,
    }
    if not args.no_bh:
        log("bh: N=2^22 two-disc collision")
        result["bh_4m"] = bench_bh(nb, args.bh_steps, 2, 0.0 if args.no_cpu else args.cpu_budget,
                                   not args.no_accuracy)
        log(f"bh: {result['bh_4m']['ms_per_step']:.2f} ms/step")
    if not args.no_batch:
        log("batch: C4 ensemble")
        result["batch_c4"] = bench_batch(nb, args.batch_steps, 2, not args.no_batch_energy, not args.no_tune,
                                         0.0 if args.no_cpu else args.cpu_budget)
        log(f"batch: {result['batch_c4']['ms_per_step']:.2f} ms/step")
    log("c1: N=4096, 100 steps")
    result["c1"] = bench_c1(nb, not args.no_cpu)
    if args.serve > 0:
        result["serve"] = bench_serve(nb, args.serve_rate, 120.0, args.serve)
    if not args.no_cpu:
        log("cpu baseline (oracle v2 restatement)")
        result["cpu_baseline"] = cpu_baseline(pos.astype(np.float32), m, args.cpu_budget)
        log(f"cpu: {result['cpu_baseline']['value']:.3e} interactions/s on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
