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
                N = 2^22, theta = 0.5 (skip with --no-bh).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "n-body_amd"))

METRIC = "interactions/s + % FP32 peak; BH 4M step ms; batched-job p50/p99 latency"
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_INTERACTION = 20
EPS2, DT = 40.0, 0.0035


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    import numpy as np

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
        "sample": f"v2 force pass (OpenMP static, -O3 -ffast-math) over {targets} targets x {n} "
                  f"sources of the same Plummer input, {el:.1f} s",
    }


def bench_bh(nb, steps: int, warmup: int):
    import numpy as np

    n = 1 << 22
    t0 = time.perf_counter()
    pos, vel, m = nb.ic_discs(2024, n)
    log(f"bh: discs generated ({time.perf_counter() - t0:.1f} s)")
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=8) as sim:
        sim.upload(pos, vel, m)
        inter = sim.count_interactions()
        sim.step(warmup)
        sim.sync()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        wall = time.perf_counter() - t0
        t = sim.timing()
    ms = 1e3 * wall / steps
    return {
        "workload": "Barnes-Hut N=4194304, two exponential discs (Rd=30, h=2, 300 apart, "
                    "closing 0.5/step), theta=0.5, leaf=8, eps2=40, dt=0.0035",
        "steps": steps,
        "ms_per_step": ms,
        "build_ms": t["build_ms"] / max(1, t["build_calls"]),
        "traverse_ms": t["force_ms"] / max(1, t["force_calls"]),
        "interactions_first_step": inter,
        "interactions_per_s": inter / (t["force_ms"] / max(1, t["force_calls"]) / 1e3),
    }


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import nbody_amd as nb

    log(nb.device_info())
    n = args.n
    pos, vel, m = nb.ic_plummer(2018, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT, bodies_per_lane=args.bodies_per_lane,
                       j_split=args.j_split) as sim:
        sim.upload(pos, vel, m)
        sim.step(args.warmup)
        sim.sync()
        sim.set_timing(True)
        t0 = time.perf_counter()
        sim.step(args.steps)
        sim.sync()
        wall = time.perf_counter() - t0
        t = sim.timing()
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
        result["bh_4m"] = bench_bh(nb, args.bh_steps, 2)
        log(f"bh: {result['bh_4m']['ms_per_step']:.2f} ms/step")
    if not args.no_cpu:
        log("cpu baseline (oracle v2 restatement)")
        result["cpu_baseline"] = cpu_baseline(pos.astype(np.float32), m, args.cpu_budget)
        log(f"cpu: {result['cpu_baseline']['value']:.3e} interactions/s on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
