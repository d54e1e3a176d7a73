"""The serving bound that bench.py prints beside its jobs/s ratio
(scheduling_bound): on small random instances its preemptive
shortest-remaining-work-first count equals the largest number of jobs any
single-server schedule can finish by the window's end (found by brute force
over subsets), and its first-in-first-out count follows the arrival order."""
import importlib.util
import itertools
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_for_tests", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def best_count(jobs, window):
    """A set of jobs can all finish by `window` (preemption allowed, released
    at their arrival) iff the work-conserving schedule of that set, which runs
    them in arrival order with idle time only while nothing is released, ends
    by the window. The answer is the largest such set."""
    best = 0
    for k in range(len(jobs), 0, -1):
        for sub in itertools.combinations(jobs, k):
            t = 0.0
            for arrive, cost in sorted(sub):
                t = max(t, arrive) + cost
            if t <= window + 1e-12:
                return k
    return best


def test_srpt_count_is_the_maximum():
    bench = load_bench()
    rng = np.random.default_rng(17)
    for trial in range(300):
        m = int(rng.integers(1, 8))
        arrive = np.sort(rng.uniform(0, 6, m))
        costs = rng.choice([0.25, 0.5, 1.0, 1.5, 2.0, 3.0, 5.0], m)
        window = float(rng.uniform(2, 10))
        jobs = [{"t": float(a), "n": k, "steps": 1} for k, a in enumerate(arrive)]
        cost = lambda n, steps, c=costs: float(c[n])
        fifo, srpt = bench.scheduling_bound(jobs, window, cost)
        released = [(float(a), float(c)) for a, c in zip(arrive, costs) if a <= window]
        assert round(srpt * window) == best_count(released, window), (trial, released, window)
        # first in first out: jobs run whole, in arrival order, until one ends late
        t, n_fifo = 0.0, 0
        for a, c in released:
            t = max(t, a) + c
            if t > window:
                break
            n_fifo += 1
        assert round(fifo * window) == n_fifo
