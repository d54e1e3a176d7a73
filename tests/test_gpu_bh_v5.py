"""The default Barnes-Hut pass on the GPU against the reference's own
Barnes-Hut semantics: the v5 sequential octree (Octree.cpp:81-338 driven by
ParticleSolverBHutCPUSeq.cpp:16-36) and, where v5 would take minutes, its
OpenMP form v6 (ParallelOctreeCPU.cpp:7-505 with
ParticleSolverBHutCPUParallel.cpp:18-41), both restated in the oracle and
pinned there (v6 == v5 on seeded inputs, theta -> 0 == direct sum).

On identical seeded inputs at theta = 0.5, eps2 = 40, the GPU's relative
force error against the FP64 direct sum must be no larger than the
reference's, at the median and at the 99th percentile."""
import numpy as np
import pytest

from helpers import mass4, rel_err

pytestmark = pytest.mark.gpu

EPS2, DT, THETA = 40.0, 0.0035, 0.5


def pm4(pos, m):
    p = np.zeros((len(m), 4), np.float32)
    p[:, :3] = pos
    p[:, 3] = m
    return p


def gpu_forces(nb, pos, vel, m, leaf=0):
    with nb.Simulation(len(m), EPS2, DT, solver=nb.SOLVER_BH, theta=THETA, leaf_size=leaf) as sim:
        sim.upload(pos, vel, m)
        return sim.forces()


def report(name, e_gpu, e_ref, ref_name):
    g = (np.median(e_gpu), np.percentile(e_gpu, 99))
    r = (np.median(e_ref), np.percentile(e_ref, 99))
    print(f"{name}: engine p50 {g[0]:.3e} p99 {g[1]:.3e} | {ref_name} p50 {r[0]:.3e} p99 {r[1]:.3e}",
          flush=True)
    return g, r


@pytest.mark.parametrize("ic", ["discs", "plummer", "ball"])
@pytest.mark.parametrize("leaf", [0, 1, 16])
def test_at_least_as_accurate_as_v5_2_16(nb, ora, ic, leaf):
    n = 1 << 16
    pos, vel, m = {"discs": lambda: nb.ic_discs(2024, n),
                   "plummer": lambda: nb.ic_plummer(2016, n, 50.0, 1e4),
                   "ball": lambda: nb.ic_ball(2017, n, 100.0, 1.0)}[ic]()
    p4 = pm4(pos, m)
    ref = ora.direct_f64(p4[:, :3].astype(np.float64), m.astype(np.float64), EPS2)
    rc, f5, _ = ora.v5_forces(p4, mass4(m), EPS2, THETA)
    assert rc == 0
    f = gpu_forces(nb, pos, vel, m, leaf)
    g, r = report(f"{ic} N=2^16 leaf {leaf or 'tuned'}", rel_err(f, ref), rel_err(f5, ref), "v5")
    assert g[0] <= r[0] and g[1] <= r[1]


def test_at_least_as_accurate_as_v6_c3(nb, ora):
    """The C3 two-disc collision at N = 2^22: 4096 targets of each disc
    (contiguous runs of the seeded input, i.e. random bodies of that disc),
    against the v6 restatement computing exactly those targets."""
    n = 1 << 22
    pos, vel, m = nb.ic_discs(2024, n)
    p4 = pm4(pos, m)
    print("C3: GPU forces", flush=True)
    f = gpu_forces(nb, pos, vel, m)
    e_gpu, e_ref = [], []
    for t0 in (0, n // 2):
        t1 = t0 + 4096
        tg = np.arange(t0, t1)
        print(f"C3: FP64 direct sum and v6 on targets [{t0}, {t1})", flush=True)
        ref = ora.direct_f64(p4[:, :3].astype(np.float64), m.astype(np.float64), EPS2, targets=tg)
        rc, f6, _, _ = ora.v6_forces(p4, mass4(m), EPS2, THETA, t0=t0, t1=t1)
        assert rc == 0
        e_gpu.append(rel_err(f[tg], ref))
        e_ref.append(rel_err(f6[tg], ref))
    g, r = report("C3 discs N=2^22 (8192 targets)", np.concatenate(e_gpu), np.concatenate(e_ref), "v6")
    assert g[0] <= r[0] and g[1] <= r[1]
