"""GPU parity of the batched-ensemble engine (nb_batch_*): every cluster of a
mixed batch against the FP64 oracle and against itself run alone."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def make_clusters(nb, count, lo, hi, seed):
    sizes = nb.log_uniform_sizes(seed, count, lo, hi)
    eps2 = nb.uniform(seed + 1, count, 1.0, 40.0)
    dt = nb.uniform(seed + 2, count, 0.001, 0.005)
    out = []
    for c in range(count):
        pos, vel, m = nb.ic_plummer(seed * 1000 + c, int(sizes[c]), 10.0, 0.1 * sizes[c])
        out.append((pos, vel, m, float(np.float32(eps2[c])), float(np.float32(dt[c]))))
    return out


def test_batch_forces_and_steps_vs_oracle(nb, ora):
    cl = make_clusters(nb, 12, 256, 3000, 7)
    with nb.Batch(sum(len(c[2]) + 256 for c in cl), 16) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        for cid, (p, v, m, e, d) in zip(ids, cl):
            f = b.forces(cid)
            ref = ora.direct_f64(p.astype(np.float32).astype(np.float64), m.astype(np.float64), e)
            assert np.percentile(rel_err(f, ref), 99) < 1e-5
        b.step(5)
        for cid, (p, v, m, e, d) in zip(ids, cl):
            gp, gv, ga = b.download(cid)
            op, ov, oa = p.copy(), v.copy(), np.zeros_like(p)
            for _ in range(5):
                f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), e)
                ora.kdk_update_f64(op, ov, oa, f, d)
            assert np.abs(gp - op).max() < 1e-9 * max(1.0, np.abs(op).max())
            assert rel_err(gv, ov).max() < 1e-5
            ek, ep = b.energy(cid)
            rk, rp, _ = ora.energy_f64(gp, gv, m.astype(np.float64), e)
            assert abs(ek - rk) <= 1e-12 * abs(rk) and abs(ep - rp) <= 1e-12 * abs(rp)


def test_cluster_alone_equals_in_mix(nb):
    cl = make_clusters(nb, 20, 256, 5000, 11)
    pick = [0, 7, 13]
    with nb.Batch(sum(len(c[2]) + 256 for c in cl), 32) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        b.step(8)
        mixed = {k: b.download(ids[k]) for k in pick}
    for k in pick:
        p, v, m, e, d = cl[k]
        with nb.Batch(len(m) + 256, 1) as b:
            cid = b.add(p, v, m, e, d)
            b.step(8)
            alone = b.download(cid)
        for x, y in zip(alone, mixed[k]):
            assert np.array_equal(x, y)


def test_add_remove_reuses_pool(nb, ora):
    cl = make_clusters(nb, 6, 300, 2000, 21)
    cap = sum((len(c[2]) + 255) // 256 * 256 for c in cl[:4])
    with nb.Batch(cap, 8) as b:
        ids = [b.add(*c[:3], c[3], c[4]) for c in cl[:4]]
        assert b.try_add(*cl[4][:3], cl[4][3], cl[4][4]) is None  # pool full
        b.step(2)
        b.remove(ids[1])
        b.remove(ids[2])
        new = b.add(*cl[4][:3], cl[4][3], cl[4][4])
        b.step(3)
        clusters, bodies, inter = b.info()
        assert clusters == 3
        p, v, m, e, d = cl[4]
        gp, gv, _ = b.download(new)
        op, ov, oa = p.copy(), v.copy(), np.zeros_like(p)
        for _ in range(3):
            f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), e)
            ora.kdk_update_f64(op, ov, oa, f, d)
        assert np.abs(gp - op).max() < 1e-9 * max(1.0, np.abs(op).max())


def test_batch_graph_equals_eager(nb):
    cl = make_clusters(nb, 8, 256, 2000, 31)
    res = []
    for flags in (0, nb.F_NO_GRAPH):
        with nb.Batch(sum(len(c[2]) + 256 for c in cl), 8, flags=flags) as b:
            ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
            b.step(6)
            res.append([b.download(i) for i in ids])
    for a, c in zip(*res):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("bpl,js", [(1, 1), (2, 2), (4, 4), (1, 16), (2, 8)])
def test_batch_launch_shapes(nb, ora, bpl, js):
    cl = make_clusters(nb, 5, 256, 1500, 41 + bpl)
    with nb.Batch(sum(len(c[2]) + 256 for c in cl), 8, bodies_per_lane=bpl, j_split=js) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        for cid, (p, v, m, e, d) in zip(ids, cl):
            ref = ora.direct_f64(p.astype(np.float32).astype(np.float64), m.astype(np.float64), e)
            assert np.percentile(rel_err(b.forces(cid), ref), 99) < 1e-5


def test_tune_picks_a_shape_and_keeps_results(nb, ora):
    cl = make_clusters(nb, 10, 256, 4000, 17)
    with nb.Batch(sum(len(c[2]) + 256 for c in cl), 16) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        p0 = [b.download(cid)[0] for cid in ids]
        t = b.tune(reps=2)
        assert (t["bodies_per_lane"], t["j_split"]) in {(bb, k) for bb in (1, 2, 4) for k in (1, 2, 4)} | \
            {(1, 8), (1, 16), (2, 8)}
        assert t["direct_ms"] > 0 and t["n"] == sum(len(c[2]) for c in cl)
        # timing the shapes did not move any body
        for cid, p in zip(ids, p0):
            assert np.array_equal(b.download(cid)[0], p)
        b.step(3)
        for cid, (p, v, m, e, d) in zip(ids, cl):
            gp, _, _ = b.download(cid)
            op, ov, oa = p.copy(), v.copy(), np.zeros_like(p)
            for _ in range(3):
                f = ora.direct_f64(op.astype(np.float32).astype(np.float64), m.astype(np.float64), e)
                ora.kdk_update_f64(op, ov, oa, f, d)
            assert np.abs(gp - op).max() < 1e-9 * max(1.0, np.abs(op).max())


def test_quotas_split_steps_without_changing_trajectories(nb):
    """Clusters with step quotas advance min(quota, nsteps) steps per call; after
    three calls with changing quotas each cluster equals itself stepped the same
    number of times alone, bit for bit. A quota of 0 freezes a cluster."""
    cl = make_clusters(nb, 6, 256, 2000, 13)
    quotas = [[1, 4, 0, 8, nb.BATCH_UNLIMITED, 3], [2, 4, 5, 8, 1, 0], [8, 1, 1, 8, 8, 8]]
    with nb.Batch(sum(len(c[2]) + 256 for c in cl), 8) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        total = [0] * len(ids)
        for q in quotas:
            b.set_quotas(dict(zip(ids, q)))
            b.step(6)
            for k in range(len(ids)):
                total[k] += min(q[k], 6)
        b.sync()
        mixed = [b.download(cid) for cid in ids]
    assert total == [9, 9, 6, 18, 13, 9]
    for k, (p, v, m, e, d) in enumerate(cl):
        with nb.Batch(len(m) + 256, 1) as b:
            cid = b.add(p, v, m, e, d)
            b.step(total[k])
            alone = b.download(cid)
        for x, y in zip(alone, mixed[k]):
            assert np.array_equal(x, y), k


def test_c4_full_ensemble(nb, ora):
    """The batched-ensemble workload at its stated shape: 512 Plummer clusters,
    N_c log-uniform in [256, 32768], eps2 ~ U[1, 40], dt ~ U[0.001, 0.005], in
    one batch. For the largest, the median and the smallest cluster: forces
    against the FP64 oracle, and 5 steps in the mix equal to 5 steps alone, bit
    for bit."""
    seed, count = 4000, 512
    sizes = nb.log_uniform_sizes(seed, count, 256, 32768)
    eps2 = nb.uniform(seed + 1, count, 1.0, 40.0)
    dts = nb.uniform(seed + 2, count, 0.001, 0.005)
    cl = []
    for c in range(count):
        pos, vel, m = nb.ic_plummer(seed * 1000 + c, int(sizes[c]), 10.0, 0.1 * int(sizes[c]))
        cl.append((pos, vel, m, float(np.float32(eps2[c])), float(np.float32(dts[c]))))
    order = np.argsort(sizes, kind="stable")
    pick = [int(order[-1]), int(order[count // 2]), int(order[0])]
    print(f"C4: {int(sizes.sum())} bodies; sampled clusters of {[int(sizes[k]) for k in pick]}", flush=True)
    with nb.Batch(sum((len(c[2]) + 255) // 256 * 256 for c in cl), count) as b:
        ids = [b.add(p, v, m, e, d) for p, v, m, e, d in cl]
        _, bodies, inter = b.info()
        assert inter == sum(int(s) * (int(s) - 1) for s in sizes)
        for k in pick:
            p, v, m, e, d = cl[k]
            f = b.forces(ids[k])
            ref = ora.direct_f64(p.astype(np.float32).astype(np.float64), m.astype(np.float64), e)
            err = rel_err(f, ref)
            print(f"C4 cluster N={len(m)}: vs FP64 p50 {np.median(err):.2e} max {err.max():.2e}", flush=True)
            assert np.median(err) < 1e-6 and np.percentile(err, 99) < 1e-5
        b.step(5)
        b.sync()
        mixed = {k: b.download(ids[k]) for k in pick}
    for k in pick:
        p, v, m, e, d = cl[k]
        with nb.Batch(len(m) + 256, 1) as b:
            cid = b.add(p, v, m, e, d)
            b.step(5)
            alone = b.download(cid)
        for x, y in zip(alone, mixed[k]):
            assert np.array_equal(x, y), len(m)
