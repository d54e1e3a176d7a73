"""GPU parity of the Barnes-Hut engine (drop-in for the reference's v7,
ParticleSolverBHutGPU.cpp:25-61) against the CPU build of the same pipeline
(oracle or_bh_*) and the FP64 direct sum, through the C ABI.

Bit-exact: Morton codes, sort permutation, octree topology (first_child, next,
body range, level). Tolerance: COM 1e-6 relative; forces 1e-5 (p99) relative to
the CPU build (both FP32, the GPU uses the hardware rsqrt); theta=0.5 error vs
FP64 no worse than the CPU build's."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

EPS2, DT = 40.0, 0.0035


def cpu_build(ora, pos4, leaf, cubic=True):
    mn, mx = ora.bbox(pos4)
    codes = ora.morton(pos4, mn, mx, cubic=cubic)
    perm, srt = ora.sort_codes(codes)
    tree = ora.build_tree(srt, leaf)
    com = ora.tree_com(tree, pos4[perm])
    return mn, mx, codes, perm, tree, com


def gpu_com(nodes):
    com = np.zeros((len(nodes), 4), np.float32)
    com[:, :3] = nodes["com"]
    com[:, 3] = nodes["mass"]
    return com


def gpu_run(nb, pos, vel, m, leaf, theta=0.5, flags=0):
    with nb.Simulation(len(m), EPS2, DT, solver=nb.SOLVER_BH, leaf_size=leaf, theta=theta,
                       flags=flags) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces()
        codes, perm = sim.debug_morton()
        nodes, bbox = sim.debug_tree()
        inter = sim.count_interactions()
    return f, codes, perm, nodes, bbox, inter


def pm4(pos, m):
    p = np.zeros((len(m), 4), np.float32)
    p[:, :3] = pos
    p[:, 3] = m
    return p


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_bit_exact_topology_and_forces(nb, ora, leaf):
    n = 20000
    pos, vel, m = nb.ic_plummer(300 + leaf, n, 50.0, 1e4)
    pos[7] = pos[8]  # coincident bodies share a code and a leaf
    p4 = pm4(pos, m)
    f, codes, perm, nodes, bbox, inter = gpu_run(nb, pos, vel, m, leaf)
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, leaf)
    np.testing.assert_array_equal(bbox[:3], mn)
    np.testing.assert_array_equal(bbox[3:], mx)
    np.testing.assert_array_equal(codes, ccodes)
    np.testing.assert_array_equal(perm, cperm)
    assert len(nodes) == len(tree["start"])
    np.testing.assert_array_equal(nodes["first_child"], tree["first_child"])
    np.testing.assert_array_equal(nodes["next"], tree["next"])
    np.testing.assert_array_equal(nodes["body_start"], tree["start"])
    np.testing.assert_array_equal(nodes["body_count"], tree["count"])
    np.testing.assert_array_equal(nodes["level"], tree["level"])
    np.testing.assert_allclose(nodes["mass"], com[:, 3], rtol=1e-6)
    np.testing.assert_allclose(nodes["com"], com[:, :3], rtol=1e-6, atol=1e-5)
    # opening radii: bit for bit the CPU formula applied to the GPU's own COMs,
    # and within rounding of the CPU build's
    srt = ccodes[cperm]
    r2 = ora.tree_open_r2(tree, gpu_com(nodes), srt, mn, mx, theta=0.5)
    np.testing.assert_array_equal(nodes["open_r2"], r2)
    np.testing.assert_allclose(nodes["open_r2"], ora.tree_open_r2(tree, com, srt, mn, mx),
                               rtol=1e-5)
    # default wave traversal vs its CPU twin (groups of NB_BH_WAVE_BODIES sorted bodies)
    fc, cinter = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, EPS2,
                                     group=nb.BH_WAVE_BODIES, codes_sorted=srt)
    e = rel_err(f, fc)
    print(f"leaf {leaf}: GPU vs CPU build p50={np.median(e):.2e} p99={np.percentile(e, 99):.2e}; "
          f"interactions {inter} vs {cinter}")
    assert np.percentile(e, 99) < 1e-5
    assert abs(inter - cinter) <= 1e-4 * cinter
    # the reference's per-body test (NB_F_BH_PER_BODY) vs the per-body CPU traversal
    fp, *_, pinter = gpu_run(nb, pos, vel, m, leaf, flags=nb.F_BH_PER_BODY)
    fcp, cpinter = ora.bh_forces(p4[cperm], cperm, mn, mx, tree, com, EPS2)
    assert np.percentile(rel_err(fp, fcp), 99) < 1e-5
    assert abs(pinter - cpinter) <= 1e-4 * cpinter
    assert inter >= pinter


def test_accuracy_vs_fp64(nb, ora):
    n = 50000
    pos, vel, m = nb.ic_plummer(11, n, 50.0, 1e4)
    tgt = np.random.default_rng(1).choice(n, 2000, replace=False)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2,
                         targets=tgt)
    errs = {}
    for name, flags in (("wave", 0), ("per-body", nb.F_BH_PER_BODY)):
        f, *_ = gpu_run(nb, pos, vel, m, 8, flags=flags)
        e = rel_err(f[tgt], ref)
        errs[name] = np.median(e)
        print(f"BH theta=0.5 N=50000 {name} vs FP64: p50={np.median(e):.2e} "
              f"p99={np.percentile(e, 99):.2e}")
        assert np.median(e) < 5e-3 and np.percentile(e, 99) < 3e-2
    assert errs["wave"] <= errs["per-body"]


def test_theta_tiny_is_direct_sum(nb, ora):
    n = 3000
    pos, vel, m = nb.ic_ball(4, n, 50.0, 1.0)
    f, *_, inter = gpu_run(nb, pos, vel, m, 8, theta=1e-7)
    assert inter == n * (n - 1)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2)
    assert rel_err(f, ref).max() < 1e-5


# 8193 and 65537 put the level scans and then the radix-histogram scans just
# past one workgroup's 8192 elements (two launches with a nearly empty tile)
@pytest.mark.parametrize("n", [1, 2, 5, 4097, 8193, 65537])
@pytest.mark.parametrize("pairs", [False, True])
def test_small_and_ragged(nb, ora, n, pairs):
    # below NB_BH_PAIR_MIN_BODIES ranges walk alone unless NB_F_BH_PAIRS
    pos, vel, m = nb.ic_ball(n, n, 10.0, 1.0)
    f, codes, perm, nodes, bbox, inter = gpu_run(nb, pos, vel, m, 4,
                                                 flags=nb.F_BH_PAIRS if pairs else 0)
    p4 = pm4(pos, m)
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, 4)
    np.testing.assert_array_equal(codes, ccodes)
    np.testing.assert_array_equal(perm, cperm)
    np.testing.assert_array_equal(nodes["next"], tree["next"])
    assert np.all(np.isfinite(f))
    if n == 1:
        assert np.all(f == 0)
        return
    # the pair walk against the CPU twin, which walks range by range
    fc, cinter = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, EPS2,
                                     group=nb.BH_WAVE_BODIES, codes_sorted=ccodes[cperm])
    assert inter == cinter
    assert np.percentile(rel_err(f, fc), 99) < 1e-5


@pytest.mark.parametrize("flags", [0, 0x20])
@pytest.mark.parametrize("n,leaf", [(50, 64), (64, 64), (33, 40)])
def test_root_leaf_with_several_ranges(nb, ora, n, leaf, flags):
    # more bodies than one range holds but no more than one leaf: the root is a
    # leaf cut into chunks (two full chunks make a pair under NB_F_BH_PAIRS =
    # 0x20), and every pair of bodies is evaluated
    pos, vel, m = nb.ic_ball(70 + n, n, 10.0, 1.0)
    f, codes, perm, nodes, bbox, inter = gpu_run(nb, pos, vel, m, leaf, flags=flags)
    assert len(nodes) == 1 and nodes["first_child"][0] == -1
    assert inter == n * (n - 1)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2)
    assert rel_err(f, ref).max() < 1e-5


@pytest.mark.parametrize("flags", [0, 0x20])
def test_unsoftened_with_coincident_bodies(nb, ora, flags):
    # eps2 = 0 selects the guarded traversal: a pair at distance zero adds
    # nothing (the CPU twin skips it), every other pair is unsoftened; single
    # ranges and (NB_F_BH_PAIRS = 0x20) pairs
    n = 20000
    pos, vel, m = nb.ic_plummer(51, n, 50.0, 1e4)
    pos[100:110] = pos[99]  # eleven bodies at one point
    p4 = pm4(pos, m)
    with nb.Simulation(n, 0.0, DT, solver=nb.SOLVER_BH, leaf_size=4, flags=flags) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces()
        inter = sim.count_interactions()
    assert np.all(np.isfinite(f))
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, 4)
    fc, cinter = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, 0.0,
                                     group=nb.BH_WAVE_BODIES, codes_sorted=ccodes[cperm])
    assert inter == cinter
    assert np.percentile(rel_err(f, fc), 99) < 1e-5


@pytest.mark.parametrize("ic,n", [("ball", 65537), ("discs", 65536), ("discs", 1 << 20)])
def test_single_ranges_equal_pairs(nb, ora, ic, n):
    # NB_F_BH_SINGLE_RANGES walks every range alone: the same interactions as
    # the pair walk, both equal to the range-by-range CPU twin (at 2^20 bodies
    # both take their work costliest first)
    pos, vel, m = nb.ic_ball(61, n, 100.0, 1.0) if ic == "ball" else nb.ic_discs(62, n)
    p4 = pm4(pos, m)
    fp, *_, ip = gpu_run(nb, pos, vel, m, 4, flags=nb.F_BH_PAIRS)
    fs, *_, is_ = gpu_run(nb, pos, vel, m, 4, flags=nb.F_BH_SINGLE_RANGES)
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, 4)
    fc, cinter = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, EPS2,
                                     group=nb.BH_WAVE_BODIES, codes_sorted=ccodes[cperm])
    assert ip == is_ == cinter
    assert np.percentile(rel_err(fs, fc), 99) < 1e-5
    assert np.percentile(rel_err(fp, fs), 99) < 1e-5


def test_deep_leaf_of_many_ranges(nb, ora):
    # 9000 coincident bodies make one level-10 leaf cut into 282 ranges (281 of
    # 32 bodies and one of 8): far more ranges of one width than a node's
    # children give, paired two by two; the rest is a ball around them
    n_pt, n_ball = 9000, 3000
    bp, _, bm = nb.ic_ball(77, n_ball, 50.0, 1.0)
    pos = np.concatenate([np.tile(np.array([[3.0, -2.0, 1.0]], np.float32), (n_pt, 1)), bp])
    m = np.concatenate([np.full(n_pt, 0.5, np.float32), bm])
    n = len(m)
    f, codes, perm, nodes, bbox, inter = gpu_run(nb, pos, np.zeros((n, 3)), m, 4,
                                                 flags=nb.F_BH_PAIRS)
    assert (nodes["body_count"][nodes["level"] == 10] >= n_pt).any()
    p4 = pm4(pos, m)
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, 4)
    np.testing.assert_array_equal(perm, cperm)
    fc, cinter = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, EPS2,
                                     group=nb.BH_WAVE_BODIES, codes_sorted=ccodes[cperm])
    assert inter == cinter
    assert np.percentile(rel_err(f, fc), 99) < 1e-5


def test_all_bodies_coincident(nb):
    n = 100
    pos = np.tile(np.array([[1.0, 2.0, 3.0]]), (n, 1))
    f, codes, perm, nodes, bbox, inter = gpu_run(nb, pos, np.zeros((n, 3)), np.ones(n, np.float32), 4)
    assert np.all(f == 0) and np.all(codes == 0)
    np.testing.assert_array_equal(perm, np.arange(n))
    assert nodes["body_count"][-1] == n and nodes["level"][-1] == 10


def test_steps_kdk_vs_cpu_pipeline(nb, ora):
    n = 8192
    pos, vel, m = nb.ic_plummer(21, n, 30.0, 1e3)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=8) as sim:
        sim.upload(pos, vel, m)
        sim.step(5)
        p, v, a = sim.download_f64()
    op, ov, oa = pos.copy(), vel.copy(), np.zeros_like(pos)
    for _ in range(5):
        f, *_ = ora.bh_pipeline(pm4(op.astype(np.float32), m), EPS2, theta=0.5, leaf_size=8,
                                group=nb.BH_WAVE_BODIES, cubic=True)
        ora.kdk_update_f64(op, ov, oa, f[:, :3].astype(np.float64), DT)
    assert np.abs(p - op).max() < 1e-6 * np.abs(op).max()
    assert rel_err(a, oa).max() < 1e-3


def test_determinism_and_graph(nb):
    n = 30000
    pos, vel, m = nb.ic_plummer(31, n, 50.0, 1e4)
    outs = []
    for chunks in ([4], [1, 1, 1, 1]):
        with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:
            sim.upload(pos, vel, m)
            for k in chunks:
                sim.step(k)
            outs.append(sim.download_f64())
    for x, y in zip(*outs):
        assert np.array_equal(x, y)


def test_forces_independent_of_traversal_order(nb):
    # the traversal takes its work items costliest first by the device time
    # they took in the previous pass: a pass with no history, one after a
    # pass of the same state, and one after passes of other states must agree
    # bit for bit
    n = 200000
    pos, vel, m = nb.ic_discs(9, n)
    # NB_F_BH_PAIRS: pairs, and with them the costliest-first order, below
    # NB_BH_PAIR_MIN_BODIES too
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=1, flags=nb.F_BH_PAIRS) as sim:
        sim.upload(pos, vel, m)
        f0 = sim.forces_f64()
        f1 = sim.forces_f64()
        sim.step(3)
        sim.upload(pos, vel, m)
        f2 = sim.forces_f64()
    assert np.array_equal(f0, f1) and np.array_equal(f0, f2)


@pytest.mark.parametrize("flags", [0, 0x20])
def test_walk_stack_overflow_is_reported(nb, flags):
    # the overflow path, forced from the host with a tiny stack region: every
    # wave stops its walk and exits normally, and the pass reports
    # NB_E_TREE_OVERFLOW (sticky until the next upload); single ranges and
    # (NB_F_BH_PAIRS = 0x20) pairs, whose deferred entries share the region
    n = 20000
    pos, vel, m = nb.ic_plummer(41, n, 50.0, 1e4)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=1, flags=flags) as sim:
        sim.upload(pos, vel, m)
        ref = sim.forces_f64()
        with pytest.raises(nb.NBodyError) as err:
            sim.debug_set_walk_stack(513)
        assert err.value.code == nb.NB_E_ARG
        sim.debug_set_walk_stack(4)
        with pytest.raises(nb.NBodyError) as err:
            sim.forces_f64()
        assert err.value.code == nb.NB_E_TREE_OVERFLOW
        sim.debug_set_walk_stack(0)
        sim.upload(pos, vel, m)
        assert np.array_equal(sim.forces_f64(), ref)


def test_c3_discs_properties(nb, ora):
    n = 1 << 22
    pos, vel, m = nb.ic_discs(2024, n)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=8) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
        codes, perm = sim.debug_morton()
        nodes, bbox = sim.debug_tree()
        sim.step(2)
        sim.sync()
    p4 = pm4(pos.astype(np.float32), m)
    mn, mx = ora.bbox(p4)
    np.testing.assert_array_equal(bbox[:3], mn)
    ccodes = ora.morton(p4, mn, mx, cubic=True)
    np.testing.assert_array_equal(codes, ccodes)
    cperm, srt = ora.sort_codes(ccodes)
    np.testing.assert_array_equal(perm, cperm)
    tree = ora.build_tree(srt, 8)
    np.testing.assert_array_equal(nodes["first_child"], tree["first_child"])
    np.testing.assert_array_equal(nodes["next"], tree["next"])
    np.testing.assert_array_equal(nodes["body_count"], tree["count"])
    assert abs(nodes["mass"][0] - m.astype(np.float64).sum()) < 1e-5 * m.sum()
    tgt = np.random.default_rng(5).choice(n, 512, replace=False)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2,
                         targets=tgt)
    e = rel_err(f[tgt], ref)
    print(f"C3 discs N=2^22 vs FP64 (512 targets): p50={np.median(e):.2e} p99={np.percentile(e, 99):.2e}")
    assert np.median(e) < 1e-2


def test_c3_mass_scaling_is_bitwise(nb):
    """At the full N = 2^22: the tree depends on positions only, and every mass
    term (FP64 prefix sums, node masses, the m r^-3 factor, the sums) scales by
    2 exactly, so doubling every mass doubles every force bit for bit."""
    n = 1 << 22
    pos, vel, m = nb.ic_discs(2025, n)

    def forces(mm):
        with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=16) as sim:
            sim.upload(pos, vel, mm)
            return sim.forces_f64()

    np.testing.assert_array_equal(forces(2.0 * m), 2.0 * forces(m))


def test_axis_box_matches_reference_quantisation(nb, ora):
    # NB_F_BH_AXIS_BOX: each axis over its own extent, the inputs of morton.glsl
    n = 6000
    pos, vel, m = nb.ic_discs(9, n)  # very flat: the axes' extents differ by ~20x
    p4 = pm4(pos, m)
    f, codes, perm, nodes, bbox, inter = gpu_run(nb, pos, vel, m, 8, flags=nb.F_BH_AXIS_BOX)
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, 8, cubic=False)
    np.testing.assert_array_equal(codes, ccodes)
    np.testing.assert_array_equal(perm, cperm)
    np.testing.assert_array_equal(nodes["next"], tree["next"])
    srt = ccodes[cperm]
    r2 = ora.tree_open_r2(tree, gpu_com(nodes), srt, mn, mx, theta=0.5, cubic=False)
    np.testing.assert_array_equal(nodes["open_r2"], r2)
    fc, cinter = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, EPS2,
                                     group=nb.BH_WAVE_BODIES, codes_sorted=srt, cubic=False)
    assert np.percentile(rel_err(f, fc), 99) < 1e-5
    assert inter == cinter
    # the same boxes walked in pairs: the same interactions
    fp, *_, ip = gpu_run(nb, pos, vel, m, 8, flags=nb.F_BH_AXIS_BOX | nb.F_BH_PAIRS)
    assert ip == cinter and np.percentile(rel_err(fp, fc), 99) < 1e-5
    # the cubic default opens fewer cells on flat systems
    _, _, _, _, _, inter_cubic = gpu_run(nb, pos, vel, m, 8)
    assert inter_cubic < inter


def test_largest_tested_size_2_24(nb, ora):
    """N = 2^24 Plummer sphere (four times the C3 size): Morton codes, permutation
    and octree topology bit-exact against the CPU build, the root mass equal to
    the total mass, and theta=0.5 forces within the C3 accuracy on seeded targets.
    Exercises the multi-block scans and a sort of 8192 tiles."""
    n = 1 << 24
    pos, vel, m = nb.ic_plummer(2424, n, 50.0, 1e4)
    print(f"N=2^24: uploading {n} bodies", flush=True)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=8) as sim:
        sim.upload(pos, vel, m)
        f = sim.forces_f64()
        codes, perm = sim.debug_morton()
        nodes, bbox = sim.debug_tree()
    print(f"N=2^24: {len(nodes)} nodes; checking against the CPU build", flush=True)
    p4 = pm4(pos.astype(np.float32), m)
    mn, mx = ora.bbox(p4)
    np.testing.assert_array_equal(bbox[:3], mn)
    np.testing.assert_array_equal(bbox[3:], mx)
    ccodes = ora.morton(p4, mn, mx, cubic=True)
    np.testing.assert_array_equal(codes, ccodes)
    cperm, srt = ora.sort_codes(ccodes)
    np.testing.assert_array_equal(perm, cperm)
    del ccodes, codes
    tree = ora.build_tree(srt, 8)
    assert len(nodes) == len(tree["start"])
    np.testing.assert_array_equal(nodes["first_child"], tree["first_child"])
    np.testing.assert_array_equal(nodes["next"], tree["next"])
    np.testing.assert_array_equal(nodes["body_start"], tree["start"])
    np.testing.assert_array_equal(nodes["body_count"], tree["count"])
    np.testing.assert_array_equal(nodes["level"], tree["level"])
    assert abs(nodes["mass"][0] - m.astype(np.float64).sum()) < 1e-5 * m.sum()
    tgt = np.random.default_rng(6).choice(n, 64, replace=False)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2,
                         targets=tgt)
    e = rel_err(f[tgt], ref)
    print(f"N=2^24 Plummer vs FP64 (64 targets): p50={np.median(e):.2e} p99={np.percentile(e, 99):.2e}")
    assert np.median(e) < 1e-2


def test_fp64_reference_on_device(nb, ora):
    """nb_direct_f64 (the device FP64 direct sum the bench prices errors with)
    against the CPU oracle's FP64 direct sum on the same FP32 positions, before
    and after a Morton reorder (targets are caller indices)."""
    n = 20000
    pos, vel, m = nb.ic_plummer(77, n, 50.0, 1e4)
    tg = np.random.default_rng(8).choice(n, 500, replace=False)
    ref = ora.direct_f64(pos.astype(np.float32).astype(np.float64), m.astype(np.float64), EPS2,
                         targets=tg)
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH) as sim:
        sim.upload(pos, vel, m)
        a = sim.direct_f64(tg)
        sim.reorder()
        b = sim.direct_f64(tg)
    assert rel_err(a, ref).max() < 1e-12
    assert rel_err(b, ref).max() < 1e-12


def test_per_body_count_and_leaf_tuning(nb, ora):
    n = 30000
    pos, vel, m = nb.ic_discs(78, n)
    p4 = pm4(pos, m)
    mn, mx, ccodes, cperm, tree, com = cpu_build(ora, p4, 8)
    _, cpb = ora.bh_forces_group(p4[cperm], cperm, mn, mx, tree, com, EPS2, group=1,
                                 codes_sorted=ccodes[cperm])
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=8) as sim:
        sim.upload(pos, vel, m)
        performed = sim.count_interactions()
        per_body = sim.count_interactions_per_body()
        print(f"discs N={n} leaf 8: performed {performed}, per body {per_body} (CPU {cpb}), "
              f"ratio {performed / per_body:.3f}")
        assert abs(per_body - cpb) <= 1e-4 * cpb
        assert performed >= per_body
        t = sim.tune_leaf(2)
        assert t["leaf_size"] in (1, 2, 4, 8, 16, 32) and t["bh_ms"] > 0
        assert t["bh_single_ranges"] in (0, 1)
        f_tuned = sim.forces_f64()
        sim.step(2)
        sim.sync()
    flags = nb.F_BH_SINGLE_RANGES if t["bh_single_ranges"] else nb.F_BH_PAIRS
    with nb.Simulation(n, EPS2, DT, solver=nb.SOLVER_BH, leaf_size=t["leaf_size"],
                       flags=flags) as ref:
        ref.upload(pos, vel, m)
        np.testing.assert_array_equal(ref.forces_f64(), f_tuned)
