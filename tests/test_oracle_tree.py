"""Pins the oracle's Morton encode, sort, octree, COM and Barnes-Hut traversal
(morton.glsl:37-98; Octree.cpp:9-350) with exhaustive checks, brute force and
limits, and the v5 restatement with theta -> 0 and monopole limits."""
import numpy as np
import pytest

from helpers import mass4, rel_err, seeded_ball, vec4


def naive_spread(v):
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def test_expand_bits_exhaustive(ora):
    for v in range(1024):
        assert ora.expand_bits(v) == naive_spread(v)


def test_morton_corners_and_digit_order(ora):
    pos = vec4(np.array([[0.0, 0, 0], [10.0, 10, 10], [10.0, 0, 0], [0.0, 10, 0], [0.0, 0, 10],
                         [5.0, 5, 5], [4.999, 0, 0]]))
    mn, mx = ora.bbox(pos)
    np.testing.assert_array_equal(mn, [0, 0, 0])
    np.testing.assert_array_equal(mx, [10, 10, 10])
    c = ora.morton(pos, mn, mx)
    assert c[0] == 0
    assert c[1] == 0x3FFFFFFF  # the max corner clamps to cell 1023 on every axis
    assert c[2] == naive_spread(1023) << 2  # x is the most significant bit of each triplet
    assert c[3] == naive_spread(1023) << 1
    assert c[4] == naive_spread(1023)
    assert c[5] == naive_spread(512) * 7  # 5/10 * 1024 = 512 on every axis
    assert c[6] == naive_spread(511) << 2  # floor(0.4999 * 1024) = 511


def test_morton_degenerate_extent(ora):
    # all x equal: extent clamps to 1e-6, (p - min) / 1e-6 = 0 -> cell 0
    pos = vec4(np.array([[3.0, 0, 0], [3.0, 1, 0], [3.0, 2, 1]]))
    mn, mx = ora.bbox(pos)
    c = ora.morton(pos, mn, mx)
    assert np.all((c >> 2) & naive_spread(1023) == 0)


def test_bbox_all_negative_axis(ora):
    # the reference's max starts at numeric_limits<float>::min() (Octree.cpp:27); the
    # restatement starts at -FLT_MAX so an all-negative axis keeps its true maximum
    pos = vec4(np.array([[-3.0, -1, -2], [-1.0, -5, -4]]))
    mn, mx = ora.bbox(pos)
    np.testing.assert_array_equal(mx, [-1, -1, -2])
    np.testing.assert_array_equal(mn, [-3, -5, -4])


def test_sort_matches_lexicographic(ora):
    rng = np.random.default_rng(2)
    codes = rng.integers(0, 64, 5000).astype(np.uint32)  # many ties
    perm, srt = ora.sort_codes(codes)
    ref = np.lexsort((np.arange(5000), codes))
    np.testing.assert_array_equal(perm, ref)
    np.testing.assert_array_equal(srt, codes[ref])


def brute_tree(codes, leaf):
    """Recursive digit split, numbered breadth-first afterwards (independent of
    the oracle's level loop)."""
    nodes = []  # (level, start, count, children list)

    def split(level, s, e):
        node = {"level": level, "start": s, "count": e - s, "kids": []}
        if e - s > leaf and level < 10:
            shift = 27 - 3 * level
            b = s
            while b < e:
                d = (int(codes[b]) >> shift) & 7
                c = b
                while c < e and ((int(codes[c]) >> shift) & 7) == d:
                    c += 1
                node["kids"].append(split(level + 1, b, c))
                b = c
        return node

    root = split(0, 0, len(codes))
    order, queue = [], [root]
    while queue:
        nd = queue.pop(0)
        order.append(nd)
        queue.extend(nd["kids"])
    idx = {id(nd): i for i, nd in enumerate(order)}
    nxt = {id(root): -1}
    for nd in order:
        for q, k in enumerate(nd["kids"]):
            nxt[id(k)] = idx[id(nd["kids"][q + 1])] if q + 1 < len(nd["kids"]) else nxt[id(nd)]
    return {
        "first_child": np.array([idx[id(nd["kids"][0])] if nd["kids"] else -1 for nd in order]),
        "next": np.array([nxt[id(nd)] for nd in order]),
        "start": np.array([nd["start"] for nd in order]),
        "count": np.array([nd["count"] for nd in order]),
        "level": np.array([nd["level"] for nd in order]),
    }


@pytest.mark.parametrize("leaf", [1, 3, 8])
def test_tree_matches_brute_force(ora, leaf):
    pos = vec4(seeded_ball(700, seed=leaf))
    pos[10] = pos[11]  # two coincident bodies share a code
    mn, mx = ora.bbox(pos)
    perm, srt = ora.sort_codes(ora.morton(pos, mn, mx))
    t = ora.build_tree(srt, leaf)
    b = brute_tree(srt, leaf)
    for k in b:
        np.testing.assert_array_equal(t[k], b[k], err_msg=k)


def rope_walk(t):
    seen, i = [], 0
    while i >= 0:
        seen.append(i)
        i = t["first_child"][i] if t["first_child"][i] >= 0 else t["next"][i]
    return seen


def test_tree_invariants_and_com(ora):
    rng = np.random.default_rng(11)
    pos = vec4(seeded_ball(3000, seed=12))
    pos[:, 3] = rng.uniform(0.5, 2.0, 3000)
    mn, mx = ora.bbox(pos)
    perm, srt = ora.sort_codes(ora.morton(pos, mn, mx))
    pos_s = pos[perm]
    t = ora.build_tree(srt, 4)
    leaves = t["first_child"] < 0
    # every body in exactly one leaf
    cover = np.zeros(3000, int)
    for s, c in zip(t["start"][leaves], t["count"][leaves]):
        cover[s:s + c] += 1
    assert np.all(cover == 1)
    # the rope walk (descend everywhere) visits every node exactly once, in pre-order
    seen = rope_walk(t)
    assert sorted(seen) == list(range(len(t["start"])))
    com = ora.tree_com(t, pos_s)
    for k in range(0, len(t["start"]), 7):
        s, c = t["start"][k], t["count"][k]
        m = pos_s[s:s + c, 3].astype(np.float64)
        x = pos_s[s:s + c, :3].astype(np.float64)
        assert abs(com[k, 3] - m.sum()) <= 1e-6 * m.sum()
        np.testing.assert_allclose(com[k, :3], (m[:, None] * x).sum(0) / m.sum(), rtol=1e-6, atol=1e-5)
    assert abs(com[0, 3] - pos[:, 3].astype(np.float64).sum()) < 1e-3


def test_bh_theta_zero_is_direct(ora):
    pos = vec4(seeded_ball(800, seed=13))
    pos[:, 3] = 1.0
    f, codes, perm, nodes, inter = ora.bh_pipeline(pos, 4.0, theta=1e-7, leaf_size=4)
    ref = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(800), 4.0)
    assert inter == 800 * 799
    assert rel_err(f, ref).max() < 1e-5


def test_bh_distant_cluster_monopole(ora):
    # Cluster B (16 bodies within radius 1) seen from cluster A 1000 units away is
    # accepted as one cell at theta = 0.5: force = M_B (c_B - x) / (d^2 + eps2)^1.5.
    rng = np.random.default_rng(14)
    a = rng.normal(scale=0.5, size=(16, 3))
    b = rng.normal(scale=0.5, size=(16, 3)) + np.array([1000.0, 0, 0])
    pos = vec4(np.vstack([a, b]))
    pos[:, 3] = 1.0
    f, _, _, _, _ = ora.bh_pipeline(pos, 0.01, theta=0.5, leaf_size=1)
    cb = b.astype(np.float32).astype(np.float64).mean(0)
    for i in range(16):
        d = cb - pos[i, :3].astype(np.float64)
        r2 = d @ d + 0.01
        mono = 16 * d / r2 ** 1.5
        near = ora.direct_f64(a.astype(np.float32).astype(np.float64), np.ones(16), 0.01)[i]
        np.testing.assert_allclose(f[i, :3], near + mono, rtol=1e-4, atol=1e-9)


def test_bh_cpu_build_accuracy(ora):
    # theta = 0.5 monopole error of the engine's tree is no worse than the v5
    # restatement's on the same input (its leaves are evaluated body by body).
    pos = vec4(seeded_ball(4000, seed=15))
    pos[:, 3] = 1.0
    ref = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(4000), 40.0)
    rc, f5, _ = ora.v5_forces(pos, mass4(np.ones(4000)), 40.0, theta=0.5)
    e5 = rel_err(f5, ref)
    for leaf in (1, 8):
        f, _, _, _, inter = ora.bh_pipeline(pos, 40.0, theta=0.5, leaf_size=leaf)
        e = rel_err(f, ref)
        assert np.median(e) < 1e-2 and np.percentile(e, 99) < 3e-2
        assert np.median(e) <= 1.02 * np.median(e5)
        assert inter < 4000 * 3999 / 5


def test_v5_theta_zero_is_direct(ora):
    pos = vec4(seeded_ball(600, seed=16))
    m = np.random.default_rng(17).uniform(0.5, 2.0, 600)
    rc, f, used = ora.v5_forces(pos, mass4(m), 4.0, theta=1e-7)
    assert rc == 0 and used > 600
    ref = ora.direct_f64(pos[:, :3].astype(np.float64), m.astype(np.float32).astype(np.float64), 4.0)
    assert rel_err(f, ref).max() < 1e-5


def test_v5_accuracy_and_agreement(ora):
    pos = vec4(seeded_ball(3000, seed=18))
    mm = mass4(np.ones(3000))
    rc, f5, _ = ora.v5_forces(pos, mm, 40.0, theta=0.5)
    assert rc == 0
    ref = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(3000), 40.0)
    e5 = rel_err(f5, ref)
    assert np.median(e5) < 1e-2
    pm = pos.copy()
    pm[:, 3] = 1.0  # the engine's layout carries the mass in w
    fb, *_ = ora.bh_pipeline(pm, 40.0, theta=0.5, leaf_size=1)
    # same opening rule and cell geometry (up to quantisation): forces agree closely
    assert np.median(rel_err(fb, f5)) < 5e-3


def test_v5_monopole_and_overflow(ora):
    rng = np.random.default_rng(19)
    a = rng.normal(scale=0.5, size=(8, 3))
    b = rng.normal(scale=0.5, size=(8, 3)) + np.array([0, 500.0, 0])
    pos = vec4(np.vstack([a, b]))
    rc, f, _ = ora.v5_forces(pos, mass4(np.ones(16)), 0.01, theta=0.5)
    assert rc == 0
    cb = b.astype(np.float32).astype(np.float64).mean(0)
    d = cb - pos[0, :3].astype(np.float64)
    mono = 8 * d / (d @ d + 0.01) ** 1.5
    near = ora.direct_f64(a.astype(np.float32).astype(np.float64), np.ones(8), 0.01)[0]
    np.testing.assert_allclose(f[0, :3], near + mono, rtol=1e-4, atol=1e-9)
    # a pool of one node per body overflows: reported, not silent (reference D5)
    rc, _, _ = ora.v5_forces(pos, mass4(np.ones(16)), 0.01, theta=0.5, pool_factor=1.0)
    assert rc == -5


def test_v6_matches_v5(ora):
    # same geometric subdivision, opening rule and child order: the parallel
    # restatement's forces equal the sequential one's (no close pairs to merge)
    pos = vec4(seeded_ball(5000, seed=20))
    mm = mass4(np.random.default_rng(21).uniform(0.5, 2.0, 5000))
    rc5, f5, _ = ora.v5_forces(pos, mm, 40.0, theta=0.5)
    rc6, f6, tb, tf = ora.v6_forces(pos, mm, 40.0, theta=0.5, nthreads=4)
    assert rc5 == 0 and rc6 == 0
    assert rel_err(f6, f5).max() < 1e-6
    # a target range computes only those bodies
    _, part, _, _ = ora.v6_forces(pos, mm, 40.0, theta=0.5, t0=100, t1=200, nthreads=2)
    np.testing.assert_array_equal(part[100:200], f6[100:200])
    assert np.all(part[:100] == 0)


def test_v6_small_inputs(ora):
    for n in (1, 2, 9, 70):
        pos = vec4(seeded_ball(n, seed=n))
        rc, f, _, _ = ora.v6_forces(pos, mass4(np.ones(n)), 1.0, theta=0.5, nthreads=2)
        ref = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(n), 1.0)
        assert rc == 0 and np.all(np.isfinite(f))
        if n <= 9:
            assert rel_err(f, ref).max() < 2e-2


def test_group_traversal_limits_and_accuracy(ora):
    pos = vec4(seeded_ball(3000, seed=22))
    pos[:, 3] = 1.0
    ref = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(3000), 40.0)
    # theta -> 0: every leaf, every pair
    f0, *_, inter0 = ora.bh_pipeline(pos, 40.0, theta=1e-7, leaf_size=8, group=64)
    assert inter0 == 3000 * 2999 and rel_err(f0, ref).max() < 1e-5
    # the modified test accepts, per body, a subset of what the reference's test
    # accepts on the same tree: more pairs, and at least as accurate
    for cubic in (False, True):
        fb, *_, ib = ora.bh_pipeline(pos, 40.0, theta=0.5, leaf_size=8, group=0, cubic=cubic)
        f1, *_, i1 = ora.bh_pipeline(pos, 40.0, theta=0.5, leaf_size=8, group=1, cubic=cubic)
        fg, *_, ig = ora.bh_pipeline(pos, 40.0, theta=0.5, leaf_size=8, group=64, cubic=cubic)
        assert ig >= i1 > ib
        e_b, e_1, e_g = (np.median(rel_err(f, ref)) for f in (fb, f1, fg))
        assert e_g <= e_1 < e_b


def test_modified_test_is_a_subset_of_the_reference_test(ora):
    """Per body and per cell: whenever the modified test accepts a cell
    ((s / theta + delta)^2 < d^2), the reference's test (s^2 < theta^2 d^2,
    Octree.cpp:320) accepts it too, since delta >= 0."""
    n = 2000
    pos = vec4(seeded_ball(n, seed=41))
    pos[:, 3] = 1.0
    mn, mx = ora.bbox(pos)
    perm, srt = ora.sort_codes(ora.morton(pos, mn, mx, cubic=True))
    t = ora.build_tree(srt, 4)
    com = ora.tree_com(t, pos[perm])
    r2 = ora.tree_open_r2(t, com, srt, mn, mx, theta=0.5)
    side = float(np.max(np.maximum(mx - mn, 1e-6)))
    s2 = (side / 2.0 ** t["level"]) ** 2
    d2 = ((com[None, :, :3].astype(np.float64) - pos[:200, None, :3]) ** 2).sum(-1)
    accept_mod = r2[None, :] < d2
    accept_ref = s2[None, :] < 0.25 * d2
    assert accept_mod.any() and not (accept_mod & ~accept_ref).any()


def test_compact_bits_inverts_expand_bits(ora):
    for v in range(1024):
        assert ora.compact_bits(ora.expand_bits(v)) == v
    # and it reads every third bit only
    assert ora.compact_bits(0xFFFFFFFF) == 0x3FF
    assert ora.compact_bits(0x92492492) == 0 and ora.compact_bits(0x24924924) == 0


@pytest.mark.parametrize("cubic", [True, False])
def test_open_radius_from_the_cell_geometry(ora, cubic):
    """r^2 = (s_l / theta + |COM - centre|)^2 with the centre found
    independently: the cell of level l holding a node's first body is
    floor(q / 2^(10-l)) with q the body's 10-bit quantised coordinate. Every
    body of the node lies in that cell."""
    rng = np.random.default_rng(42)
    n = 3000
    pos = vec4(rng.uniform(-1, 1, (n, 3)) * np.array([80.0, 30.0, 5.0]))
    pos[:, 3] = rng.uniform(0.5, 2.0, n)
    pos[::97, 3] = 0.0  # a few massless bodies
    pos[5:25, :3] = pos[5, :3]  # and one massless-free coincident clump
    mn, mx = ora.bbox(pos)
    perm, srt = ora.sort_codes(ora.morton(pos, mn, mx, cubic=cubic))
    t = ora.build_tree(srt, 3)
    ps = pos[perm]
    com = ora.tree_com(t, ps)
    theta = 0.6
    r2 = ora.tree_open_r2(t, com, srt, mn, mx, theta=theta, cubic=cubic)
    ext = np.maximum(mx.astype(np.float64) - mn, 1e-6)
    side = ext.max()
    ext = np.full(3, side) if cubic else ext
    q = np.clip(np.floor(np.clip((ps[:, :3] - mn) / ext, 0, 1) * 1024), 0, 1023).astype(np.int64)
    for k in range(len(t["start"])):
        lv, st, c = t["level"][k], t["start"][k], t["count"][k]
        if com[k, 3] <= 0:
            assert r2[k] == -1.0
            continue
        idx = q[st] >> (10 - lv)
        assert np.all(q[st:st + c] >> (10 - lv) == idx)  # the node's bodies share the cell
        h = ext / 2.0 ** lv
        ctr = mn + (idx + 0.5) * h
        want = (side / 2.0 ** lv / theta + np.linalg.norm(com[k, :3] - ctr)) ** 2
        assert abs(r2[k] - want) <= 1e-5 * want


def test_root_leaf_is_cut_into_ranges(ora):
    """Fewer bodies than the leaf size but more than one range: the root is a
    leaf and is cut into chunks; the group traversal then evaluates every pair."""
    n = 50
    pos = vec4(seeded_ball(n, seed=43))
    pos[:, 3] = 1.0
    mn, mx = ora.bbox(pos)
    perm, srt = ora.sort_codes(ora.morton(pos, mn, mx, cubic=True))
    t = ora.build_tree(srt, 64)
    assert len(t["start"]) == 1 and t["first_child"][0] == -1
    rs, rc = ora.bh_wave_ranges(t, n, 32)
    assert list(rs) == [0, 32] and list(rc) == [32, 18]
    ref = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(n), 40.0)
    f, *_, inter = ora.bh_pipeline(pos, 40.0, theta=0.5, leaf_size=64, group=32, cubic=True)
    assert inter == n * (n - 1) and rel_err(f, ref).max() < 1e-5


def test_cubic_box_is_the_reference_formula_on_a_cube(ora):
    # with min at the origin, the cubic box is morton.glsl's formula with
    # max = (E, E, E), E the largest extent
    rng = np.random.default_rng(23)
    pos = vec4(rng.uniform(0, 1, (2000, 3)) * np.array([100.0, 10.0, 1.0]))
    pos[0, :3] = 0
    pos[1, :3] = [100.0, 10.0, 1.0]
    mn, mx = ora.bbox(pos)
    cub = ora.morton(pos, mn, mx, cubic=True)
    ref = ora.morton(pos, mn, np.float32([100.0, 100.0, 100.0]), cubic=False)
    np.testing.assert_array_equal(cub, ref)
    assert not np.array_equal(cub, ora.morton(pos, mn, mx, cubic=False))
    # every cubic-box pipeline limit still holds
    pos[:, 3] = 1.0
    f0, *_, inter0 = ora.bh_pipeline(pos, 1.0, theta=1e-7, leaf_size=8, group=64, cubic=True)
    ref_f = ora.direct_f64(pos[:, :3].astype(np.float64), np.ones(2000), 1.0)
    assert inter0 == 2000 * 1999 and rel_err(f0, ref_f).max() < 1e-5


def test_wave_ranges_partition_the_bodies(ora):
    rng = np.random.default_rng(24)
    pos = vec4(seeded_ball(5000, seed=25))
    pos[100:300, :3] = pos[100, :3]  # 200 coincident bodies: one unsplittable leaf
    mn, mx = ora.bbox(pos)
    perm, srt = ora.sort_codes(ora.morton(pos, mn, mx, cubic=True))
    t = ora.build_tree(srt, 8)
    rs, rc = ora.bh_wave_ranges(t, 5000, 64)
    assert rc.min() >= 1 and rc.max() <= 64
    cover = np.zeros(5000, int)
    for a, c in zip(rs, rc):
        cover[a:a + c] += 1
    assert np.all(cover == 1)
    # each range lies in one cell: find the deepest node holding the whole range;
    # the range's codes share that node's prefix, and unless the node is a leaf
    # (a chunk of it), the range is a run of whole children of a node that holds
    # more than the capacity
    fc, st, cnt, lv = t["first_child"], t["start"], t["count"], t["level"]
    for a, c in zip(rs, rc):
        k = 0
        while fc[k] >= 0:
            j = fc[k]
            inside = None
            while j < len(st) and lv[j] == lv[k] + 1 and st[j] < st[k] + cnt[k]:
                if st[j] <= a and a + c <= st[j] + cnt[j]:
                    inside = j
                j += 1
            if inside is None:
                break
            k = inside
        shift = 30 - 3 * int(lv[k])
        assert (int(srt[a]) >> shift) == (int(srt[a + c - 1]) >> shift)
        if a == st[k] and c == cnt[k]:
            continue  # the range is one whole cell
        assert cnt[k] > 64  # a part of a larger node: a run of its children, or a chunk of a leaf
        if fc[k] >= 0:
            j, starts, ends = fc[k], set(), set()
            while j < len(st) and lv[j] == lv[k] + 1 and st[j] < st[k] + cnt[k]:
                starts.add(int(st[j]))
                ends.add(int(st[j] + cnt[j]))
                j += 1
            assert int(a) in starts and int(a + c) in ends
    # a small system is one range; capacity 1 gives singletons
    rs1, rc1 = ora.bh_wave_ranges(ora.build_tree(srt[:50], 8), 50, 64)
    assert list(rs1) == [0] and list(rc1) == [50]
    rs2, rc2 = ora.bh_wave_ranges(t, 5000, 1)
    assert len(rs2) == 5000 and np.all(rc2 == 1)
