"""Pins the oracle's linear-octree restatement (oracle/linear_octree.cpp),
the FP64 direct sum, the metrics and the multipole formulas (oracle/direct.cpp).

No reference test holds key, sort or topology fixtures (the reference's GPU
tests only check replay and envelopes), so these are pinned by closed forms,
brute force on small grids and integer invariants.
"""
import numpy as np
import pytest


def spread_python(v):
    r = 0
    for k in range(21):
        r |= ((v >> k) & 1) << (3 * k)
    return r


def test_encode_closed_form(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rng.integers(0, 2 ** 21, 3)
        k = oracle.encode(int(q[0]), int(q[1]), int(q[2]), 0)
        assert k == (spread_python(int(q[0])) << 2) | (spread_python(int(q[1])) << 1) | \
            spread_python(int(q[2]))
        # decode
        back = [0, 0, 0]
        for b in range(21):
            back[0] |= ((k >> (3 * b + 2)) & 1) << b
            back[1] |= ((k >> (3 * b + 1)) & 1) << b
            back[2] |= ((k >> (3 * b)) & 1) << b
        assert back == [int(x) for x in q]
    assert oracle.encode(2 ** 21 - 1, 2 ** 21 - 1, 2 ** 21 - 1, 0) == 2 ** 63 - 1


def test_reference_key_keeps_low_ten_bits(oracle):
    # OctMortonCodes.inl:9-15 runs the interleave on 32-bit lanes with the
    # 10-bit masks: the key equals the full interleave of (q mod 1024).
    rng = np.random.default_rng(2)
    for _ in range(200):
        q = [int(v) for v in rng.integers(0, 2 ** 21, 3)]
        k1 = oracle.encode(q[0], q[1], q[2], 1)
        k0 = oracle.encode(q[0] & 1023, q[1] & 1023, q[2] & 1023, 0)
        assert k1 == k0 and k1 < 2 ** 30


def test_brute_force_morton_order_small_grid(oracle):
    # On an 8x8x8 grid, sorting by key equals sorting by the interleaved
    # binary digits (z-order): compare against an independent digit sort.
    pts = [(x, y, z) for x in range(8) for y in range(8) for z in range(8)]
    keys = [oracle.encode(x << 18, y << 18, z << 18, 0) for x, y, z in pts]

    def digits(p):
        x, y, z = p
        return tuple(((x >> b) & 1, (y >> b) & 1, (z >> b) & 1) for b in (2, 1, 0))
    by_key = [p for _, p in sorted(zip(keys, pts))]
    assert by_key == sorted(pts, key=digits)


def test_quantisation_frame(oracle):
    rng = np.random.default_rng(4)
    pos = rng.standard_normal((1000, 3))
    k, q = oracle.keys(pos, 0, with_q=True)
    for a in range(3):
        assert q[np.argmin(pos[:, a]), a] == 0
        assert q[np.argmax(pos[:, a]), a] == 2 ** 21 - 1
    # monotone along each axis
    for a in range(3):
        o = np.argsort(pos[:, a].astype(np.float32), kind="stable")
        assert np.all(np.diff(q[o, a].astype(np.int64)) >= 0)
    # key is the encode of q
    for i in range(0, 1000, 97):
        assert int(k[i]) == oracle.encode(int(q[i, 0]), int(q[i, 1]), int(q[i, 2]), 0)


def test_stable_sort(oracle):
    rng = np.random.default_rng(5)
    k = rng.integers(0, 50, 5000).astype(np.uint64)
    sk, order = oracle.stable_sort(k)
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(order, ref.astype(np.uint32))
    assert np.array_equal(sk, k[ref])


@pytest.mark.parametrize("n,cap,depth", [(65536, 256, 3), (1048576, 256, 4), (8388608, 256, 5),
                                         (16777216, 256, 6), (10, 256, 1), (1048576, 16, 6)])
def test_leaf_depth_rule(oracle, n, cap, depth):
    assert oracle.leaf_depth(n, cap) == depth


def _tree_inputs(oracle, n, seed=6, key_mode=0):
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((n, 3)) ** 3  # clustered
    m = (0.5 + rng.random(n)).astype(np.float32)
    k = oracle.keys(pos, key_mode)
    sk, order = oracle.stable_sort(k)
    return pos, m, sk, order


def test_reference_tree_invariants(oracle):
    n = 4000
    pos, m, sk, order = _tree_inputs(oracle, n)
    d = oracle.leaf_depth(n, 16)
    t = oracle.reference_tree(sk, order, pos, m, d)
    leaves = t["level"] == d
    assert t["count"][leaves].sum() == n
    # leaves tile [0, n) in order
    st = t["start"][leaves]
    assert st[0] == 0 and np.all(np.diff(st) == t["count"][leaves][:-1])
    # parent prefix relation and mass conservation
    for i in range(len(t["level"])):
        kids = [c for c in t["child"][i] if c >= 0]
        for c in kids:
            assert t["prefix"][c] >> 3 == t["prefix"][i]
            assert t["parent"][c] == i
        if kids:
            assert t["msum"][i] == pytest.approx(sum(t["msum"][c] for c in kids), rel=1e-12)
    root = len(t["level"]) - 1
    assert t["msum"][root] == pytest.approx(float(m.astype(np.float64).sum()), rel=1e-12)
    # next links: pre-order walk from the root visits every node once
    first_child = [next((c for c in t["child"][i] if c >= 0), -1) for i in range(root + 1)]
    seen, node = [], root
    while node >= 0:
        seen.append(node)
        node = first_child[node] if first_child[node] >= 0 else t["next"][node]
    assert sorted(seen) == list(range(root + 1))


def test_reference_tree_with_reference_keys_is_one_leaf(oracle):
    # With the 30-bit reference key every level-<=11 prefix is 0.
    n = 5000
    pos, m, sk, order = _tree_inputs(oracle, n, key_mode=1)
    d = oracle.leaf_depth(n, 256)
    t = oracle.reference_tree(sk, order, pos, m, d)
    assert len(t["level"]) == 1 and t["count"][0] == n


def test_hip_tree_rule(oracle):
    n = 6000
    pos, m, sk, order = _tree_inputs(oracle, n)
    d = oracle.leaf_depth(n, 16)
    b = 8
    t = oracle.hip_tree(sk, d, b)
    lvl, cnt, nc, fc = t["level"], t["count"], t["nchild"], t["first_child"]
    leaf = nc == 0
    assert cnt[leaf].sum() == n
    # leaves satisfy the rule; internal nodes were required to split
    assert np.all((cnt[leaf] <= b) | (lvl[leaf] >= d) | (lvl[leaf] == 21))
    assert np.all(lvl[leaf] >= d)
    for i in np.nonzero(~leaf)[0]:
        kids = np.arange(fc[i], fc[i] + nc[i])
        assert cnt[kids].sum() == cnt[i]
        assert np.all(t["prefix"][kids] >> 3 == t["prefix"][i])
        assert np.all(lvl[kids] == lvl[i] + 1)
    # breadth-first order = (level, prefix) order
    key = lvl.astype(np.int64) * 0 + lvl.astype(np.int64)
    assert np.all(np.diff(key) >= 0)
    for L in np.unique(lvl):
        assert np.all(np.diff(t["prefix"][lvl == L].astype(np.float64)) > 0)


def test_hip_tree_truncates_to_reference_tree(oracle):
    n = 6000
    pos, m, sk, order = _tree_inputs(oracle, n)
    d = oracle.leaf_depth(n, 16)
    ref = oracle.reference_tree(sk, order, pos, m, d)
    hip = oracle.hip_tree(sk, d, 8)
    rset = {(int(a), int(b), int(c), int(e)) for a, b, c, e in
            zip(ref["level"], ref["prefix"], ref["start"], ref["count"])}
    top = min(ref["level"])
    hset = {(int(a), int(b), int(c), int(e)) for a, b, c, e in
            zip(hip["level"], hip["prefix"], hip["start"], hip["count"]) if top <= a <= d}
    assert rset == hset


def test_direct_sum_closed_forms(oracle):
    pos = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    m = np.array([1.0, 3.0])
    a = oracle.direct_fp64(pos, m, [0, 1], eps=0.1)
    assert a[0, 0] == pytest.approx(3.0 * 2.0 / (4.01) ** 1.5, rel=1e-14)
    assert a[1, 0] == pytest.approx(-1.0 * 2.0 / (4.01) ** 1.5, rel=1e-14)
    # Newton's third law: sum m a = 0
    rng = np.random.default_rng(8)
    pos = rng.standard_normal((500, 3))
    m = rng.random(500)
    a = oracle.direct_fp64(pos, m, np.arange(500), eps=0.01)
    assert np.abs((m[:, None] * a).sum(0)).max() < 1e-10
    # a point inside a uniform shell feels no force (sampled shell, loose)
    rng = np.random.default_rng(9)
    v = rng.standard_normal((20000, 3))
    shell = v / np.linalg.norm(v, axis=1, keepdims=True)
    pts = np.vstack([[0.1, 0.2, -0.1], shell])
    mm = np.concatenate([[0.0], np.full(20000, 1 / 20000)])
    assert np.linalg.norm(oracle.direct_fp64(pts, mm, [0], eps=0.0)[0]) < 0.02


def test_force_error_metrics(oracle):
    ref = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 4.0]])
    approx = ref.copy()
    approx[0, 0] += 0.1
    approx[1, 1] += 0.4
    m = oracle.force_error(approx, ref)
    assert m["median"] == pytest.approx(0.1)
    assert m["max"] == pytest.approx(0.2)
    assert m["p99"] == pytest.approx(0.2)
    assert m["linf"] == pytest.approx(0.4 / 4.0)
    assert m["l2"] == pytest.approx(np.sqrt(0.01 + 0.16) / np.sqrt(1 + 4 + 16))


def test_moments_two_point_and_parallel_axis(oracle):
    pos = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    m = np.array([1.0, 1.0])
    mo = oracle.moments(pos, m)
    assert mo[0] == 2.0 and np.allclose(mo[1:4], 0)
    # Q_xx = sum m (3x^2 - r^2) = 2 * 2 = 4; Q_yy = Q_zz = -2; traceless
    assert np.allclose(mo[4:10], [4, 0, 0, -2, 0, -2])
    # shifting the set shifts only the COM
    mo2 = oracle.moments(pos + np.array([3.0, -2.0, 1.0]), m)
    assert np.allclose(mo2[4:], mo[4:]) and np.allclose(mo2[1:4], [3, -2, 1])


def test_multipole_convergence_orders(oracle):
    # Far field of a small random cluster: the monopole error falls like
    # (s/r)^2 (dipole vanishes about the COM), the quadrupole one like (s/r)^3.
    rng = np.random.default_rng(10)
    src = rng.standard_normal((50, 3)) * 0.5
    m = rng.random(50) + 0.2
    mo = oracle.moments(src, m)
    errs = {1: [], 2: []}
    for r in (8.0, 16.0, 32.0):
        x = mo[1:4] + np.array([r, 0.3 * r, -0.2 * r])
        allp = np.vstack([x, src])
        mm = np.concatenate([[0.0], m])
        exact = oracle.direct_fp64(allp, mm, [0], eps=0.0)[0]
        for order in (1, 2):
            a = oracle.multipole_accel(mo, x, 0.0, order)
            errs[order].append(np.linalg.norm(a - exact) / np.linalg.norm(exact))
    s1 = np.log2(errs[1][0] / errs[1][2]) / 2
    s2 = np.log2(errs[2][0] / errs[2][2]) / 2
    assert 1.7 < s1 < 2.3
    assert 2.6 < s2 < 3.4
    assert errs[2][0] < errs[1][0]
