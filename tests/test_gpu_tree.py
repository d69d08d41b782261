"""GPU parity of keys, sort order and tree topology (bit-exact) and of the
node moments (FP32 vs the oracle's FP64), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _solver(n, **kw):
    import blitzar_amd as B
    return B.OctreeHip(n, **kw)


def _clustered(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((n, 3)) ** 3 * np.array([1.0, 0.7, 1.3])
    m = (0.5 + rng.random(n)).astype(np.float32) / n
    return pos, m


@pytest.mark.parametrize("n,key_mode", [(1, 0), (2, 0), (63, 0), (64, 0), (65, 0), (3000, 0),
                                        (50000, 0), (3000, 1), (50000, 1)])
def test_keys_and_sorted_order_bit_exact(oracle, n, key_mode):
    say(f"case keys n={n} key_mode={key_mode}")
    pos, m = _clustered(n, 100 + n)
    with _solver(n, key_mode=key_mode, leaf_capacity=16, bucket_capacity=8) as s:
        s.upload(pos, None, m)
        s.build_tree()
        t = s.dump_tree()
    k = oracle.keys(pos, key_mode)
    sk, order = oracle.stable_sort(k)
    assert np.array_equal(t["sorted_keys"], sk)
    assert np.array_equal(t["sorted_ids"], order)


def test_sorted_order_ties_by_particle_index(oracle):
    # Coincident particles tie on the key; after a step the store is in Morton
    # order, and the tie order must still follow the caller index.
    n = 4000
    say(f"case ties n={n}")
    rng = np.random.default_rng(7)
    base = rng.standard_normal((n // 4, 3))
    pos = np.repeat(base, 4, axis=0)[rng.permutation(n)]
    m = np.full(n, 1.0 / n, np.float32)
    with _solver(n, leaf_capacity=16, bucket_capacity=8) as s:
        s.upload(pos, None, m)
        s.build_tree()
        s.build_tree()  # second build starts from the permuted store
        t = s.dump_tree()
    sk, order = oracle.stable_sort(oracle.keys(pos, 0))
    assert np.array_equal(t["sorted_keys"], sk)
    assert np.array_equal(t["sorted_ids"], order)


@pytest.mark.parametrize("n,leaf_cap,bucket", [(1, 256, 16), (2, 256, 16), (5, 256, 16),
                                               (3000, 16, 8), (4096, 16, 16), (4097, 16, 3),
                                               (40000, 256, 16), (40000, 16, 1), (40000, 64, 64),
                                               (40000, 16, 65), (40000, 16, 200),
                                               (1500000, 16, 1)])
def test_topology_bit_exact(oracle, n, leaf_cap, bucket):
    say(f"case topology n={n} leaf_cap={leaf_cap} bucket={bucket}")
    pos, m = _clustered(n, 200 + n)
    with _solver(n, leaf_capacity=leaf_cap, bucket_capacity=bucket) as s:
        s.upload(pos, None, m)
        s.build_tree()
        t = s.dump_tree()
    sk, _ = oracle.stable_sort(oracle.keys(pos, 0))
    d = oracle.leaf_depth(n, leaf_cap)
    assert t["leaf_depth"] == d
    ref = oracle.hip_tree(sk, d, bucket)
    assert len(t["level"]) == len(ref["level"])
    for key in ("level", "prefix", "start", "count", "first_child", "nchild"):
        assert np.array_equal(t[key].astype(np.int64), ref[key].astype(np.int64)), key


def test_topology_bit_exact_hernquist_merger_8388608(oracle):
    """The benchmark state: 8,388,608 particles in two Hernquist halos (2,048
    topology tiles, about 2.1 million nodes down to level 17), the whole tree
    against the oracle."""
    import blitzar_amd as B
    n = 8388608
    say(f"case topology hernquist merger n={n}")
    pos, vel, m = B.hernquist_merger(n // 2, seed=11, separation=20.0, pericentre=2.0)
    with _solver(n, leaf_capacity=16, bucket_capacity=16) as s:
        s.upload(pos, vel, m)
        s.build_tree()
        t = s.dump_tree()
    sk, order = oracle.stable_sort(oracle.keys(pos, 0))
    assert np.array_equal(t["sorted_keys"], sk)
    assert np.array_equal(t["sorted_ids"], order)
    ref = oracle.hip_tree(sk, oracle.leaf_depth(n, 16), 16)
    say(f"  nodes {len(ref['level'])}, deepest level {int(ref['level'].max())}")
    assert len(t["level"]) == len(ref["level"])
    for key in ("level", "prefix", "start", "count", "first_child", "nchild"):
        assert np.array_equal(t[key].astype(np.int64), ref[key].astype(np.int64)), key


@pytest.mark.parametrize("bucket", [1, 16, 64])
def test_topology_bit_exact_coincident_runs(oracle, bucket):
    """Runs of exactly coincident particles (equal keys down to level 21)
    beside a uniform background: 5,000 of them, longer than a topology tile
    of 4,096 keys and so across a tile boundary; 65, one more than the tiles'
    halo; 64, the halo exactly; 17, one more than the default bucket. Each run
    ends as one level-21 leaf holding all of it."""
    n = 12000
    say(f"case topology coincident runs n={n} bucket={bucket}")
    rng = np.random.default_rng(77)
    pos = rng.random((n, 3))
    pos[100:5100] = (0.25, 0.5, 0.75)
    pos[6000:6065] = (0.8, 0.1, 0.3)
    pos[7000:7064] = (0.6, 0.6, 0.6)
    pos[8000:8017] = (0.1, 0.9, 0.2)
    m = np.full(n, 1.0 / n, np.float32)
    with _solver(n, leaf_capacity=16, bucket_capacity=bucket) as s:
        s.upload(pos, None, m)
        s.build_tree()
        t = s.dump_tree()
    sk, order = oracle.stable_sort(oracle.keys(pos, 0))
    assert np.array_equal(t["sorted_keys"], sk)
    assert np.array_equal(t["sorted_ids"], order)
    ref = oracle.hip_tree(sk, oracle.leaf_depth(n, 16), bucket)
    assert len(t["level"]) == len(ref["level"])
    for key in ("level", "prefix", "start", "count", "first_child", "nchild"):
        assert np.array_equal(t[key].astype(np.int64), ref[key].astype(np.int64)), key
    # the 5,000-particle run is one leaf at level 21
    leaf = (ref["level"] == 21) & (ref["count"] == 5000) & (ref["nchild"] == 0)
    assert leaf.sum() == 1


@pytest.mark.parametrize("key_mode", [0, 1])
def test_reference_tree_is_truncation(oracle, key_mode):
    n = 30000
    say(f"case truncation key_mode={key_mode}")
    pos, m = _clustered(n, 31)
    with _solver(n, key_mode=key_mode, leaf_capacity=16, bucket_capacity=8) as s:
        s.upload(pos, None, m)
        s.build_tree()
        t = s.dump_tree()
    sk, order = oracle.stable_sort(oracle.keys(pos, key_mode))
    d = oracle.leaf_depth(n, 16)
    ref = oracle.reference_tree(sk, order, pos, m, d)
    top = int(ref["level"].min())
    rset = {(int(a), int(b), int(c), int(e)) for a, b, c, e in
            zip(ref["level"], ref["prefix"], ref["start"], ref["count"])}
    hset = {(int(a), int(b), int(c), int(e)) for a, b, c, e in
            zip(t["level"], t["prefix"], t["start"], t["count"]) if top <= a <= d}
    assert rset == hset
    # FP64 subtree masses equal the reference node masses
    lvl = t["level"]
    sel = (lvl >= top) & (lvl <= d)
    mine = {(int(a), int(b)): float(mm) for a, b, mm in
            zip(lvl[sel], t["prefix"][sel], t["moments"][sel, 3])}
    for a, b, mm in zip(ref["level"], ref["prefix"], ref["msum"]):
        if key_mode == 0:
            assert mine[(int(a), int(b))] == pytest.approx(mm, rel=2e-5)


def test_moments_match_fp64(oracle):
    n = 20000
    say(f"case moments n={n}")
    import blitzar_amd as B
    pos, vel, m = B.plummer(n, 5)
    with _solver(n, leaf_capacity=16, bucket_capacity=16) as s:
        s.upload(pos, vel, m)
        s.build_tree()
        t = s.dump_tree()
    g = t["geom"]
    lo, unit = g[6:9], g[9:12]
    mo = t["moments"]
    lat = t["lattice"].astype(np.float64)
    ids = t["sorted_ids"]
    rng = np.random.default_rng(0)
    nodes = rng.choice(len(t["level"]), 300, replace=False)
    nodes = np.unique(np.concatenate([nodes, [0]]))
    for i in nodes:
        members = ids[t["start"][i]:t["start"][i] + t["count"][i]]
        ref = oracle.moments(pos[members], m[members].astype(np.float64))
        com = lo + lat[i] * unit + mo[i, 0:3]
        assert mo[i, 3] == pytest.approx(ref[0], rel=1e-5)
        size = mo[i, 10]
        assert np.abs(com - ref[1:4]).max() <= 1e-5 * size + 1e-12
        q = np.array([mo[i, 4], mo[i, 5], mo[i, 6], mo[i, 7], mo[i, 8], -mo[i, 4] - mo[i, 7]])
        qref = ref[[4, 5, 6, 7, 8, 9]]
        scale = ref[0] * size * size
        assert np.abs(q - qref).max() <= 2e-4 * scale + 1e-12
        # trace of the second moment (the softened expansion's eps^2 term)
        y = pos[members] - ref[1:4]
        trs = float(np.sum(m[members] * np.sum(y * y, axis=1)))
        assert abs(mo[i, 11] - trs) <= 2e-4 * scale + 1e-12
        # bmax encloses every particle
        r = np.linalg.norm(pos[members] - ref[1:4], axis=1).max()
        assert mo[i, 9] >= r * (1 - 1e-5) - 1e-6 * size


def test_sort_pairs_edge_cases():
    import blitzar_amd as B
    rng = np.random.default_rng(11)
    cases = [
        ("single", np.array([5], np.uint64)),
        ("pair", np.array([9, 3], np.uint64)),
        ("2047", rng.integers(0, 2 ** 63, 2047, dtype=np.uint64)),
        ("2049", rng.integers(0, 2 ** 63, 2049, dtype=np.uint64)),
        ("tile-1", rng.integers(0, 2 ** 63, 4095, dtype=np.uint64)),
        ("tile", rng.integers(0, 2 ** 63, 4096, dtype=np.uint64)),
        ("tile+1", rng.integers(0, 2 ** 63, 4097, dtype=np.uint64)),
        ("dups", rng.integers(0, 7, 100000).astype(np.uint64)),
        ("equal", np.full(30000, 12345, np.uint64)),
        ("max", np.full(5000, 2 ** 64 - 1, np.uint64)),
        ("big", rng.integers(0, 2 ** 64 - 1, 300000, dtype=np.uint64)),
        # 733 tiles: the digit-count scan spans 92 blocks, so its look-back
        # crosses more than one 64-block window
        ("huge", rng.integers(0, 2 ** 64 - 1, 3000000, dtype=np.uint64)),
    ]
    for name, k in cases:
        say(f"case sort {name} n={k.size}")
        v = np.arange(k.size, dtype=np.uint32)[::-1].copy()
        ko, vo = B.sort_pairs(k, v, 64)
        o = np.argsort(k, kind="stable")
        assert np.array_equal(ko, k[o]), name
        assert np.array_equal(vo, v[o]), name
    k = rng.integers(0, 2 ** 40, 50000, dtype=np.uint64)
    ko, vo = B.sort_pairs(k, np.arange(k.size, dtype=np.uint32), 30)
    low = k & np.uint64(2 ** 30 - 1)
    o = np.argsort(low, kind="stable")
    assert np.array_equal(vo, o.astype(np.uint32))


def test_topology_with_a_dense_clump(oracle):
    """A level run thread-per-node (more than 4,096 nodes) that also holds
    nodes of more than 2,048 keys: those are listed and searched by whole
    wavefronts in a second kernel. 160,000 uniform particles give the level
    past the leaf depth 32,768 nodes; a clump of 240,000 particles in a
    1e-4-wide cube puts nodes of up to 240,000 keys on the same levels."""
    rng = np.random.default_rng(77)
    n_u, n_c = 160000, 240000
    pos = np.vstack([rng.random((n_u, 3)), 0.3 + 1e-4 * rng.random((n_c, 3))])
    n = len(pos)
    m = np.full(n, 1.0 / n, np.float32)
    say(f"case topology dense clump n={n}")
    with _solver(n, leaf_capacity=16, bucket_capacity=16) as s:
        s.upload(pos, None, m)
        s.build_tree()
        t = s.dump_tree()
    sk, _ = oracle.stable_sort(oracle.keys(pos, 0))
    d = oracle.leaf_depth(n, 16)
    ref = oracle.hip_tree(sk, d, 16)
    lv, cnt = ref["level"], ref["count"]
    per_level = np.bincount(lv)
    big_levels = [L for L in range(len(per_level))
                  if per_level[L] > 4096 and np.any(cnt[lv == L] > 2048)]
    assert big_levels  # the case exercises the large-node list
    assert len(t["level"]) == len(ref["level"])
    for key in ("level", "prefix", "start", "count", "first_child", "nchild"):
        assert np.array_equal(t[key].astype(np.int64), ref[key].astype(np.int64)), key


def test_node_capacity_overflow_recovers(oracle):
    """Near-duplicate pairs split in single-child chains down to level 20:
    16,153 nodes for 2,000 particles, twice the initial capacity (2n + 4096).
    The build that overflows drops the level that would not fit (flagged,
    every write inside the node arrays); the call sees the flag right after
    the build, doubles the capacity and builds again, so it returns the full
    tree, equal to the oracle's. A force evaluation that meets the overflow
    gives the forces of a handle that already had the capacity, bit for bit."""
    import blitzar_amd as B
    n = 2000
    rng = np.random.default_rng(9)
    base = rng.random((n // 2, 3))
    pos = np.repeat(base, 2, axis=0)
    pos[1::2] += 1.0e-6
    m = np.full(n, 1.0 / n, np.float32)
    say("case node capacity overflow n=2000")
    with _solver(n, leaf_capacity=16, bucket_capacity=1) as s:
        s.upload(pos, None, m)
        s.build_tree()
        t = s.dump_tree()
    sk, _ = oracle.stable_sort(oracle.keys(pos, 0))
    ref = oracle.hip_tree(sk, oracle.leaf_depth(n, 16), 1)
    assert len(t["level"]) == len(ref["level"]) == 16153
    for key in ("level", "prefix", "start", "count", "first_child", "nchild"):
        assert np.array_equal(t[key].astype(np.int64), ref[key].astype(np.int64)), key
    say("case node capacity overflow inside a force evaluation and a step")
    with _solver(n, leaf_capacity=16, bucket_capacity=1) as a, \
            _solver(n, leaf_capacity=16, bucket_capacity=1) as b:
        a.upload(pos, None, m)
        b.upload(pos, None, m)
        b.build_tree()  # grows b's capacity first
        fa = a.compute_accelerations()
        fb = b.compute_accelerations()
        assert np.array_equal(fa, fb)
        a.upload(pos, None, m)
        c = _solver(n, leaf_capacity=16, bucket_capacity=1)
        c.upload(pos, None, m)
        c.step(1e-4, 3)   # overflows at its first build, recovers, completes
        a.step(1e-4, 3)
        pa, va = a.download()
        pc, vc = c.download()
        c.close()
        assert np.array_equal(pa, pc) and np.array_equal(va, vc)


def test_failed_node_growth_leaves_the_handle_unusable(monkeypatch):
    """A node regrow whose allocation fails (forced by the test switch, which
    asks for an impossible size) releases the node arrays and leaves the
    handle unusable: the call and every later one return an error, none
    touches the released memory."""
    import blitzar_amd as B
    n = 2000
    rng = np.random.default_rng(9)
    pos = np.repeat(rng.random((n // 2, 3)), 2, axis=0)
    pos[1::2] += 1.0e-6
    m = np.full(n, 1.0 / n, np.float32)
    say("case failed node growth n=2000")
    monkeypatch.setenv("BH_HIP_TEST_FAIL_NODE_GROWTH", "1")
    with _solver(n, leaf_capacity=16, bucket_capacity=1) as s:
        monkeypatch.delenv("BH_HIP_TEST_FAIL_NODE_GROWTH")
        s.upload(pos, None, m)
        with pytest.raises(B.BhHipError, match="unusable"):
            s.compute_accelerations()
        for call in (lambda: s.step(1e-3, 1), s.download, s.build_tree, s.dump_tree,
                     lambda: s.upload(pos, None, m)):
            with pytest.raises(B.BhHipError, match="unusable"):
                call()
