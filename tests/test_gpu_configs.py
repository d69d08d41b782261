"""GPU parity at the full sizes of the five benchmark workloads, through the C
ABI, against the oracle.

At these sizes the oracle checks sampled targets: the FP64 direct sum
(oracle/direct.cpp, the counterpart of TpmCpu.cpp:78-116) on up to 512
targets on the CPU, which also pins the GPU FP64 direct sum used for the full
4,096-target samples; the octree_cpu restatement (oracle/octree_cpu.cpp) on
4,096 targets; the linear-octree restatement (oracle/linear_octree.cpp) for
the whole 1M topology; the FP64 periodic erfc sum and the FP64 numpy mesh
(oracle/pm.py) for the periodic box. Size-independent properties (energy,
momentum, schedule counts) cover what a sample cannot.

Tolerances (the walk is FP32):
  * force error vs the FP64 direct sum: median <= 1e-3 (the metric's bar),
    P99 <= 1e-2 (open walks) or 2e-2 (periodic short range);
  * GPU FP64 direct sum vs the CPU FP64 direct sum: 1e-10 relative;
  * vs octree_cpu at the same theta: median <= 2e-3 (octree_cpu is monopole
    only, its own error dominates the difference);
  * keys, sorted order and topology: bit-exact;
  * energy drift of 100 leapfrog steps: <= 1e-5 relative;
  * particle mesh vs the FP64 mesh: median <= 3e-4, L2 <= 3e-4;
  * periodic short-range walk of the 256^3 displaced lattice vs the FP64
    periodic erfc sum: median <= 1e-4 at theta 0.3, <= 2e-2 at the
    workload's theta 0.6 (a near-cancelling sum, see the test).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


def sample(n, k, seed=2024):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, k, replace=False)).astype(np.uint32)


def check_against_direct(oracle, s, pos, m, t, eps, cpu_targets=256, median=1e-3, p99=1e-2):
    """GPU walk forces at targets t vs the GPU FP64 direct sum, itself checked
    against the CPU oracle on the first cpu_targets of t."""
    a = s.compute_accelerations()
    ref = s.direct_forces_fp64(t)
    tc = t[:cpu_targets]
    ref_cpu = oracle.direct_fp64(pos, m.astype(np.float64), tc.astype(np.int32),
                                 eps=float(np.float32(eps)))
    assert rel(ref[:cpu_targets], ref_cpu).max() < 1e-10
    r = rel(a[t], ref)
    say(f"  vs FP64 direct: median={np.median(r):.3e} p99={np.percentile(r, 99):.3e}")
    assert np.median(r) <= median and np.percentile(r, 99) <= p99
    return a


def test_plummer_65536_seed42_kdk_100_steps(oracle):
    import blitzar_amd as B
    n, eps = 65536, 0.01
    say(f"case plummer n={n} seed=42 theta=0.5 eps={eps} kdk 100 steps")
    pos, vel, m = B.plummer(n, seed=42)
    t = sample(n, 4096)
    with B.OctreeHip(n, theta=0.5, softening=eps, integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        a = check_against_direct(oracle, s, pos, m, t, eps)
        acpu = oracle.octree_forces_subset(oracle.hot(pos, m), t.astype(np.int32),
                                           oracle.make_params(theta=0.5, softening=eps))
        r = rel(a[t], acpu)
        say(f"  vs octree_cpu: median={np.median(r):.3e}")
        assert np.median(r) <= 2e-3
        d0 = s.diagnostics(sample_limit=n)
        s.step(1e-3, 100)
        d1 = s.diagnostics(sample_limit=n)
        p1, v1 = s.download()
    # the GPU diagnostics equal the (pinned) FP64 oracle on the final state
    ref = oracle.energy(p1, v1, m.astype(np.float64), float(np.float32(eps)))
    assert d1["ke"] == pytest.approx(ref["ke"], rel=1e-10)
    assert d1["pe"] == pytest.approx(ref["pe"], rel=1e-10)
    e0, e1 = d0["ke"] + d0["pe"], d1["ke"] + d1["pe"]
    drift = abs(e1 - e0) / abs(e0)
    say(f"  energy drift over 100 steps: {drift:.3e}")
    assert drift <= 1e-5
    # momentum: the walk is not exactly antisymmetric, but close
    assert np.linalg.norm(np.asarray(d1["p"]) - np.asarray(d0["p"])) <= 1e-4


def test_plummer_1048576_seed7_theta05_steps_0_and_20(oracle):
    import blitzar_amd as B
    n, eps = 1 << 20, 0.005
    say(f"case plummer n={n} seed=7 theta=0.5 eps={eps} dt=5e-4 20 steps")
    pos, vel, m = B.plummer(n, seed=7)
    t = sample(n, 4096)
    params = oracle.make_params(theta=0.5, softening=eps)
    with B.OctreeHip(n, theta=0.5, softening=eps, integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        say("  step 0")
        a0 = check_against_direct(oracle, s, pos, m, t, eps)
        acpu = oracle.octree_forces_subset(oracle.hot(pos, m), t.astype(np.int32), params)
        r = rel(a0[t], acpu)
        say(f"  vs octree_cpu: median={np.median(r):.3e}")
        assert np.median(r) <= 2e-3
        # the whole 1M topology, bit-exact against the linear-octree restatement
        tree = s.dump_tree()
        sk, order = oracle.stable_sort(oracle.keys(pos, 0))
        assert np.array_equal(tree["sorted_keys"], sk)
        assert np.array_equal(tree["sorted_ids"], order)
        ref = oracle.hip_tree(sk, oracle.leaf_depth(n, 256), 16)
        assert len(tree["level"]) == len(ref["level"])
        for key in ("level", "prefix", "start", "count", "first_child", "nchild"):
            assert np.array_equal(tree[key].astype(np.int64), ref[key].astype(np.int64)), key
        s.step(5e-4, 20)
        p20, _ = s.download()
        say("  step 20")
        a20 = check_against_direct(oracle, s, p20, m, t, eps)
        acpu = oracle.octree_forces_subset(oracle.hot(p20, m), t.astype(np.int32), params)
        assert np.median(rel(a20[t], acpu)) <= 2e-3


def test_hernquist_merger_8388608_block_steps_64_base_steps(oracle):
    import blitzar_amd as B
    n, eps, L = 8388608, 0.01, 8
    base_dt, steps = 2.0 ** -6, 64
    say(f"case hernquist merger n={n} block steps levels={L} theta=0.5 {steps} base steps")
    pos, vel, m = B.hernquist_merger(n // 2, seed=11, separation=20.0, pericentre=2.0)
    t = sample(n, 4096)
    with B.OctreeHip(n, theta=0.5, softening=eps, integrator="leapfrog",
                     adaptive_max_level=L, adaptive_eta=0.02, adaptive_flags=0) as s:
        s.upload(pos, vel, m)
        s.set_timing(True)
        d0 = s.diagnostics(sample_limit=4096)
        s.step(base_dt, steps)
        tm = s.timings()
        levels = s.adaptive_levels()
        d1 = s.diagnostics(sample_limit=4096)
        p1, _ = s.download()
        hist = np.bincount(levels, minlength=L + 1)
        say(f"  levels {hist.tolist()} particle steps {tm['particle_steps']} slices {tm['walks']}")
        assert levels.min() >= 0 and levels.max() <= L and hist.sum() == n
        # every particle is corrected at the last tick of each base step
        # (each cadence divides 2^L): at least n corrections per base step,
        # and about sum_l count_l 2^l of them while the levels hold
        ps = tm["particle_steps"]
        assert ps >= steps * n
        expect = steps * float(np.sum(hist * 2.0 ** np.arange(L + 1)))
        assert 0.5 * expect <= ps <= 2.0 * expect
        # momentum of the merger (its total is zero to rounding at the start)
        dp = np.linalg.norm(np.asarray(d1["p"]) - np.asarray(d0["p"]))
        say(f"  momentum change {dp:.3e}")
        assert dp <= 1e-3
        assert np.all(np.isfinite(p1))
        # force error of the theta 0.5 walk on the evolved state
        s.set_params(adaptive_max_level=0)
        say("  forces of the final state")
        check_against_direct(oracle, s, p1, m, t, eps, cpu_targets=64)


def test_zeldovich_256cubed_periodic_treepm(oracle):
    import blitzar_amd as B
    from oracle import pm
    ng = 256
    h = 1.0 / ng
    rs, rc, eps = 1.25 * h, 4.5 * 1.25 * h, 0.05 * h
    say(f"case periodic treepm zeldovich {ng}^3 mesh {ng} theta=0.6")
    pos = B.zeldovich_lattice(ng)
    n = pos.shape[0]
    m = np.full(n, 1.0 / n, np.float32)
    t = sample(n, 512, seed=5)
    f = lambda x: float(np.float32(x))
    ref = oracle.direct_tpm_fp64(pos, m.astype(np.float64), t.astype(np.int32), f(eps), f(rs),
                                 f(rc), box=1.0)
    # The short-range force of a displaced lattice is a near-cancelling sum
    # (|a| is a few nearest-neighbour forces), so the geometric criterion's
    # relative error is large at the workload's theta = 0.6: its accepted
    # nodes are as wide as r_s, where the erfc kernel bends. At theta = 0.3
    # (and 0.05) the walk equals the FP64 sum to FP32 rounding.
    for theta, med, p99 in ((0.3, 1e-4, 2e-2), (0.6, 2e-2, 1e-1)):
        kw = dict(theta=theta, softening=eps, treepm_split_scale=rs, treepm_cutoff=rc,
                  periodic_box=1.0)
        with B.OctreeHip(n, **kw) as s:
            s.upload(pos, None, m)
            a_short = s.compute_accelerations()
        r = rel(a_short[t], ref)
        say(f"  theta {theta}: short range vs FP64 periodic sum: median={np.median(r):.3e} "
            f"p99={np.percentile(r, 99):.3e}")
        assert np.median(r) <= med and np.percentile(r, 99) <= p99
    del a_short
    kw = dict(theta=0.6, softening=eps, treepm_split_scale=rs, treepm_cutoff=rc,
              periodic_box=1.0)
    with B.OctreeHip(n, pm_mesh=ng, **kw) as s:
        s.upload(pos, None, m)
        a_pm = s.pm_accelerations()
        s.step(1e-3, 10)
        p10, _ = s.download()
    ref = pm.long_range_accel(pos, m.astype(np.float64), 1.0, ng, f(rs), targets=t)
    r = rel(a_pm[t], ref)
    l2 = np.linalg.norm(a_pm[t] - ref) / np.linalg.norm(ref)
    say(f"  mesh vs FP64 mesh: median={np.median(r):.3e} l2={l2:.3e}")
    assert np.median(r) <= 3e-4 and l2 <= 3e-4
    # ten steps keep every particle in the box and finite
    assert np.all(np.isfinite(p10)) and np.all((p10 >= 0.0) & (p10 < 1.0))


def test_hernquist_merger_8388608_learned_correction_theta09(oracle):
    """The learned far-field correction at theta 0.9 (band 0.75, nodes of at
    least 64 particles) on the 8M merger: trained on a second realisation
    (seed 12), applied to the workload's state (seed 11), against the GPU
    FP64 direct sum (pinned against the CPU oracle on 64 targets) on 4,096
    targets. The corrected walk meets the metric's median bar, the network
    removes at least 15 % of the median error of the same walk with a zero
    network, and the theta 0.5 walk it is compared with is within the bar."""
    import blitzar_amd as B
    from oracle import mlp
    n, eps = 8388608, 0.01
    say(f"case learned correction merger n={n} theta=0.9 band=0.75")
    p_tr, v_tr, m_tr = B.hernquist_merger(n // 2, seed=12, separation=20.0, pericentre=2.0)
    pos, vel, m = B.hernquist_merger(n // 2, seed=11, separation=20.0, pericentre=2.0)
    t = sample(n, 4096)
    with B.OctreeHip(n, theta=0.5, softening=eps) as s:
        s.upload(pos, vel, m)
        say("  theta 0.5")
        check_against_direct(oracle, s, pos, m, t, eps, cpu_targets=64)
        ref = s.direct_forces_fp64(t)
        s.set_params(theta=0.9)
        s.lff_configure(theta=0.9, band=0.75, min_count=64)
        s.upload(p_tr, v_tr, m_tr)
        tuples = s.lff_collect(group_stride=1, max_tuples=1 << 24)
        loss = s.lff_train(epochs=6, batch=16384, lr=2e-3, seed=3)
        say(f"  tuples {tuples} loss {np.array2string(loss, precision=4)}")
        assert tuples > 1_000_000 and np.all(np.isfinite(loss)) and loss[-1] < 0.5 * loss[0]
        params = s.lff_params()
        s.upload(pos, vel, m)
        s.lff_set_params(np.zeros(mlp.NP, np.float32))
        s.lff_enable(True)
        e0 = rel(s.compute_accelerations()[t], ref)
        s.lff_set_params(params)
        e1 = rel(s.compute_accelerations()[t], ref)
    say(f"  zero network median {np.median(e0):.3e} p99 {np.percentile(e0, 99):.3e} | "
        f"learned median {np.median(e1):.3e} p99 {np.percentile(e1, 99):.3e}")
    assert np.median(e1) <= 1e-3
    assert np.median(e1) <= 0.85 * np.median(e0)
    assert np.percentile(e1, 99) <= np.percentile(e0, 99)
