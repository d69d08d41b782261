"""GPU parity of the walk (forces), the FP64 direct sum, kick/drift and the
FP64 diagnostics, through the C ABI.

Tolerances (FP32 walk):
  * near-exact opening (theta = 0.05): median relative error vs FP64 direct
    sum <= 1e-5 (FP32 rounding only)
  * theta = 0.5 quadrupole: median <= 1e-3 (the metric's bar), P99 <= 1e-2
  * theta = 0.5 vs the octree_cpu restatement at the same theta: median
    relative difference <= 1e-3 (octree_cpu is monopole-only, so the
    difference is dominated by its own error)
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, read_xyz

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


def test_two_body_closed_form():
    import blitzar_amd as B
    say("case two_body")
    pos, mass = read_xyz(os.path.join(GOLDEN, "two_body_rest.xyz"))
    with B.OctreeHip(2, softening=0.01) as s:
        s.upload(pos, None, mass.astype(np.float32))
        a = s.compute_accelerations()
    d = 2.0
    expect = d / (d * d + 1e-4) ** 1.5
    assert a[0, 0] == pytest.approx(expect, rel=2e-6)
    assert a[1, 0] == pytest.approx(-expect, rel=2e-6)
    assert np.all(a[:, 1:] == 0.0)


def test_self_interaction_is_exactly_zero():
    # A target must meet itself as a source at exactly zero separation: its
    # position is rebuilt from lattice offsets in both roles, and a one-ulp
    # difference would add m * ulp / eps^3 (large for small eps, large |x|).
    import blitzar_amd as B
    rng = np.random.default_rng(17)
    eps = 1e-3
    for case in range(20):
        pos = rng.uniform(-300.0, 300.0, (2, 3))
        m = rng.uniform(0.5, 2.0, 2).astype(np.float32)
        with B.OctreeHip(2, softening=eps) as s:
            s.upload(pos, None, m)
            a = s.compute_accelerations()
        d = pos[1] - pos[0]
        r2 = d @ d + float(np.float32(eps)) ** 2
        want0 = float(m[1]) * d / r2 ** 1.5
        want1 = -float(m[0]) * d / r2 ** 1.5
        assert np.linalg.norm(a[0] - want0) <= 1e-5 * np.linalg.norm(want0), case
        assert np.linalg.norm(a[1] - want1) <= 1e-5 * np.linalg.norm(want1), case


def test_single_particle_has_no_force():
    import blitzar_amd as B
    with B.OctreeHip(1) as s:
        s.upload(np.array([[0.3, -0.2, 5.0]]), None, np.array([1.0], np.float32))
        a = s.compute_accelerations()
    assert np.all(a == 0.0)


def test_direct_fp64_matches_oracle(oracle):
    import blitzar_amd as B
    n = 20000
    say(f"case direct n={n}")
    pos, vel, m = B.plummer(n, 3)
    t = np.arange(0, n, 97, dtype=np.uint32)
    with B.OctreeHip(n, softening=0.01) as s:
        s.upload(pos, vel, m)
        a = s.direct_forces_fp64(t)
    # the solver's softening is an FP32 parameter: pass the same value
    ref = oracle.direct_fp64(pos, m.astype(np.float64), t.astype(np.int32),
                             eps=float(np.float32(0.01)), min_d2=float(np.float32(1e-12)))
    assert rel(a, ref).max() < 1e-12


@pytest.mark.parametrize("order", [1, 2])
def test_small_theta_is_direct_sum(oracle, order):
    import blitzar_amd as B
    n = 8000
    say(f"case theta0.05 order={order}")
    pos, vel, m = B.plummer(n, 9)
    with B.OctreeHip(n, theta=0.05, softening=0.01, multipole_order=order,
                     leaf_capacity=16, bucket_capacity=16) as s:
        s.upload(pos, vel, m)
        a = s.compute_accelerations()
    t = np.arange(0, n, 7)
    ref = oracle.direct_fp64(pos, m.astype(np.float64), t, eps=0.01)
    r = rel(a[t], ref)
    rms = np.sqrt((ref ** 2).sum(1).mean())
    say(f"median {np.median(r):.2e} p99 {np.percentile(r, 99):.2e} "
        f"max/rms {np.linalg.norm(a[t] - ref, axis=1).max() / rms:.2e}")
    # At theta = 0.05 the truncation error is far below FP32 accumulation
    # error (~1e-7 * sum |a_j| / |a|, larger where the net force cancels).
    assert np.median(r) < 1e-5
    assert np.percentile(r, 99) < 5e-5
    assert np.linalg.norm(a[t] - ref, axis=1).max() / rms < 5e-5


@pytest.mark.parametrize("criterion", ["com", "bounds"])
@pytest.mark.parametrize("n", [5000, 60000])
def test_theta_half_accuracy(oracle, criterion, n):
    import blitzar_amd as B
    say(f"case theta0.5 n={n} criterion={criterion}")
    pos, vel, m = B.plummer(n, 7)
    with B.OctreeHip(n, theta=0.5, softening=0.005, criterion=criterion) as s:
        s.upload(pos, vel, m)
        a = s.compute_accelerations()
        tm = s.timings()
    rng = np.random.default_rng(1)
    t = rng.choice(n, 1024, replace=False)
    ref = oracle.direct_fp64(pos, m.astype(np.float64), t, eps=0.005)
    r = rel(a[t], ref)
    say(f"  median={np.median(r):.3e} p99={np.percentile(r, 99):.3e} "
        f"pp={tm['pp_interactions']} pc={tm['pc_interactions']}")
    assert np.median(r) < 1e-3
    assert np.percentile(r, 99) < 1e-2
    # quadrupole beats the monopole-only octree_cpu at the same theta
    x = oracle.hot(pos, m)
    acpu = oracle.octree_forces_subset(x, t, oracle.make_params(theta=0.5, softening=0.005,
                                                                criterion=1 if criterion == "bounds" else 0))
    rc = rel(acpu, ref)
    assert np.median(r) <= np.median(rc)
    assert np.median(rel(a[t], acpu)) < max(1e-3, 1.5 * np.median(rc))


def test_clamp_and_min_distance(oracle):
    import blitzar_amd as B
    pos = np.array([[0.0, 0, 0], [1e-3, 0, 0], [5.0, 0, 0]])
    m = np.ones(3, np.float32)
    with B.OctreeHip(3, softening=1e-4, max_acceleration=64.0) as s:
        s.upload(pos, None, m)
        a = s.compute_accelerations()
    assert np.linalg.norm(a[0]) == pytest.approx(64.0, rel=1e-5)
    # coincident pair with zero softening: skipped by min_distance2, no NaN
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0]])
    with B.OctreeHip(3, softening=0.0, min_softening=0.0, min_distance2=1e-12) as s:
        s.upload(pos, None, m)
        a = s.compute_accelerations()
    assert np.all(np.isfinite(a))
    assert a[0, 0] == pytest.approx(1.0, rel=1e-5)


def test_kdk_energy_and_reversibility():
    import blitzar_amd as B
    n = 4096
    say(f"case kdk n={n}")
    pos, vel, m = B.plummer(n, 42)
    with B.OctreeHip(n, theta=0.4, softening=0.01, integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        e0 = s.diagnostics()
        s.step(1e-3, 50)
        e1 = s.diagnostics()
        p1, v1 = s.download()
        # time reversal: flip velocities, step back
        s.upload(p1, -v1, m)
        s.step(1e-3, 50)
        p2, v2 = s.download()
    E0 = e0["ke"] + e0["pe"]
    E1 = e1["ke"] + e1["pe"]
    assert abs(E1 - E0) / abs(E0) < 1e-4
    assert np.abs(p2 - pos).max() < 1e-6
    assert np.abs(e1["p"]).max() < 1e-6


def test_euler_matches_reference_update():
    import blitzar_amd as B
    pos = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    m = np.ones(2, np.float32)
    dt = 0.01
    with B.OctreeHip(2, softening=0.01, integrator="euler") as s:
        s.upload(pos, None, m)
        a = s.compute_accelerations()
        assert s.update(dt)
        p, v = s.download()
    assert v[0, 0] == pytest.approx(a[0, 0] * dt, rel=1e-6)
    assert p[0, 0] == pytest.approx(-1.0 + a[0, 0] * dt * dt, rel=1e-6)


def test_diagnostics_match_oracle(oracle):
    import blitzar_amd as B
    n = 3000
    pos, vel, m = B.plummer(n, 2)
    with B.OctreeHip(n, softening=0.01) as s:
        s.upload(pos, vel, m)
        d = s.diagnostics()
        ds = s.diagnostics(sample_limit=500)
    ref = oracle.energy(pos, vel, m.astype(np.float64), 0.01, float(np.float32(1e-12)))
    assert d["ke"] == pytest.approx(ref["ke"], rel=1e-12)
    assert d["pe"] == pytest.approx(ref["pe"], rel=1e-10)
    assert np.allclose(d["p"], ref["p"], atol=1e-12)
    assert np.allclose(d["l"], ref["l"], atol=1e-12)
    assert ds["pe"] == pytest.approx(ref["pe"], rel=0.1)


def test_handles_coexist():
    import blitzar_amd as B
    a_pos, a_vel, a_m = B.plummer(3000, 1)
    b_pos, b_vel, b_m = B.plummer(5000, 2)
    with B.OctreeHip(3000) as s1, B.OctreeHip(5000) as s2:
        s1.upload(a_pos, a_vel, a_m)
        s2.upload(b_pos, b_vel, b_m)
        x1 = s1.compute_accelerations()
        x2 = s2.compute_accelerations()
        y1 = s1.compute_accelerations()
    assert np.array_equal(x1, y1)
    assert x2.shape == (5000, 3)


@pytest.mark.parametrize("theta,alpha,factor", [(0.5, 0.6, 1.5), (0.5, 0.9, 2.5),
                                                (0.8, 0.75, 1.2)])
def test_group_far_field_accuracy(theta, alpha, factor):
    """Accepted nodes far from the whole group enter the group's local
    expansion; the forces keep the walk's accuracy against the FP64 sum."""
    import blitzar_amd as B
    n = 60000
    say(f"case far field n={n} theta={theta} alpha={alpha}")
    pos, vel, m = B.plummer(n, 7)
    t = np.arange(0, n, 29, dtype=np.uint32)
    with B.OctreeHip(n, theta=theta, softening=0.005) as s:
        s.set_timing(True)
        s.upload(pos, vel, m)
        ref = s.direct_forces_fp64(t)
        a0 = s.compute_accelerations()
        s.set_params(far_alpha=alpha)
        a1 = s.compute_accelerations()
        tm = s.timings()
    e0, e1 = rel(a0[t], ref), rel(a1[t], ref)
    say(f"  plain median {np.median(e0):.2e} p99 {np.percentile(e0, 99):.2e} | far field "
        f"median {np.median(e1):.2e} p99 {np.percentile(e1, 99):.2e} "
        f"far nodes {tm['far_interactions']} cells {tm['pc_interactions']}")
    assert tm["far_interactions"] > 0
    # far_alpha trades accuracy for speed: bounded relative to the plain walk
    assert np.median(e1) < factor * np.median(e0) + 1e-5
    assert np.percentile(e1, 99) < 2 * factor * np.percentile(e0, 99) + 1e-4


def _box_cluster(rng, n, centre, half):
    return np.asarray(centre) + rng.uniform(-1.0, 1.0, (n, 3)) * np.asarray(half)


@pytest.mark.parametrize("eps", [0.01, 0.8])
def test_group_far_field_matches_local_expansion_oracle(eps):
    """Two source clusters, each the only content of one root octant, and a
    group of 40 light targets in a third: at theta 0.7 both source octants
    are accepted whole, and at far_alpha 0.9 both enter the group's local
    expansion. The walk must then equal the FP64 local expansion of
    oracle/expansion.py (sum over the two nodes, about the target box
    centre) to FP32 rounding, while the expansion itself differs from the
    direct sum by its third-order truncation error (checked to be 30x the
    tolerance, so a dropped or mis-signed term cannot pass)."""
    import blitzar_amd as B
    from oracle import expansion as X
    say(f"case far field vs local expansion eps={eps}")
    rng = np.random.default_rng(41)
    tgt = _box_cluster(rng, 40, (-1.0, -1.0, -1.0), (0.4, 0.4, 0.4))
    src_a = _box_cluster(rng, 100, (1.2, 1.2, 1.2), (0.3, 0.1, 0.15))
    src_b = _box_cluster(rng, 80, (1.2, -1.0, 1.2), (0.1, 0.25, 0.1))
    m_a = rng.uniform(0.5, 1.5, 100) * 0.01
    m_b = rng.uniform(0.5, 1.5, 80) * 0.01
    pos = np.vstack([tgt, src_a, src_b])
    m = np.concatenate([np.full(40, 1e-15), m_a, m_b]).astype(np.float32)
    with B.OctreeHip(len(pos), theta=0.7, softening=eps, multipole_order=2) as s:
        s.set_timing(True)
        s.upload(pos, None, m)
        s.set_params(far_alpha=0.9)
        acc = s.compute_accelerations()[:40].astype(np.float64)
        tm = s.timings()
    assert tm["far_interactions"] >= 2
    # the FP64 expansion of both nodes about the centre of the target box
    gc = 0.5 * (tgt.min(0) + tgt.max(0))
    want = np.zeros((40, 3))
    direct = np.zeros((40, 3))
    for xs, ms in ((src_a, m_a), (src_b, m_b)):
        ms = ms.astype(np.float32).astype(np.float64)
        M = ms.sum()
        com = (ms[:, None] * xs).sum(0) / M
        y = xs - com
        r2 = (y * y).sum(1)
        S = np.einsum("n,ni,nj->ij", ms, y, y)
        Q = 3.0 * S - r2 @ ms * np.eye(3)
        q6 = (Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2])
        A, J, H = X.local_expansion(M, com - gc, q6, float(r2 @ ms), eps)
        want += X.evaluate(A, J, H, tgt - gc)
        direct += X.direct(tgt, xs, ms, eps)
    scale = np.linalg.norm(want, axis=1)
    d_gpu = np.linalg.norm(acc - want, axis=1) / scale
    d_trunc = np.linalg.norm(direct - want, axis=1) / scale
    say(f"  GPU vs expansion max {d_gpu.max():.2e}; expansion vs direct max {d_trunc.max():.2e}")
    tol = 2e-5
    assert d_trunc.max() > 30 * tol
    assert d_gpu.max() < tol


def test_relative_criterion_accuracy_and_priming():
    """The relative criterion (a node's estimated error below alpha times the
    group's smallest previous |a|): the first evaluation after an upload is
    primed by a COM walk, a tighter alpha gives a smaller error, and the
    error stays bounded after KDK steps that permute the store (the |a| of
    the last evaluation follows the permutation)."""
    import blitzar_amd as B
    n = 60000
    say(f"case relative criterion n={n}")
    pos, vel, m = B.plummer(n, 7)
    t = np.arange(0, n, 29, dtype=np.uint32)
    med = {}
    with B.OctreeHip(n, theta=1.1, softening=0.005, criterion="relative",
                     relative_alpha=0.004, integrator="leapfrog") as s:
        s.upload(pos, vel, m)
        ref = s.direct_forces_fp64(t)
        first = s.compute_accelerations()  # COM priming walk + relative walk
        assert np.all(np.isfinite(first))
        for alpha in (0.016, 0.004, 0.001):
            s.set_params(relative_alpha=alpha)
            e = rel(s.compute_accelerations()[t], ref)
            med[alpha] = float(np.median(e))
            say(f"  alpha {alpha}: median {med[alpha]:.2e} p99 {np.percentile(e, 99):.2e}")
        assert med[0.001] < med[0.004] < med[0.016]
        assert med[0.004] < 1e-3
        s.step(1e-3, 3)
        p3, _ = s.download()
        s.set_params(relative_alpha=0.004)
        a3 = s.compute_accelerations()
        r3 = s.direct_forces_fp64(t)
    e3 = rel(a3[t], r3)
    say(f"  after 3 steps: median {np.median(e3):.2e}")
    assert np.median(e3) < 1.5 * med[0.004] + 1e-5
    with B.OctreeHip(n, theta=1.1, softening=0.005, criterion="relative",
                     relative_alpha=0.004) as s2:
        s2.upload(p3, None, m)
        b3 = s2.compute_accelerations()  # fresh state: primed again
    assert np.median(rel(b3[t], r3)) < 1.5 * med[0.004] + 1e-5


def test_relative_point_mass_cells():
    """relative_mono: accepted nodes whose quadrupole term is within
    relative_mono * alpha |a| are evaluated as softened point masses. A small
    fraction keeps the error of the second-order walk; the point-mass count
    grows with the fraction; 0 leaves only nodes with bmax = 0 (single
    particles, exact as point masses)."""
    import blitzar_amd as B
    n = 60000
    say(f"case relative point masses n={n}")
    pos, vel, m = B.plummer(n, 7)
    t = np.arange(0, n, 29, dtype=np.uint32)
    out = {}
    with B.OctreeHip(n, theta=2.0, softening=0.005, criterion="relative",
                     relative_alpha=0.004) as s:
        s.set_timing(True)
        s.upload(pos, vel, m)
        ref = s.direct_forces_fp64(t)
        s.compute_accelerations()
        for k in (0.0, 0.1, 1.0):
            s.set_params(relative_mono=k)
            a = s.compute_accelerations()
            tm = s.timings()
            out[k] = (float(np.median(rel(a[t], ref))), tm["mono_interactions"],
                      tm["pc_interactions"])
            say(f"  mono {k}: median {out[k][0]:.2e} point masses {out[k][1]} cells {out[k][2]}")
    assert out[0.1][1] > out[0.0][1] and out[1.0][1] > out[0.1][1]
    assert out[1.0][2] < out[0.1][2] < out[0.0][2]
    assert out[0.1][0] < 1.3 * out[0.0][0] + 1e-5
    assert out[1.0][0] < 1e-2


@pytest.mark.parametrize("criterion", ["com", "relative"])
def test_ticket_queues_do_not_change_results(criterion):
    """Walks of at least 16 groups per resident wave hand groups out through
    eight per-XCD ticket queues, smaller walks through one counter. Every
    group is walked exactly once either way, and a group's forces do not
    depend on the wave that walks it: a residency of 2 waves per CU (queues)
    and the default (one counter) give bit-identical accelerations, also
    after steps that permute the store."""
    import blitzar_amd as B
    n = 600000
    say(f"case ticket queues criterion={criterion} n={n}")
    pos, vel, m = B.plummer(n, 21)
    out = []
    for wpc in (2, 24):
        with B.OctreeHip(n, theta=0.6, softening=0.005, criterion=criterion,
                         relative_alpha=0.004, waves_per_cu=wpc) as s:
            s.set_timing(True)
            s.upload(pos, vel, m)
            a0 = s.compute_accelerations()
            s.step(1e-3, 2)
            a1 = s.compute_accelerations()
            groups = s.timings()["groups"]
            out.append((a0, a1, s.download()[0]))
        say(f"  waves/CU {wpc}: {groups} groups")
    assert groups >= 16 * 2 * 256 > 0  # the residency-2 walk used the queues
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
