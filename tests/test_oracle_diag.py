"""Pins the oracle's FP64 diagnostics (dx_energy), its kick-drift-kick step
(oc_kdk_step) and its octree_cpu tree export (oc_octree_export).

dx_energy (the counterpart of FmmMetrics.cpp:47-83 and
SrvSnapshotAndEnergy.cpp:195-310) is checked against hand-worked values:
  * two bodies: KE = sum m v^2 / 2, PE = -m1 m2 / sqrt(d^2 + eps^2),
    P = sum m v, L = sum m x x v, each with numbers worked out below;
  * three bodies on a line (pair sum i < j, each pair counted once);
  * min_distance2 drops a coincident pair from PE;
  * PE scales as m^2 and as 1/length, KE as m v^2; P and L flip with v.
oc_kdk_step (CudParticleKernels.inl:94-220 sequence) is checked by a free
particle moving on a straight line, time reversal, the velocity clamp and a
circular two-body orbit against its closed form.
oc_octree_export (OctreeBuild.cpp:147-216) is checked by the tree's own
invariants: every particle in exactly one leaf, children in the octant given
by the >= test, child cubes of half the parent's half size, mass and COM
sums, the leaf rule and pre-order numbering.
"""
import numpy as np
import pytest


def test_energy_two_body_hand_worked(oracle):
    pos = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    vel = np.array([[0.0, 0.5, 0.0], [0.0, -0.25, 0.0]])
    m = np.array([1.0, 2.0])
    d = oracle.energy(pos, vel, m, eps=0.5)
    # KE = 0.5*1*0.25 + 0.5*2*0.0625 = 0.125 + 0.0625
    assert d["ke"] == pytest.approx(0.1875, rel=1e-14)
    # PE = -1*2 / sqrt(4 + 0.25)
    assert d["pe"] == pytest.approx(-2.0 / np.sqrt(4.25), rel=1e-14)
    # P = (0, 0.5 - 0.5, 0) = 0
    assert np.allclose(d["p"], [0.0, 0.0, 0.0], atol=1e-15)
    # L_z = 1*(1*0.5 - 0) + 2*((-1)*(-0.25) - 0) = 0.5 + 0.5
    assert np.allclose(d["l"], [0.0, 0.0, 1.0], atol=1e-15)


def test_energy_three_bodies_each_pair_once(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    vel = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    m = np.array([1.0, 2.0, 3.0])
    d = oracle.energy(pos, vel, m, eps=0.0)
    # pairs (0,1): 1*2/1, (0,2): 1*3/3, (1,2): 2*3/2
    assert d["pe"] == pytest.approx(-(2.0 + 1.0 + 3.0), rel=1e-14)
    assert d["ke"] == pytest.approx(0.5 * (1.0 + 2.0 + 3.0 * 4.0), rel=1e-14)
    assert np.allclose(d["p"], [2.0, 6.0, 1.0], atol=1e-15)
    # L = sum m x x v: body 2 at (3,0,0) with v (0,2,0) gives 3*(3*2) along z;
    # body 0 at the origin and body 1 moving along its own axis give nothing
    assert np.allclose(d["l"], [0.0, 0.0, 18.0], atol=1e-14)


def test_energy_min_distance_drops_coincident_pair(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    vel = np.zeros((3, 3))
    m = np.ones(3)
    d = oracle.energy(pos, vel, m, eps=0.0, min_d2=1e-12)
    assert d["pe"] == pytest.approx(-1.0, rel=1e-14)  # two pairs at distance 2


def test_energy_scaling_laws(oracle):
    rng = np.random.default_rng(5)
    pos = rng.standard_normal((50, 3))
    vel = rng.standard_normal((50, 3))
    m = rng.random(50) + 0.5
    d0 = oracle.energy(pos, vel, m, eps=0.1)
    d1 = oracle.energy(2.0 * pos, vel, 3.0 * m, eps=0.2)   # lengths x2, eps x2, masses x3
    assert d1["pe"] == pytest.approx(d0["pe"] * 9.0 / 2.0, rel=1e-12)
    assert d1["ke"] == pytest.approx(d0["ke"] * 3.0, rel=1e-12)
    d2 = oracle.energy(pos, -vel, m, eps=0.1)
    assert np.allclose(d2["p"], -d0["p"], rtol=1e-12, atol=1e-14)
    assert np.allclose(d2["l"], -d0["l"], rtol=1e-12, atol=1e-14)
    assert d2["ke"] == pytest.approx(d0["ke"], rel=1e-14)


def test_kdk_free_particle_and_clamp(oracle):
    x = np.array([[0.25, -0.5, 1.0, 1.0]], np.float32)
    v = np.array([[0.5, 0.25, -1.0]], np.float32)
    a = np.zeros((1, 3), np.float32)
    for _ in range(4):
        oracle.kdk_step(x, v, a, 0.125, 0.0)
    assert np.allclose(x[0, :3], [0.25 + 0.25, -0.5 + 0.125, 1.0 - 0.5], atol=1e-7)
    assert np.allclose(v[0], [0.5, 0.25, -1.0]) and np.all(a == 0)
    # velocity clamp: |v| > max_speed is scaled down to max_speed after the kick
    x = np.array([[0.0, 0.0, 0.0, 1.0]], np.float32)
    v = np.array([[3.0, 4.0, 0.0]], np.float32)
    oracle.kdk_step(x, v, a, 0.1, 1.0)
    assert np.allclose(v[0], [0.6, 0.8, 0.0], atol=1e-6)
    assert np.allclose(x[0, :3], [0.06, 0.08, 0.0], atol=1e-6)


def _two_body_circular(eps=0.0):
    # two unit masses at +-1 on x: a = m / (2^2 + eps^2)^1.5 * 2 (toward the
    # other body at distance 2), circular speed v^2 / 1 = |a|
    acc = 2.0 / (4.0 + eps * eps) ** 1.5
    v = np.sqrt(acc)
    x = np.array([[1.0, 0.0, 0.0, 1.0], [-1.0, 0.0, 0.0, 1.0]], np.float32)
    vel = np.array([[0.0, v, 0.0], [0.0, -v, 0.0]], np.float32)
    return x, vel, v


def test_kdk_circular_orbit_closed_form(oracle):
    x, vel, v = _two_body_circular()
    p = oracle.make_params(theta=0.5, softening=0.0, min_softening=0.0)
    acc, _ = oracle.octree_forces(x, p)
    dt, steps = 0.01, 200
    for _ in range(steps):
        oracle.kdk_step(x, vel, acc, dt, 0.0, p)
    phi = v * dt * steps  # angular speed v / r with r = 1
    want = np.array([np.cos(phi), np.sin(phi), 0.0])
    assert np.linalg.norm(x[0, :3] - want) < 2e-4      # second-order phase error
    assert np.linalg.norm(x[1, :3] + want) < 2e-4
    assert np.hypot(x[0, 0], x[0, 1]) == pytest.approx(1.0, abs=1e-4)


def test_kdk_time_reversal(oracle):
    rng = np.random.default_rng(9)
    n = 64
    x = np.concatenate([rng.standard_normal((n, 3)), np.full((n, 1), 1.0 / n)], 1).astype(np.float32)
    vel = (0.1 * rng.standard_normal((n, 3))).astype(np.float32)
    x0 = x.copy()
    p = oracle.make_params(theta=0.05, softening=0.1)
    acc, _ = oracle.octree_forces(x, p)
    for _ in range(20):
        oracle.kdk_step(x, vel, acc, 0.01, 0.0, p)
    vel *= -1.0
    for _ in range(20):
        oracle.kdk_step(x, vel, acc, 0.01, 0.0, p)
    assert np.abs(x - x0).max() < 2e-5


def test_octree_export_invariants(oracle):
    rng = np.random.default_rng(4)
    n = 3000
    pos = np.concatenate([rng.standard_normal((n - 200, 3)),
                          0.001 * rng.standard_normal((200, 3)) + 0.5]).astype(np.float32)
    m = (0.5 + rng.random(n)).astype(np.float32)
    x = oracle.hot(pos, m)
    t = oracle.octree_export(x)
    node, child, count, first = t["node"], t["child"], t["count"], t["first"]
    nn = node.shape[0]
    assert nn > 1
    leaves = np.where((child < 0).all(1))[0]
    # every particle in exactly one leaf
    idx = np.concatenate([t["leaf_idx"][first[i]:first[i] + count[i]] for i in leaves])
    assert np.array_equal(np.sort(idx), np.arange(n))
    for i in range(nn):
        c, h = node[i, :3], node[i, 3]
        kids = child[i][child[i] >= 0]
        if len(kids) == 0:
            pts = t["leaf_idx"][first[i]:first[i] + count[i]]
            # leaf rule: <= 16 particles, or the cell reached the 0.01 half
            # size, or every particle falls in one octant (the clump of 200
            # does); depth 24 is not reached by this cloud
            oct_ = ((pos[pts, 0] >= c[0]).astype(int) + 2 * (pos[pts, 1] >= c[1]) +
                    4 * (pos[pts, 2] >= c[2]))
            assert count[i] <= 16 or h <= 0.01 + 1e-7 or len(set(oct_.tolist())) == 1
            assert node[i, 4] == pytest.approx(m[pts].sum(), rel=1e-5)
            com = (pos[pts].astype(np.float64) * m[pts, None]).sum(0) / m[pts].sum()
            assert np.allclose(node[i, 5:8], com, atol=1e-5)
            continue
        assert np.all(kids > i)  # pre-order numbering
        assert node[i, 4] == pytest.approx(node[kids, 4].sum(), rel=1e-5)
        for o in range(8):
            k = child[i, o]
            if k < 0:
                continue
            # octant bits x = 1, y = 2, z = 4 with the >= test; child half = h / 2
            sgn = np.array([1 if o & 1 else -1, 1 if o & 2 else -1, 1 if o & 4 else -1])
            assert np.allclose(node[k, :3], c + sgn * 0.5 * h, atol=1e-6)
            assert node[k, 3] == pytest.approx(0.5 * h, rel=1e-6)
    # the root cube contains every particle
    assert np.all(np.abs(pos - node[0, :3]) <= node[0, 3] + 1e-6)
