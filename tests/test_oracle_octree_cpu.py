"""Pins the oracle's octree_cpu restatement (oracle/octree_cpu.cpp).

Checks, each chosen so a plausible slip (dropped softening, wrong sign,
swapped stack order, wrong clamp) fails one of them:
  * golden vectors of the reference's own build (tests/golden/octree_cpu_ref.npz)
  * oracle/_ref bit-exactness when the reference build is present here
  * two_body_rest.xyz closed form |a| = m d / (d^2 + eps^2)^1.5
  * TST_UNT_PHYS_019: bounds criterion error <= com criterion error + 1e-6
  * theta -> 0.05: octree force == pairwise sum of the same law
  * policy clamps (PhyForceLawPolicy.cpp:22-32) and the acceleration clamp
  * one Euler update (OctParticleSystemHost.cpp:61-69) in closed form
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, read_xyz


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "octree_cpu_ref.npz"))


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-30)


def test_golden_vectors_of_reference_build(oracle, golden):
    # rsqrtss is vendor specific (Intel vs AMD): compare with 1e-6 relative.
    x9 = golden["criterion9_xyzm"]
    for crit in (0, 1):
        p = oracle.make_params(theta=0.9, softening=0.01, criterion=crit)
        a, _ = oracle.octree_forces(x9, p)
        assert _rel(a, golden[f"criterion9_acc_c{crit}"]).max() < 1e-6
    a, _ = oracle.octree_forces(golden["two_body_xyzm"], oracle.make_params(theta=0.5))
    assert _rel(a, golden["two_body_acc"]).max() < 1e-6
    xc = golden["cloud2k_xyzm"]
    for th in (0.5, 0.7):
        for crit in (0, 1):
            p = oracle.make_params(theta=th, softening=0.01, criterion=crit)
            a, _ = oracle.octree_forces(xc, p)
            r = _rel(a, golden[f"cloud2k_acc_t{int(th * 10)}_c{crit}"])
            assert np.median(r) < 1e-6 and r.max() < 1e-4


def test_bit_exact_against_reference_build(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(3)
    for n in (1, 2, 17, 1000, 5000):
        pos = rng.standard_normal((n, 3)).astype(np.float32)
        m = (0.5 + rng.random(n)).astype(np.float32)
        x = oracle.hot(pos, m)
        for crit in (0, 1):
            p = oracle.make_params(theta=0.6, softening=0.02, criterion=crit)
            a, _ = oracle.octree_forces(x, p)
            assert np.array_equal(a, oracle.ref_octree_forces(x, p))


def test_two_body_closed_form(oracle):
    pos, mass = read_xyz(os.path.join(GOLDEN, "two_body_rest.xyz"))
    x = oracle.hot(pos, mass)
    eps = 0.01
    a, _ = oracle.octree_forces(x, oracle.make_params(theta=0.5, softening=eps))
    d = 2.0
    expect = mass[1] * d / (d * d + eps * eps) ** 1.5
    assert a[0, 0] == pytest.approx(expect, rel=1e-6)
    assert a[1, 0] == pytest.approx(-expect, rel=1e-6)
    assert np.all(a[:, 1:] == 0.0)


def test_criterion_ordering_fixture(oracle, golden):
    # TST_UNT_PHYS_019 (octree_criteria.cpp:91-112)
    x9 = golden["criterion9_xyzm"]
    pos = x9[:, :3].astype(np.float64)
    m = x9[:, 3].astype(np.float64)
    ref = oracle.direct_fp64(pos, m, [0], eps=0.01, min_d2=1e-12)[0]
    errs = []
    for crit in (0, 1):
        a, _ = oracle.octree_forces(x9, oracle.make_params(theta=0.9, softening=0.01,
                                                           criterion=crit))
        errs.append(np.linalg.norm(a[0] - ref))
    assert errs[1] <= errs[0] + 1e-6


def test_small_theta_equals_pairwise(oracle):
    rng = np.random.default_rng(5)
    n = 300
    pos = rng.standard_normal((n, 3))
    m = rng.random(n) + 0.1
    x = oracle.hot(pos, m)
    a, _ = oracle.octree_forces(x, oracle.make_params(theta=0.05, softening=0.05))
    ref = oracle.direct_fp64(x[:, :3].astype(np.float64), x[:, 3].astype(np.float64),
                             np.arange(n), eps=0.05)
    assert _rel(a, ref).max() < 2e-5


def test_policy_resolution(oracle):
    t, s, d = oracle.resolve_policy(oracle.make_params(theta=0.01, softening=0.0, min_softening=0.003,
                                                       min_theta=0.2, min_distance2=-1.0))
    assert t == pytest.approx(0.2) and s == pytest.approx(0.003) and d == 0.0
    t, _, _ = oracle.resolve_policy(oracle.make_params(theta=9.0, min_theta=0.0))
    assert t == pytest.approx(4.0)
    t, _, _ = oracle.resolve_policy(oracle.make_params(theta=0.3, min_theta=0.0))
    assert t == pytest.approx(0.3)  # min_theta clamps up to 0.05 first


def test_acceleration_clamp(oracle):
    pos = np.array([[0, 0, 0], [0.001, 0, 0]], np.float64)
    x = oracle.hot(pos, [1.0, 1.0])
    a, _ = oracle.octree_forces(x, oracle.make_params(softening=1e-4, max_acceleration=64.0))
    assert np.linalg.norm(a[0]) == pytest.approx(64.0, rel=1e-6)
    assert a[0, 0] > 0 and a[1, 0] < 0


def test_euler_update_closed_form(oracle):
    pos, mass = read_xyz(os.path.join(GOLDEN, "two_body_rest.xyz"))
    x = oracle.hot(pos, mass)
    v = np.zeros((2, 3), np.float32)
    dt = 0.01
    acc = oracle.euler_update(x, v, dt, max_speed=0.0, params=oracle.make_params(softening=0.01))
    a0 = acc[0, 0]
    assert v[0, 0] == pytest.approx(a0 * dt, rel=1e-6)
    assert x[0, 0] == pytest.approx(-1.0 + a0 * dt * dt, rel=1e-6)
    # velocity clamp
    x = oracle.hot(pos, mass)
    v = np.array([[5.0, 0, 0], [0, 0, 0]], np.float32)
    oracle.euler_update(x, v, dt, max_speed=1.0)
    assert np.linalg.norm(v[0]) == pytest.approx(1.0, rel=1e-6)
