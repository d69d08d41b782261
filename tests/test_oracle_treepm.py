"""Pins the TreePM short-range weight of the oracle (octree_cpu restatement
with short_range_scale, and the FP64 erfc direct sum) by closed forms:
w(0) -> 1, w decays to ~0 beyond a few r_s, and the weighted pair force of
the octree equals the FP64 weighted direct sum at small theta."""
import math

import numpy as np
import pytest


def w(r, rs):
    a = 0.5 * r / rs
    return math.erfc(a) + r / (rs * math.sqrt(math.pi)) * math.exp(-a * a)


def test_weight_closed_form(oracle):
    rs = 0.1
    for d in (0.001, 0.05, 0.2, 0.45):
        p = oracle.make_params(softening=1e-4, short_range_scale=rs)
        a = oracle.pair_accel([0, 0, 0], [d, 0, 0], 1.0, p)[0]
        r = math.sqrt(d * d + 1e-8)
        assert a == pytest.approx(d / r ** 3 * w(r, rs), rel=2e-5)
    assert w(1e-9, rs) == pytest.approx(1.0)
    assert w(4.5 * rs * 2, rs) < 1e-8  # beyond the 4.5 r_s cut (in units of 2 r_s)


def test_octree_short_range_equals_direct(oracle):
    rng = np.random.default_rng(4)
    n = 3000
    pos = rng.random((n, 3))
    m = np.full(n, 1.0 / n)
    rs, rc, eps = 0.02, 0.09, 0.002
    x = oracle.hot(pos, m)
    p = oracle.make_params(theta=0.05, softening=eps, short_range_scale=rs, cutoff2=rc * rc)
    t = np.arange(0, n, 11)
    a = oracle.octree_forces_subset(x, t, p)
    ref = oracle.direct_tpm_fp64(x[:, :3].astype(np.float64), x[:, 3].astype(np.float64), t,
                                 eps, rs, rc)
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(err) < 1e-5


def test_periodic_two_particles_across_the_boundary(oracle):
    # L = 1: particles at x = 0.01 and x = 0.99 are 0.02 apart through the face
    rs, rc, eps = 0.05, 0.2, 1e-3
    pos = np.array([[0.01, 0.5, 0.5], [0.99, 0.5, 0.5]])
    m = np.array([1.0, 2.0])
    a = oracle.direct_tpm_fp64(pos, m, np.array([0, 1]), eps, rs, rc, box=1.0)
    r = math.sqrt(0.02 ** 2 + eps * eps)
    mag = 0.02 / r ** 3 * w(r, rs)
    assert a[0, 0] == pytest.approx(-2.0 * mag, rel=1e-9)  # pulled through x = 0
    assert a[1, 0] == pytest.approx(1.0 * mag, rel=1e-9)
    assert np.all(np.abs(a[:, 1:]) < 1e-12)
    # without the box the pair is 0.98 apart: beyond the cutoff
    a_open = oracle.direct_tpm_fp64(pos, m, np.array([0, 1]), eps, rs, rc)
    assert np.all(a_open == 0.0)


def test_periodic_translation_invariance(oracle):
    rng = np.random.default_rng(9)
    n = 2000
    pos = rng.random((n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    rs, rc, eps = 0.02, 0.09, 0.002
    t = np.arange(0, n, 7)
    a = oracle.direct_tpm_fp64(pos, m, t, eps, rs, rc, box=1.0)
    shifted = np.mod(pos + np.array([0.37, 0.61, 0.13]), 1.0)
    b = oracle.direct_tpm_fp64(shifted, m, t, eps, rs, rc, box=1.0)
    np.testing.assert_allclose(b, a, rtol=1e-9, atol=1e-12 * np.abs(a).max())
    # far from every face the box changes nothing
    inner = np.arange(n)[np.all((pos > 0.1) & (pos < 0.9), axis=1)][:50]
    np.testing.assert_array_equal(
        oracle.direct_tpm_fp64(pos, m, inner, eps, rs, rc, box=1.0),
        oracle.direct_tpm_fp64(pos, m, inner, eps, rs, rc))
