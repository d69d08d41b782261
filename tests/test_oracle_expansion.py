"""Pins the FP64 local expansion of a softened node field (oracle/expansion.py)
by convergence orders against the direct sum, harmonic structure, symmetry and
agreement with the walk's own second-order cell formula."""
import numpy as np
import pytest

from oracle import expansion as ex
from oracle import mlp


def _cluster(seed, n=30, size=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, (n, 3)) * size
    m = rng.uniform(0.5, 1.5, n)
    M, com, q, trs = mlp.node_moments(pos, m)
    return pos, m, M, com, q, trs


@pytest.mark.parametrize("eps", [0.0, 0.3])
def test_single_mass_taylor_error_is_third_order(eps):
    r = np.array([3.0, -1.0, 2.0])
    A, J, H = ex.local_expansion(1.7, r, np.zeros(6), 0.0, eps)
    dirn = np.array([0.2, 0.7, -0.68])
    errs = []
    for h in (0.2, 0.1, 0.05):
        x = h * dirn
        exact = ex.direct(x, r[None, :], [1.7], eps)
        errs.append(np.linalg.norm(ex.evaluate(A, J, H, x) - exact))
    assert 7.0 < errs[0] / errs[1] < 9.0 and 7.0 < errs[1] / errs[2] < 9.0


def test_centre_value_is_the_walk_cell_formula():
    pos, m, M, com, q, trs = _cluster(1)
    x0 = com + np.array([4.0, 1.0, -2.5])
    for eps in (0.0, 0.4):
        A, _, _ = ex.local_expansion(M, com - x0, q, trs, eps)
        want = mlp.multipole_accel(x0, com, M, q, eps, trs)[0]
        np.testing.assert_allclose(A, want, rtol=1e-12)


def test_harmonic_and_symmetric_without_softening():
    pos, m, M, com, q, trs = _cluster(2)
    A, J, H = ex.local_expansion(M, np.array([3.0, 2.0, -4.0]), q, trs, 0.0)
    assert abs(np.trace(J)) < 1e-12 * np.abs(J).max()  # div a = 0 outside the sources
    np.testing.assert_allclose(J, J.T, rtol=1e-12, atol=1e-15)
    for perm in ((1, 0, 2), (0, 2, 1), (2, 1, 0)):
        np.testing.assert_allclose(H, np.transpose(H, perm), atol=1e-15)
    np.testing.assert_allclose(np.einsum("ijj->i", H), 0.0, atol=1e-12 * np.abs(H).max())


def test_cluster_field_error_orders():
    # halving the group offset delta: the local Taylor error falls ~8x; moving the
    # cluster twice as far (same delta / d): the whole error falls ~2^4 x
    pos, m, M, com, q, trs = _cluster(3, size=0.5)
    dirn = np.array([0.6, -0.64, 0.48])
    d = 6.0
    x0 = com + d * dirn
    A, J, H = ex.local_expansion(M, com - x0, q, trs, 0.0)
    probe = np.array([0.3, 0.2, -0.25])
    errs = []
    for h in (1.0, 0.5, 0.25):
        x = x0 + h * probe
        errs.append(np.linalg.norm(ex.evaluate(A, J, H, h * probe)[0] - ex.direct(x, pos, m, 0.0)[0]))
    assert errs[1] < errs[0] and errs[2] < errs[1]
    rel = errs[0] / np.linalg.norm(ex.direct(x0, pos, m, 0.0)[0])
    assert rel < 3e-3


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_tpm_cell_expansion_is_the_newton_one_without_a_split(eps):
    # r_s -> infinity: w = 1, the softened Newtonian second-order expansion
    pos, m, M, com, q, trs = _cluster(3, size=0.3)
    x = com + np.array([[1.3, -0.4, 0.8], [-0.9, 1.1, 0.2]])
    got = ex.tpm_cell_accel(x, com, M, q, trs, eps, 1e7)
    want = mlp.multipole_accel(x, com, M, q, eps, trs)
    np.testing.assert_allclose(got, want, rtol=1e-6)


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_tpm_cell_expansion_error_is_third_order(eps):
    # a cluster of size a at 1.5 r_s from the target: the error against the
    # direct split sum falls by ~8 per halving of a (second-order expansion),
    # and a single particle is exact
    rs = 1.0
    dirn = np.array([0.48, 0.6, 0.64])
    errs = []
    for size in (0.4, 0.2, 0.1):
        pos, m, M, com, q, trs = _cluster(4, n=25, size=size)
        x = com - 1.5 * rs * dirn
        exact = ex.tpm_direct(x, pos, m, eps, rs)
        got = ex.tpm_cell_accel(x, com, M, q, trs, eps, rs)
        errs.append(np.linalg.norm(got - exact) / np.linalg.norm(exact))
    assert 6.0 < errs[0] / errs[1] < 10.0 and 6.0 < errs[1] / errs[2] < 10.0
    one = ex.tpm_cell_accel(np.zeros((1, 3)), dirn, 1.3, np.zeros(6), 0.0, eps, rs)
    np.testing.assert_allclose(one, ex.tpm_direct(np.zeros((1, 3)), dirn[None, :], [1.3], eps, rs),
                               rtol=1e-9)
    # multiplying the Newtonian expansion by w(R) instead is only second
    # order: its error falls by ~4.5 per halving where the expansion's falls
    # by ~8
    prod_err, cell_err = [], []
    for size in (0.1, 0.05):
        pos, m, M, com, q, trs = _cluster(4, n=25, size=size)
        x = com - 1.5 * rs * dirn
        exact = ex.tpm_direct(x, pos, m, eps, rs)
        prod = mlp.multipole_accel(x, com, M, q, eps, trs) * ex.tpm_weight(
            float(np.sqrt(1.5 ** 2 + eps * eps)), rs)
        prod_err.append(np.linalg.norm(prod - exact))
        cell_err.append(np.linalg.norm(ex.tpm_cell_accel(x, com, M, q, trs, eps, rs) - exact))
    assert prod_err[0] / prod_err[1] < 5.5 < 6.5 < cell_err[0] / cell_err[1]
