"""Pins the FP64 learned-far-field oracle (oracle/mlp.py) by what the
mathematics fixes: finite differences, closed forms, invariances."""
import numpy as np
import pytest

from oracle import mlp


def _params(rng, scale=0.3):
    p = rng.normal(0.0, scale, mlp.NP)
    return p


def test_loss_grad_matches_central_differences_everywhere():
    rng = np.random.default_rng(5)
    p = _params(rng)
    X = rng.normal(size=(7, mlp.NF))
    Y = rng.normal(size=(7, mlp.NO))
    L, g = mlp.loss_grad(p, X, Y)
    assert L == pytest.approx(mlp.loss(p, X, Y), rel=1e-14)
    h = 1e-6
    fd = np.empty(mlp.NP)
    for k in range(mlp.NP):
        pp = p.copy()
        pp[k] += h
        lp = mlp.loss(pp, X, Y)
        pp[k] -= 2 * h
        fd[k] = (lp - mlp.loss(pp, X, Y)) / (2 * h)
    err = np.abs(fd - g)
    assert err.max() <= 1e-7 * max(1.0, np.abs(g).max())
    # every block of the layout is exercised
    for lo, hi in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                   (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.NP)):
        assert np.abs(g[lo:hi]).max() > 1e-6


def test_zero_weight_network_closed_form():
    rng = np.random.default_rng(6)
    X = rng.normal(size=(11, mlp.NF))
    Y = rng.normal(size=(11, mlp.NO))
    b1 = rng.normal(size=mlp.NH)
    b2 = rng.normal(size=mlp.NH)
    b3 = rng.normal(size=mlp.NO)
    p = mlp.pack(np.zeros((mlp.NF, mlp.NH)), b1, np.zeros((mlp.NH, mlp.NH)), b2,
                 np.zeros((mlp.NH, mlp.NO)), b3)
    L, g = mlp.loss_grad(p, X, Y)
    e = b3[None, :] - Y
    assert L == pytest.approx(np.mean(np.sum(e * e, 1)), rel=1e-14)
    W1, gb1, W2, gb2, gW3, gb3 = mlp.unpack(g)
    np.testing.assert_allclose(gb3, 2 * e.mean(0), rtol=1e-13)
    h2 = b2 / (1 + np.exp(-b2))  # silu(b2)
    np.testing.assert_allclose(gW3, np.outer(h2, 2 * e.mean(0)), rtol=1e-12)
    assert np.all(W1 == 0) and np.all(gb1 == 0) and np.all(gb2 == 0)


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    np.testing.assert_allclose(mlp.dsilu(z), (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h),
                               atol=1e-9)


def test_bf16_rounding():
    torch = pytest.importorskip("torch")
    u = 2.0 ** -7
    assert mlp.to_bf16(1.0) == 1.0
    assert mlp.to_bf16(1.0 + u / 2) == 1.0                 # tie -> even
    assert mlp.to_bf16(1.0 + 1.5 * u) == 1.0 + 2 * u       # tie -> even
    assert mlp.to_bf16(1.0 + 0.75 * u) == 1.0 + u
    assert mlp.to_bf16(-3.0) == -3.0
    rng = np.random.default_rng(1)
    x = (rng.normal(size=20000) * 10.0 ** rng.uniform(-6, 6, 20000)).astype(np.float32)
    t = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.to_bf16(x), t)


def test_bf16_forward_within_rounding_of_fp64():
    rng = np.random.default_rng(8)
    p = _params(rng, 0.15)
    X = rng.normal(size=(500, mlp.NF))
    y = mlp.forward(p, X)
    yb = mlp.forward_bf16(p, X)
    scale = np.sqrt(np.mean(y * y))
    assert np.abs(yb - y).max() < 0.05 * scale
    assert np.abs(yb - y).max() > 0  # the emulation does round


def test_residual_single_particle_node_is_zero():
    x = np.array([[3.0, -1.0, 2.0], [0.5, 0.2, -0.1]])
    r = mlp.node_residual(x, np.array([[0.1, 0.2, 0.3]]), np.array([2.5]), 0.01)
    assert np.abs(r).max() < 1e-13


@pytest.mark.parametrize("u", [0.3, 0.1, 0.02])
def test_residual_two_particle_closed_forms(u):
    a = 0.7
    D = a / u
    pos = np.array([[0, 0, a], [0, 0, -a]])
    m = np.array([1.3, 1.3])
    # on the axis: normalised exact = -((1-u)^-2 + (1+u)^-2)/2, multipole = -(1 + 3u^2)
    r_ax = mlp.node_residual(np.array([[0, 0, D]]), pos, m, 0.0)[0]
    want = -(((1 - u) ** -2 + (1 + u) ** -2) / 2 - 1 - 3 * u * u)
    assert r_ax[2] == pytest.approx(want, rel=1e-7, abs=1e-15)
    assert abs(r_ax[0]) < 1e-15 and abs(r_ax[1]) < 1e-15
    # broadside: exact = -(1+u^2)^-3/2, multipole = -(1 - 3u^2/2)
    r_bs = mlp.node_residual(np.array([[D, 0, 0]]), pos, m, 0.0)[0]
    want = -((1 + u * u) ** -1.5 - 1 + 1.5 * u * u)
    assert r_bs[0] == pytest.approx(want, rel=1e-7, abs=1e-15)


def test_residual_falls_off_as_cube_of_size_over_distance():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.5, 0.5, (40, 3))
    m = rng.uniform(0.5, 1.5, 40)
    _, com, _, _ = mlp.node_moments(pos, m)
    dirn = np.array([0.6, -0.48, 0.64])
    r1 = np.linalg.norm(mlp.node_residual(com + 60.0 * dirn, pos, m, 0.0))
    r2 = np.linalg.norm(mlp.node_residual(com + 120.0 * dirn, pos, m, 0.0))
    assert 7.0 < r1 / r2 < 9.0


def test_features_and_residual_invariances():
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.5, 0.5, (25, 3))
    m = rng.uniform(0.5, 1.5, 25)
    x = np.array([[1.2, 0.4, -0.3], [0.1, 1.0, 0.9]])
    M, com, q, _ = mlp.node_moments(pos, m)
    o = mlp.node_third_moments(pos, m)
    centre = np.zeros(3)
    f = mlp.node_features(x, com, M, q, 1.0, centre, 0.8, 25, 0.02, o)
    r = mlp.node_residual(x, pos, m, 0.02)
    k, c, t = 3.7, 0.25, np.array([5.0, -2.0, 1.0])
    M2, com2, q2, _ = mlp.node_moments(k * pos + t, c * m)
    o2 = mlp.node_third_moments(k * pos + t, c * m)
    f2 = mlp.node_features(k * x + t, com2, M2, q2, k, k * centre + t, k * 0.8, 25, k * 0.02, o2)
    r2 = mlp.node_residual(k * x + t, k * pos + t, c * m, k * 0.02)
    np.testing.assert_allclose(f2, f, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(r2, r, rtol=1e-9, atol=1e-13)
    # f15 = |f12..f14|; f6 = -(f1 + f4) (traceless); f0 = log2(count)/16; pads
    np.testing.assert_allclose(f[:, 15], np.linalg.norm(f[:, 12:15], axis=1), rtol=1e-14)
    np.testing.assert_allclose(f[:, 6], -(f[:, 1] + f[:, 4]), rtol=1e-12, atol=1e-15)
    assert f[0, 0] == pytest.approx(np.log2(25) / 16)
    assert np.all(f[:, list(mlp.PADS)] == 0.0)
    d2 = np.sum((x - com) ** 2, axis=1)
    np.testing.assert_allclose(f[:, 28], 0.02 ** 2 / (d2 + 0.02 ** 2), rtol=1e-14)
    # the octupole features are the expansion's third-order term: at a
    # distant target the second-order residual is that term to within
    # (size / distance)^4
    far = com + 40.0 * np.array([[0.36, 0.48, 0.8]])
    ff = mlp.node_features(far, com, M, q, 1.0, centre, 0.8, 25, 0.0, o)[0]
    rf = mlp.node_residual(far, pos, m, 0.0)[0]
    assert np.linalg.norm(rf - ff[29:32]) < 0.05 * np.linalg.norm(rf)


def test_third_moments_two_particle_closed_form_and_parallel_axis():
    # m1 at +a1 e_z, m2 at -a2 e_z with m1 a1 = m2 a2 (COM at the origin):
    # T_zzz = m1 a1^3 - m2 a2^3, every other component zero
    m1, a1, a2 = 1.5, 0.4, 0.6
    m2 = m1 * a1 / a2
    o = mlp.node_third_moments(np.array([[0, 0, a1], [0, 0, -a2]]), np.array([m1, m2]))
    want = np.zeros(10)
    want[9] = m1 * a1 ** 3 - m2 * a2 ** 3
    np.testing.assert_allclose(o, want, atol=1e-15)
    # parallel-axis shift of two children (what the tree's upward pass does)
    rng = np.random.default_rng(9)
    pa = rng.normal(size=(30, 3)) * 0.2 + [0.3, 0.1, -0.2]
    pb = rng.normal(size=(20, 3)) * 0.1 + [-0.4, 0.2, 0.5]
    ma, mb = rng.uniform(0.5, 1.5, 30), rng.uniform(0.5, 1.5, 20)
    allp, allm = np.vstack([pa, pb]), np.concatenate([ma, mb])
    _, com, _, _ = mlp.node_moments(allp, allm)
    shifted = np.zeros(10)
    for p_, m_ in ((pa, ma), (pb, mb)):
        Mc, cc, qc, trs = mlp.node_moments(p_, m_)
        S = (np.array([[qc[0], qc[1], qc[2]], [qc[1], qc[3], qc[4]], [qc[2], qc[4], qc[5]]])
             + trs * np.eye(3)) / 3.0
        D = cc - com
        Tc = mlp.node_third_moments(p_, m_)
        for k, (i, j, l) in enumerate(mlp._OCT_IDX):
            shifted[k] += (Tc[k] + D[i] * S[j, l] + D[j] * S[i, l] + D[l] * S[i, j]
                           + Mc * D[i] * D[j] * D[l])
    np.testing.assert_allclose(shifted, mlp.node_third_moments(allp, allm), rtol=1e-12,
                               atol=1e-14)


@pytest.mark.parametrize("u", [0.2, 0.05])
def test_octupole_term_on_axis_and_order(u):
    # on the axis of the pair above the u^3 term of -sum m / (D -+ a)^2 is
    # -4 T_zzz / D^5
    m1, a1, a2 = 1.5, 0.4, 0.6
    m2 = m1 * a1 / a2
    pos = np.array([[0, 0, a1], [0, 0, -a2]])
    o = mlp.node_third_moments(pos, np.array([m1, m2]))
    D = a2 / u
    got = mlp.octupole_accel(np.array([[0, 0, D]]), np.zeros(3), o)[0]
    assert got[2] == pytest.approx(-4 * o[9] / D ** 5, rel=1e-12)
    assert abs(got[0]) < 1e-15 and abs(got[1]) < 1e-15
    # a random node: (second-order residual - octupole term) is fourth order
    rng = np.random.default_rng(11)
    p = rng.uniform(-0.5, 0.5, (40, 3))
    m = rng.uniform(0.5, 1.5, 40)
    M, com, _, _ = mlp.node_moments(p, m)
    oo = mlp.node_third_moments(p, m)
    dirn = np.array([0.6, -0.48, 0.64])
    rest = []
    for dist in (30.0, 60.0):
        x = com + dist * dirn
        mono = M / dist ** 2
        res = mlp.node_residual(x, p, m, 0.0)[0] * mono
        rest.append(np.linalg.norm(res - mlp.octupole_accel(x, com, oo)[0]) / mono)
    assert 12.0 < rest[0] / rest[1] < 20.0


@pytest.mark.parametrize("eps", [0.0, 0.3, 1.0, 4.0])
def test_softened_expansion_two_particle_closed_form(eps):
    # masses m at com +- a e_z, target on the axis at distance D: the exact
    # force is -m [f(D - a) + f(D + a)], f(z) = z (z^2 + eps^2)^-3/2, whose
    # expansion to second order in a is -m [2 f(D) + a^2 f''(D)] with
    # f''(z) = z (6 z^2 - 9 eps^2) (z^2 + eps^2)^-7/2.
    m, a, D = 0.7, 0.05, 1.3
    pos = np.array([[0, 0, a], [0, 0, -a]])
    M, com, q, trs = mlp.node_moments(pos, np.array([m, m]))
    assert trs == pytest.approx(2 * m * a * a)
    got = mlp.multipole_accel(np.array([[0, 0, D]]), com, M, q, eps, trs)[0]
    f = D * (D * D + eps * eps) ** -1.5
    f2 = D * (6 * D * D - 9 * eps * eps) * (D * D + eps * eps) ** -3.5
    want = -m * (2 * f + a * a * f2)
    assert got[2] == pytest.approx(want, rel=1e-12)
    assert abs(got[0]) < 1e-15 and abs(got[1]) < 1e-15
    # broadside (target at D e_x): exact -2 m D ((D^2 + a^2 + eps^2))^-3/2,
    # second order in a: -2 m D [R^-3 - 3/2 a^2 R^-5], R^2 = D^2 + eps^2
    got = mlp.multipole_accel(np.array([[D, 0, 0]]), com, M, q, eps, trs)[0]
    R2 = D * D + eps * eps
    want = -2 * m * D * (R2 ** -1.5 - 1.5 * a * a * R2 ** -2.5)
    assert got[0] == pytest.approx(want, rel=1e-12)
    # without the eps^2 trs term the softened expansion is wrong at second order
    if eps > 0:
        bad = mlp.multipole_accel(np.array([[D, 0, 0]]), com, M, q, eps, 0.0)[0]
        assert abs(bad[0] - want) > 1e-3 * abs(want - (-2 * m * D * R2 ** -1.5))
