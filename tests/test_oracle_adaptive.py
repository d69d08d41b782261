"""Pins the dyadic block time-step restatement (oracle/adaptive.cpp) by closed
forms: the level rule, free motion, the single-level update, the end-of-step
synchronisation and energy convergence on a circular two-body orbit."""
import numpy as np
import pytest


@pytest.mark.parametrize("a,v,eta,soft,dt,L,vc,want", [
    # eta sqrt(soft/|a|) = 0.5 sqrt(0.01/4) = 0.025: first 1/2^l <= 0.025 is l = 6
    ((4.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.5, 0.01, 1.0, 12, 1, 6),
    # clipped at the deepest level
    ((4.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.5, 0.01, 1.0, 3, 1, 3),
    # |a| below 1e-6: the global step
    ((1e-7, 0.0, 0.0), (0.0, 0.0, 0.0), 0.5, 0.01, 1.0, 12, 1, 0),
    # velocity criterion: eta soft / |v| = 0.5 * 0.01 / 2 = 0.0025 -> l = 9 (1/512)
    ((1e-7, 0.0, 0.0), (0.0, 2.0, 0.0), 0.5, 0.01, 1.0, 12, 1, 9),
    ((1e-7, 0.0, 0.0), (0.0, 2.0, 0.0), 0.5, 0.01, 1.0, 12, 0, 0),
    # exactly on a level boundary: dt / 2^l <= stable is inclusive
    ((1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.25, 1.0, 1.0, 12, 0, 2),
])
def test_level_rule(oracle, a, v, eta, soft, dt, L, vc, want):
    assert oracle.choose_level(a, v, eta, soft, dt, L, vc) == want


def test_free_particles_move_on_straight_lines(oracle):
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(20, 3))
    vel = rng.normal(size=(20, 3))
    st = oracle.AdaptiveState(pos, vel, np.zeros(20), dt=0.125, max_level=4, eta=0.1,
                              soft=0.01, eps=0.01)
    st.steps(3)
    np.testing.assert_allclose(st.pos, pos + vel * 0.375, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(st.vel, vel)


def _direct(pos, mass, eps):
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d * d).sum(-1) + eps * eps
    np.fill_diagonal(r2, np.inf)
    return (d * (mass[None, :] / r2 ** 1.5)[:, :, None]).sum(1)


def test_single_level_update_closed_form(oracle):
    rng = np.random.default_rng(2)
    pos = rng.normal(size=(5, 3))
    vel = rng.normal(size=(5, 3)) * 0.1
    m = rng.uniform(0.5, 1.5, 5)
    dt, eps = 0.01, 0.05
    st = oracle.AdaptiveState(pos, vel, m, dt=dt, max_level=0, eta=0.1, soft=eps, eps=eps)
    st.steps(1)
    a0 = _direct(pos, m, eps)
    xp = pos + vel * dt + a0 * dt * dt / 2
    a1 = _direct(xp, m, eps)
    np.testing.assert_allclose(st.pos, pos + vel * dt + a1 * dt * dt / 2, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.vel, vel + a1 * dt, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(st.acc, a1, rtol=1e-12)


def test_every_particle_is_synchronised_at_the_end_of_a_step(oracle):
    rng = np.random.default_rng(3)
    pos = rng.normal(size=(64, 3)) * 0.3
    vel = rng.normal(size=(64, 3)) * 0.2
    m = np.full(64, 1 / 64)
    st = oracle.AdaptiveState(pos, vel, m, dt=1 / 64, max_level=6, eta=0.05, soft=0.01,
                              eps=0.01)
    st.steps(2)
    assert st.tick.value == 2 * 64
    assert np.all(st.last == st.tick.value)
    assert len(np.unique(st.level)) > 1  # the schedule is genuinely multi-level
    # acc is the force at the predicted positions; the correction then moves
    # each particle by (a' - a) h^2 / 2, a small fraction of a step
    np.testing.assert_allclose(st.acc, _direct(st.pos, m, 0.01), rtol=2e-2, atol=1e-3)


def _two_body_energy_error(oracle, eta):
    # circular orbit, separation 1, masses 0.5 + 0.5: speed 0.5 each, period 2 pi
    pos = np.array([[0.5, 0, 0], [-0.5, 0, 0]])
    vel = np.array([[0, 0.5, 0], [0, -0.5, 0]])
    m = np.array([0.5, 0.5])
    st = oracle.AdaptiveState(pos, vel, m, dt=2 * np.pi / 16, max_level=8, eta=eta, soft=0.1,
                              eps=0.0, velocity_criterion=False)

    def energy(p, v):
        r = np.linalg.norm(p[0] - p[1])
        return 0.5 * (m * (v * v).sum(1)).sum() - m[0] * m[1] / r

    e0 = energy(st.pos, st.vel)
    st.steps(16)
    return abs(energy(st.pos, st.vel) - e0) / abs(e0), st


def test_two_body_energy_converges_at_first_order(oracle):
    # The predictor-corrector uses the force at the predicted end point for
    # both updates (x1 = x0 + v0 h + a1 h^2/2, v1 = v0 + a1 h): first order,
    # so halving the slice roughly halves the energy error of one orbit.
    errs = []
    for eta in (0.2, 0.1, 0.05):
        e, st = _two_body_energy_error(oracle, eta)
        errs.append(e)
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]
    assert errs[2] > 0.2 * errs[1]  # not second order
    assert np.linalg.norm(st.pos[0] - st.pos[1]) == pytest.approx(1.0, rel=0.15)
