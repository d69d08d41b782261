"""Pins the FP64 particle-mesh oracle (oracle/pm.py) by what the split fixes:
the long-range force of a point mass, momentum conservation, the force-free
uniform lattice and exact shift covariance."""
import math

import numpy as np
import pytest

from oracle import pm


def test_point_mass_long_range_force():
    # A unit mass at the box centre, probes along a diagonal-free direction:
    # a(r) = m (1 - w(r)) / r^2 toward the mass for r_s << r << L / 2.
    L, N = 1.0, 64
    h = L / N
    rs = 1.25 * h
    src = np.array([[0.5, 0.5, 0.5]]) + 0.5 * h  # a cell corner: symmetric CIC
    dirn = np.array([0.8, 0.36, 0.48])
    dirn /= np.linalg.norm(dirn)
    rs_list = [4 * rs, 6 * rs, 8 * rs, 0.08]
    probes = src + np.outer(rs_list, dirn)
    pos = np.vstack([src, probes])
    mass = np.array([1.0] + [0.0] * len(rs_list))
    acc = pm.long_range_accel(pos, mass, L, N, rs)[1:]
    for r, a in zip(rs_list, acc):
        want = (1.0 - pm.short_range_weight(r, rs)) / r ** 2
        radial = -a @ dirn
        assert radial == pytest.approx(want, rel=0.03), r
        assert np.linalg.norm(a + radial * dirn) < 0.03 * want  # no tangential part


def test_total_force_vanishes():
    rng = np.random.default_rng(2)
    pos = rng.random((300, 3))
    m = rng.uniform(0.5, 1.5, 300)
    acc = pm.long_range_accel(pos, m, 1.0, 32, 1.25 / 32)
    f = (m[:, None] * acc).sum(0)
    assert np.abs(f).max() < 1e-10 * np.abs(m[:, None] * acc).sum()


def test_uniform_lattice_is_force_free():
    N = 16
    g = (np.arange(N) + 0.5) / N
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    acc = pm.long_range_accel(pos, np.full(len(pos), 1.0 / len(pos)), 1.0, N, 1.25 / N)
    assert np.abs(acc).max() < 1e-12


def test_whole_cell_shift_is_exact():
    rng = np.random.default_rng(3)
    pos = rng.random((200, 3))
    m = rng.uniform(0.5, 1.5, 200)
    N = 32
    a = pm.long_range_accel(pos, m, 1.0, N, 1.25 / N)
    b = pm.long_range_accel(np.mod(pos + np.array([3, 5, 11]) / N, 1.0), m, 1.0, N, 1.25 / N)
    np.testing.assert_allclose(b, a, rtol=1e-9, atol=1e-12 * np.abs(a).max())


def test_weight_complement():
    rs = 0.02
    assert pm.short_range_weight(1e-9, rs) == pytest.approx(1.0)
    assert pm.short_range_weight(10 * rs, rs) < 1e-9
    assert 0 < pm.short_range_weight(2 * rs, rs) < 1
    assert math.isclose(pm.short_range_weight(2 * rs, rs),
                        math.erfc(1.0) + 2 / math.sqrt(math.pi) * math.exp(-1.0))
