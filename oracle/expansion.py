"""FP64 local (Taylor) expansion of a node's softened field about a group
centre (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module, as the checker of the far-field path of
blitzar_amd/csrc/k_walk.hip. It has no counterpart in the reference (whose
walks evaluate every accepted node per target); it is the standard
cell-to-local step of a fast multipole method written for the
Plummer-softened kernel g(u) = (|u|^2 + eps^2)^-1/2:

  a_i(x0 + delta) = A_i + J_ij delta_j + 1/2 H_ijk delta_j delta_k + O(delta^3),
  with r = com - x0, R^2 = |r|^2 + eps^2, S = (Q + trs I) / 3 the second
  moment about the COM (Q traceless, trs = sum m |y|^2):
  monopole   A_i = -M d_i g,  J_ij = M d_ij g,  H_ijk = -M d_ijk g
  second     A_i += -1/2 S_jk d_ijk g,  J_il += 1/2 S_jk d_ijkl g
  d_i g = -r_i R^-3, d_ij g = 3 r_i r_j R^-5 - delta_ij R^-3,
  d_ijk g = -15 r_i r_j r_k R^-7 + 3 (delta_ij r_k + delta_ik r_j + delta_jk r_i) R^-5,
  d_ijkl g = 105 r_i r_j r_k r_l R^-9 - 15 (6 terms delta r r) R^-7
             + 3 (delta_ij delta_kl + delta_ik delta_jl + delta_il delta_jk) R^-5.

Pinned in tests/test_oracle_expansion.py by convergence orders against the
direct sum, the harmonic (trace-free) structure at eps = 0, symmetry, and
A matching the walk's own second-order cell formula at delta = 0.
"""
from __future__ import annotations

import itertools

import numpy as np

I3 = np.eye(3)


def _derivs(r, eps):
    r = np.asarray(r, np.float64)
    R2 = r @ r + eps * eps
    R = np.sqrt(R2)
    g1 = -r / R ** 3
    g2 = 3 * np.einsum("i,j->ij", r, r) / R ** 5 - I3 / R ** 3
    g3 = -15 * np.einsum("i,j,k->ijk", r, r, r) / R ** 7 + 3 * (
        np.einsum("ij,k->ijk", I3, r) + np.einsum("ik,j->ijk", I3, r) +
        np.einsum("jk,i->ijk", I3, r)) / R ** 5
    rr = np.einsum("i,j->ij", r, r)
    t6 = np.zeros((3, 3, 3, 3))
    for (a, b), (c, d) in (((0, 1), (2, 3)), ((0, 2), (1, 3)), ((0, 3), (1, 2)),
                           ((1, 2), (0, 3)), ((1, 3), (0, 2)), ((2, 3), (0, 1))):
        # delta over index positions (a, b), r r over (c, d)
        for i, j, k, l in itertools.product(range(3), repeat=4):
            ind = (i, j, k, l)
            t6[ind] += I3[ind[a], ind[b]] * rr[ind[c], ind[d]]
    t3 = (np.einsum("ij,kl->ijkl", I3, I3) + np.einsum("ik,jl->ijkl", I3, I3) +
          np.einsum("il,jk->ijkl", I3, I3))
    g4 = 105 * np.einsum("i,j,k,l->ijkl", r, r, r, r) / R ** 9 - 15 * t6 / R ** 7 + 3 * t3 / R ** 5
    return g1, g2, g3, g4


def quad_matrix(q6):
    qxx, qxy, qxz, qyy, qyz, qzz = q6
    return np.array([[qxx, qxy, qxz], [qxy, qyy, qyz], [qxz, qyz, qzz]], np.float64)


def local_expansion(M, r, q6, trs, eps):
    """(A[3], J[3,3], H[3,3,3]) of a node (mass M, COM at r from the expansion
    centre, traceless quadrupole q6 = xx xy xz yy yz zz, trs) at the centre."""
    g1, g2, g3, g4 = _derivs(r, eps)
    S = (quad_matrix(q6) + trs * I3) / 3.0
    A = -M * g1 - 0.5 * np.einsum("jk,ijk->i", S, g3)
    J = M * g2 + 0.5 * np.einsum("jk,ijkl->il", S, g4)
    H = -M * g3
    return A, J, H


def evaluate(A, J, H, delta):
    delta = np.atleast_2d(np.asarray(delta, np.float64))
    return (A[None, :] + delta @ J.T + 0.5 * np.einsum("ijk,nj,nk->ni", H, delta, delta))


def direct(x, pos, mass, eps):
    x = np.atleast_2d(np.asarray(x, np.float64))
    d = np.asarray(pos, np.float64)[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(-1) + eps * eps
    return (d * (np.asarray(mass, np.float64)[None, :] / r2 ** 1.5)[:, :, None]).sum(1)


# ---- TreePM short-range cells ------------------------------------------------
# The second-order expansion about a node's COM of the split short-range
# kernel F(d) = g(R) d, g = w(R) / R^3, R^2 = |d|^2 + eps^2, with the erfc
# weight w(R) = erfc(R / 2 r_s) + R / (r_s sqrt(pi)) exp(-R^2 / 4 r_s^2)
# (engine/physics/octree/force/OctreeForce.cpp:49-57). Summed over the node's
# particles at com + y (sum m y = 0, S = sum m y y^T):
#   a = M F(d) + 1/2 S_jk d_j d_k F(d),
# with d_j F_i = G1 d_i d_j + g delta_ij and
# d_k d_j F_i = G2 d_i d_j d_k + G1 (delta_jk d_i + delta_ik d_j + delta_ij d_k),
# G1 = g'(R) / R, G2 = G1'(R) / R. The derivatives of g are taken here by
# central differences of g in R (step 1e-4 R), so this restatement does not
# share the closed forms it checks.

def tpm_weight(R, rs):
    from math import erfc, exp, pi, sqrt
    return erfc(0.5 * R / rs) + R / (rs * sqrt(pi)) * exp(-R * R / (4.0 * rs * rs))


def tpm_cell_accel(x, com, mass, quad6, trs, eps, rs):
    """Second-order expansion of the split kernel of a node (mass, COM,
    traceless quadrupole quad6 = xx xy xz yy yz zz, trs = sum m |y|^2) at the
    targets x; d = com - x."""
    x = np.atleast_2d(np.asarray(x, np.float64))
    qxx, qxy, qxz, qyy, qyz, qzz = np.asarray(quad6, np.float64)
    Q = np.array([[qxx, qxy, qxz], [qxy, qyy, qyz], [qxz, qyz, qzz]])
    S = (Q + trs * I3) / 3.0
    g = lambda R: tpm_weight(R, rs) / R ** 3
    out = np.zeros((x.shape[0], 3))
    for n, xi in enumerate(x):
        d = np.asarray(com, np.float64) - xi
        R = float(np.sqrt(d @ d + eps * eps))
        h = 1e-4 * R
        g0 = g(R)
        gp = (g(R + h) - g(R - h)) / (2 * h)
        G1 = gp / R
        G1f = lambda r: (g(r + h) - g(r - h)) / (2 * h) / r
        G2 = (G1f(R + h) - G1f(R - h)) / (2 * h) / R
        dSd = d @ S @ d
        trS = np.trace(S)
        Sd = S @ d
        out[n] = mass * g0 * d + 0.5 * (G2 * dSd * d + G1 * (trS * d + 2.0 * Sd))
    return out


def tpm_direct(x, pos, mass, eps, rs):
    """FP64 split-kernel force of particles at x (no cutoff)."""
    x = np.atleast_2d(np.asarray(x, np.float64))
    out = np.zeros((x.shape[0], 3))
    for n, xi in enumerate(x):
        d = np.asarray(pos, np.float64) - xi
        R = np.sqrt(np.sum(d * d, axis=1) + eps * eps)
        w = np.array([tpm_weight(r, rs) for r in R])
        out[n] = np.sum((np.asarray(mass, np.float64) * w / R ** 3)[:, None] * d, axis=0)
    return out
