"""Writes tests/golden/octree_cpu_ref.npz: input/output vectors of the
reference's own CPU octree (oracle/_ref/libref.so, compiled from the reference
sources by oracle/Makefile), used to pin the oracle's octree_cpu restatement on
machines where the reference sources are absent.

Run from the repository root (needs oracle/_ref/libref.so):
    python tests/golden/make_golden.py
Cases:
  criterion9   the 9-body scene of TST_UNT_PHYS_019
               (engine/physics/octree/tests/octree_criteria.cpp:46-62),
               theta 0.9, eps 0.01, both opening criteria
  two_body     tests/data/two_body_rest.xyz (copied as two_body_rest.xyz)
  cloud2k_*    2000-particle anisotropic Gaussian cloud (numpy seed 20261202),
               theta 0.5 and 0.7, eps 0.01, both criteria
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402


def criterion9():
    pos = np.array([[0, 0, 0], [6, -4, -4], [6, 4, -4], [6, -4, 4], [6, 4, 4], [14, -4, -4],
                    [14, 4, -4], [14, -4, 4], [14, 4, 4]], np.float32)
    m = np.array([1.0] + [0.5 + 0.1 * i for i in range(8)], np.float32)
    return O.hot(pos, m)


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref.so is missing: run `make -C oracle` here first")
    out = {}
    x9 = criterion9()
    out["criterion9_xyzm"] = x9
    for crit in (0, 1):
        p = O.make_params(theta=0.9, softening=0.01, min_softening=1e-4, min_distance2=1e-12,
                          min_theta=0.05, criterion=crit)
        out[f"criterion9_acc_c{crit}"] = O.ref_octree_forces(x9, p)
    pos2 = np.array([[-1, 0, 0], [1, 0, 0]], np.float32)
    x2 = O.hot(pos2, np.ones(2, np.float32))
    out["two_body_xyzm"] = x2
    out["two_body_acc"] = O.ref_octree_forces(x2, O.make_params(theta=0.5, softening=0.01))
    rng = np.random.default_rng(20261202)
    pos = (rng.standard_normal((2000, 3)) * np.array([1.0, 0.6, 0.3])).astype(np.float32)
    mass = (0.5 + rng.random(2000)).astype(np.float32) / 2000.0
    xc = O.hot(pos, mass)
    out["cloud2k_xyzm"] = xc
    for th in (0.5, 0.7):
        for crit in (0, 1):
            p = O.make_params(theta=th, softening=0.01, criterion=crit)
            out[f"cloud2k_acc_t{int(th * 10)}_c{crit}"] = O.ref_octree_forces(xc, p)
    np.savez_compressed(os.path.join(HERE, "octree_cpu_ref.npz"), **out)
    print("wrote", os.path.join(HERE, "octree_cpu_ref.npz"), sorted(out))


if __name__ == "__main__":
    main()
