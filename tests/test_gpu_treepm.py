"""GPU parity of the TreePM short-range walk (erfc split, cutoff at 4.5 r_s,
theta 0.6) against the FP64 erfc-weighted direct sum and against the oracle's
octree_cpu with the same short-range weight and cutoff
(engine/physics/octree/force/OctreeForce.cpp:49-57,
engine/physics/octree/traversal/OctreeTraversal.cpp:156-207)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.mark.parametrize("theta", [0.3, 0.6])
def test_short_range_walk(oracle, theta):
    import blitzar_amd as B
    n = 40000
    print(f"case treepm n={n} theta={theta}", flush=True)
    rng = np.random.default_rng(12)
    pos = rng.random((n, 3))
    m = np.full(n, 1.0 / n, np.float32)
    cell = 1.0 / 64.0
    rs = 1.25 * cell
    rc = 4.5 * rs
    eps = 0.1 * cell
    with B.OctreeHip(n, theta=theta, softening=eps, treepm_split_scale=rs, treepm_cutoff=rc) as s:
        s.upload(pos, None, m)
        a = s.compute_accelerations()
    t = np.arange(0, n, 37)
    ref = oracle.direct_tpm_fp64(pos, m.astype(np.float64), t, float(np.float32(eps)),
                                 float(np.float32(rs)), float(np.float32(rc)))
    r = rel(a[t], ref)
    print(f"  median={np.median(r):.3e} p99={np.percentile(r, 99):.3e}", flush=True)
    assert np.median(r) < (2e-4 if theta < 0.5 else 1e-3)
    assert np.percentile(r, 99) < 1e-2
    x = oracle.hot(pos, m)
    acpu = oracle.octree_forces_subset(x, t, oracle.make_params(
        theta=theta, softening=eps, short_range_scale=rs, cutoff2=rc * rc))
    assert np.median(rel(a[t], acpu)) < 2e-3
