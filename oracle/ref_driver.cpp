// TEST INFRASTRUCTURE ONLY. C entry points over the reference's own CPU octree
// (compiled from /root/reference sources by oracle/Makefile into oracle/_ref/).
// Nothing here restates the algorithm: it only marshals arrays into the
// reference's ParticleHotData / ForceLawPolicy types and calls
//   Octree::build(const std::vector<ParticleHotData>&)   OctreeBuild.cpp:54
//   Octree::computeForceOn(hot, i, policy, criterion)     OctreeTraversal.cpp:24
//   resolveForceLawPolicy(...)                            PhyForceLawPolicy.cpp:22
#include "physics/core/force/PhyForceLawPolicy.hpp"
#include "physics/core/particle/PhyParticleHotData.hpp"
#include "physics/octree/model/Octree.hpp"

#include <vector>

extern "C" {
// Unclamped octree accelerations on every particle (serial, deterministic).
int ref_octree_forces(int n, const float* xyzm, float theta, float softening,
                      float min_softening, float min_distance2, float min_theta, int criterion,
                      float cutoff2, float* acc3)
{
    std::vector<ParticleHotData> hot(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        hot[static_cast<size_t>(i)] =
            ParticleHotData(xyzm[4 * i], xyzm[4 * i + 1], xyzm[4 * i + 2], xyzm[4 * i + 3]);
    const ForceLawPolicy policy =
        resolveForceLawPolicy(theta, softening, min_softening, min_distance2, min_theta);
    Octree tree;
    tree.build(hot);
    const OctreeOpeningCriterion crit =
        criterion == 1 ? OctreeOpeningCriterion::Bounds : OctreeOpeningCriterion::CenterOfMass;
    for (int i = 0; i < n; ++i) {
        const Vector3 a = tree.computeForceOn(hot, static_cast<size_t>(i), policy, crit, cutoff2);
        acc3[3 * i] = a.x;
        acc3[3 * i + 1] = a.y;
        acc3[3 * i + 2] = a.z;
    }
    return static_cast<int>(tree.getNodeCount());
}
}

extern "C" {
// Octree accelerations on `nt` listed targets, OpenMP-parallel over targets
// (the reference's OctForces.cpp:120-126 with deterministic=false). Returns the
// seconds spent in Octree::build and in the walk through *t_build / *t_walk.
int ref_octree_forces_subset(int n, const float* xyzm, float theta, float softening,
                             float min_softening, float min_distance2, float min_theta,
                             int criterion, int nt, const int* targets, float* acc3,
                             int nthreads, double* t_build, double* t_walk);
}

#include <omp.h>

int ref_octree_forces_subset(int n, const float* xyzm, float theta, float softening,
                             float min_softening, float min_distance2, float min_theta,
                             int criterion, int nt, const int* targets, float* acc3,
                             int nthreads, double* t_build, double* t_walk)
{
    std::vector<ParticleHotData> hot(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        hot[static_cast<size_t>(i)] =
            ParticleHotData(xyzm[4 * i], xyzm[4 * i + 1], xyzm[4 * i + 2], xyzm[4 * i + 3]);
    const ForceLawPolicy policy =
        resolveForceLawPolicy(theta, softening, min_softening, min_distance2, min_theta);
    const OctreeOpeningCriterion crit =
        criterion == 1 ? OctreeOpeningCriterion::Bounds : OctreeOpeningCriterion::CenterOfMass;
    Octree tree;
    const double t0 = omp_get_wtime();
    tree.build(hot);
    const double t1 = omp_get_wtime();
    const int threads = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int k = 0; k < nt; ++k) {
        const Vector3 a =
            tree.computeForceOn(hot, static_cast<size_t>(targets[k]), policy, crit, 0.0f);
        acc3[3 * k] = a.x;
        acc3[3 * k + 1] = a.y;
        acc3[3 * k + 2] = a.z;
    }
    const double t2 = omp_get_wtime();
    if (t_build)
        *t_build = t1 - t0;
    if (t_walk)
        *t_walk = t2 - t1;
    return static_cast<int>(tree.getNodeCount());
}
