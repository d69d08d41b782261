/*
 * Deterministic initial-condition generators of octree_hip (C ABI, host code).
 *
 * The reference generates scenes through buildGeneratedState /
 * buildPlummerModel (engine/server/simulation/generation/SrvPlummer.cpp:17-106,
 * Gaussian velocities, std::mt19937). It ships no generator for the
 * distribution-function Plummer sphere, Hernquist halos or Zel'dovich
 * lattices, so these are this library's own. Every particle draws from its
 * own counter-based random stream (seed, stream, particle index), so the
 * output is independent of the thread count. Units: G = 1.
 * All arrays are caller-owned host memory: pos3/vel3 are n x 3, mass is n.
 */
#ifndef BH_IC_H_
#define BH_IC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Plummer sphere sampled from its distribution function (Aarseth, Henon &
 * Wielen 1974): total mass M, scale radius a, equal masses M/n; centred on
 * the centre of mass and the mean velocity. */
void bh_ic_plummer(uint64_t n, uint64_t seed, double total_mass, double scale_radius,
                   double* pos3, double* vel3, float* mass);

/* Hernquist (1990) halo with isotropic velocities drawn from its
 * distribution function by rejection; radii truncated at r_max (> 0).
 * `stream` separates halos generated with the same seed. Centred. */
void bh_ic_hernquist(uint64_t n, uint64_t seed, uint64_t stream, double total_mass,
                     double scale_radius, double r_max, double* pos3, double* vel3,
                     float* mass);

/* Two equal Hernquist halos (mass 1, scale radius 1, r_max 200 each) of
 * n_per_halo particles, centres `separation` apart along x on a parabolic
 * orbit with pericentre distance `pericentre`, in the centre-of-mass frame.
 * Arrays hold 2 * n_per_halo particles (halo 0 first). */
void bh_ic_hernquist_merger(uint64_t n_per_halo, uint64_t seed, double separation,
                            double pericentre, double* pos3, double* vel3, float* mass);

/* Uniform random cube [-h, h]^3, zero velocities, equal masses M/n. */
void bh_ic_uniform_cube(uint64_t n, uint64_t seed, double half_extent, double total_mass,
                        double* pos3, double* vel3, float* mass);

#ifdef __cplusplus
}
#endif

#endif /* BH_IC_H_ */
