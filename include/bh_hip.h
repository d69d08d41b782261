/*
 * octree_hip: MI355X-native Barnes-Hut gravity solver, C ABI.
 *
 * This is the boundary a host (the reference's ParticleSystem with a new
 * SolverMode::OctreeHip, or any FFI) binds. It replaces, for the octree_hip
 * solver, the CUDA octree path that ParticleSystem::update dispatches to:
 *   engine/physics/core/particle/PhyParticleSystem.hpp:118-123,163  (SolverMode, update)
 *   engine/physics/core/cuda/integration/CudProfilerUpdate.inl:90-125 (dispatch)
 *   engine/physics/octree/cuda/linear/OctBuild.inl:14-248            (buildLinearOctreeGpu)
 *   engine/physics/octree/cuda/force/OctGpu.inl:100-276              (GPU walk)
 *   engine/physics/core/cuda/prelude/CudParticleKernels.inl:94-220   (KDK kernels)
 * Conventions follow the reference FFI (runtime/ffi/core/FfiApi.hpp:15-26,83-100):
 * an opaque handle, a result enum, caller-owned error buffers, no exceptions
 * across the boundary. Plain pointers and sizes only; all arrays are host
 * memory owned by the caller and are copied on entry/exit. Positions and
 * velocities are interleaved xyz triples (n x 3) in the caller's particle
 * order; that order is preserved across calls whatever order the solver keeps
 * internally.
 *
 * Threading: one handle per caller thread; handles are independent (each owns
 * its device buffers and one HIP stream) and may coexist
 * (engine/physics/octree/tests/reentrancy.cpp:12-35).
 */
#ifndef BH_HIP_H_
#define BH_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same values as blitzar_core_result_t (runtime/ffi/core/FfiApi.hpp:15-21). */
typedef enum bh_hip_result {
    BH_HIP_OK = 0,
    BH_HIP_INVALID_ARGUMENT = 1,
    BH_HIP_NOT_READY = 2,
    BH_HIP_TIMEOUT = 3,
    BH_HIP_INTERNAL_ERROR = 4
} bh_hip_result_t;

enum { BH_HIP_INTEGRATOR_LEAPFROG = 0, BH_HIP_INTEGRATOR_EULER = 1 };
/* COM and BOUNDS are the reference's octree(criterion) values; EXTENT opens
 * by bmax / distance (bmax = radius about the COM enclosing the node).
 * RELATIVE (no counterpart in the reference) accepts a node when its
 * estimated truncation error is below relative_alpha times the group's
 * smallest acceleration of the previous force evaluation:
 *   quadrupole walk: M bmax^3 <= relative_alpha |a|min (d^2 + eps^2)^(5/2)
 *   monopole walk:   M bmax^2 <= relative_alpha |a|min (d^2 + eps^2)^2
 * (d: distance from the node's COM to the group's target box), and also
 * passes the COM test at theta, which then bounds the opening angle. The
 * first force evaluation after an upload primes |a| with a COM walk at
 * min(theta, 0.7). */
enum {
    BH_HIP_CRITERION_COM = 0,
    BH_HIP_CRITERION_BOUNDS = 1,
    BH_HIP_CRITERION_EXTENT = 2,
    BH_HIP_CRITERION_RELATIVE = 3
};
enum { BH_HIP_KEYS_MORTON63 = 0, BH_HIP_KEYS_REFERENCE30 = 1 };

/*
 * Solver configuration: the setters ParticleSystem receives from
 * configureSystem (engine/batch/runner/BatRunner.cpp:48-97). theta and
 * softening are resolved like resolveForceLawPolicy
 * (engine/physics/core/force/PhyForceLawPolicy.cpp:22-32):
 *   softening = max(softening, max(0, min_softening))
 *   theta     = clamp(theta, clamp(min_theta, 0.05, 4), 4)
 */
typedef struct bh_hip_params {
    float theta;             /* octree(theta) */
    float softening;         /* octree(softening): Plummer epsilon */
    float min_softening;     /* physics(min_softening) */
    float min_distance2;     /* physics(min_distance2): pairs closer are skipped */
    float min_theta;         /* physics(min_theta) */
    float max_acceleration;  /* physics(max_acceleration); <= 0 disables the clamp */
    float max_speed;         /* sph(max_speed) velocity clamp; <= 0 disables */
    int32_t criterion;       /* BH_HIP_CRITERION_* (octree(criterion)) */
    int32_t leaf_capacity;   /* octree(leaf_capacity): reference leaf depth rule */
    int32_t bucket_capacity; /* max particles per leaf after refinement (>= 1) */
    int32_t multipole_order; /* 1 = monopole, 2 = monopole + quadrupole */
    int32_t key_mode;        /* BH_HIP_KEYS_*; REFERENCE30 builds topology only */
    int32_t integrator;      /* BH_HIP_INTEGRATOR_* */
    float treepm_split_scale;/* r_s of the erfc short-range split; 0 = Newton */
    float treepm_cutoff;     /* short-range cutoff radius; 0 = none */
    int32_t waves_per_cu;    /* walk residency per CU (0 = default) */
    /* Dyadic block time steps (ParticleSystem::setAdaptiveTimeStepParameters,
     * engine/physics/octree/adaptive/OctAdaptive.cpp:17-133): 0 = off, else
     * each bh_hip_step of dt runs 2^min(max_level, 12) slices. */
    int32_t adaptive_max_level;
    float adaptive_eta;      /* level rule: eta sqrt(eps / |a|) (and eta eps / |v|) */
    int32_t adaptive_flags;  /* BH_HIP_ADAPTIVE_* */
    /* Periodic cube [0, L)^3 for the TreePM short-range walk (0 = open
     * boundaries). Needs treepm_split_scale > 0 and 0 < treepm_cutoff <= L/4;
     * every pair is taken at its nearest image and positions are wrapped
     * into the box after each drift. The reference has no periodic
     * short-range walk (its periodic mode is PM only). */
    float periodic_box;
    /* Particle-mesh long-range force of the periodic split (rocFFT): mesh
     * cells per side (0 = short range only). Needs periodic_box and the
     * short-range split; even, 8..1024. */
    int32_t pm_mesh;
    /* Group far field (no counterpart in the reference): an accepted node with
     * (bmax + 3 group radius) < far_alpha * distance from the group centre enters
     * a second-order local expansion about the group centre, evaluated once
     * per target, instead of being evaluated per target. 0 = off. Quadrupole
     * walk without TreePM only. */
    float far_alpha;
    /* Error tolerance of BH_HIP_CRITERION_RELATIVE (> 0; ignored by the other
     * criteria). */
    float relative_alpha;
    /* BH_HIP_CRITERION_RELATIVE with the quadrupole: an accepted node whose
     * quadrupole term is also small, M bmax^2 <= relative_mono *
     * relative_alpha |a|min (d^2 + eps^2)^2, is evaluated as a point mass
     * (softened monopole) instead of a second-order cell. 0 = off. */
    float relative_mono;
} bh_hip_params_t;

/* adaptive_flags:
 *   VELOCITY: also bound the step by eta * eps / |v| (the reference does);
 *   REBUILD:  build the tree at every force evaluation of a global step. By
 *             default it is built at the first one and refitted at the others
 *             (moments of the predicted positions on the same topology, cell
 *             sizes widened to cover particles that left their cells), as the
 *             reference keeps one tree per global step
 *             (CudOctreeGpuAdaptive.inl:89-113,147). */
enum { BH_HIP_ADAPTIVE_VELOCITY = 1, BH_HIP_ADAPTIVE_REBUILD = 2 };

typedef struct bh_hip_handle bh_hip_handle_t;

/* Per-phase device time summed over the force evaluations of the last call
 * (milliseconds, HIP events on the handle's stream; enable with
 * bh_hip_set_timing) and walk interaction counts. */
typedef struct bh_hip_timings {
    double build_ms;      /* AABB + keys + sort + permute + tree + moments */
    double sort_ms;       /* radix sort alone */
    double walk_ms;       /* tree walk */
    double integrate_ms;  /* kick/drift kernels */
    uint64_t pp_interactions; /* particle-particle interactions (sum over timed walks) */
    uint64_t pc_interactions; /* particle-cell interactions (sum over timed walks) */
    uint64_t node_count;      /* nodes of the last tree */
    int32_t leaf_depth;       /* reference leaf depth of the last tree */
    int32_t max_level;        /* deepest populated level of the last tree */
    int32_t walks;            /* force evaluations the sums above cover */
    int32_t groups;           /* walk groups (target sets) of the last tree */
    uint64_t particle_steps;  /* block time steps: particles corrected (sum over slices) */
    uint64_t far_interactions;  /* group far field: node-group expansions (sum over walks) */
    uint64_t far_evaluations;   /* group far field: per-target expansion evaluations */
    uint64_t mono_interactions; /* relative_mono: accepted nodes evaluated as point masses
                                   (per target; also counted in pp_interactions) */
    /* parts of build_ms */
    double keys_ms;     /* AABB + Morton keys */
    double permute_ms;  /* equal-key tie order + store permutation */
    double tree_ms;     /* topology, from the end of the sort (it runs beside the
                           permutation, on a second stream) */
    double moments_ms;  /* moments and leaf packing, after the topology, until the
                           walk groups (second stream, beside them) are done */
    uint64_t lff_pairs; /* learned correction: (target, node) pairs evaluated */
    uint64_t lff_tiles; /* learned correction: 16-pair MFMA tiles run */
    double lff_ms;      /* learned correction: the MLP pass (part of walk_ms) */
} bh_hip_timings_t;

/* Fill `out` with the defaults (reference config defaults, CfgConfig.hpp). */
void bh_hip_default_params(bh_hip_params_t* out);

/* Create a solver for n particles on the current HIP device. On failure
 * returns an error, writes a message into err (when non-NULL) and sets *out
 * to NULL. */
bh_hip_result_t bh_hip_create(const bh_hip_params_t* params, uint64_t n, bh_hip_handle_t** out,
                              char* err, size_t err_cap);
void bh_hip_destroy(bh_hip_handle_t* h);

/* The configuration a handle runs with for `in`: the resolution above
 * (resolveForceLawPolicy, engine/physics/core/force/PhyForceLawPolicy.cpp:22-32)
 * plus the solver's own clamps (adaptive level <= 12 and eta in [0.01, 1] as
 * CfgScenarioRuntime.cpp:63-68; far_alpha in [0, 1]; defaults for unset
 * fields). Host only: needs no device. */
void bh_hip_resolve_params(const bh_hip_params_t* in, bh_hip_params_t* out);

/* The resolved configuration of a handle (the values ParticleSystem's
 * getters return after configureSystem, engine/batch/runner/BatRunner.cpp:48-97). */
bh_hip_result_t bh_hip_get_params(bh_hip_handle_t* h, bh_hip_params_t* out);

/* Replace the configuration (same resolution rules as create). */
bh_hip_result_t bh_hip_set_params(bh_hip_handle_t* h, const bh_hip_params_t* params);

/* Copy state in (caller order). vel3 may be NULL (zero velocities). */
bh_hip_result_t bh_hip_upload(bh_hip_handle_t* h, const double* pos3, const double* vel3,
                              const float* mass);
/* Copy state out (caller order). Either pointer may be NULL. */
bh_hip_result_t bh_hip_download(bh_hip_handle_t* h, double* pos3, double* vel3);

/* Build the tree for the current positions and compute accelerations
 * (the force pass of ParticleSystem::update). acc3 (caller order, n x 3) may
 * be NULL to keep the result on the device only. */
bh_hip_result_t bh_hip_compute_accelerations(bh_hip_handle_t* h, float* acc3);

/* Advance nsteps steps of size dt with the configured integrator
 * (leapfrog: kick-drift-kick reusing the previous force; euler: the
 * reference CPU update v = clamp(v + a dt), x += v dt). */
bh_hip_result_t bh_hip_step(bh_hip_handle_t* h, double dt, int32_t nsteps);

/* Build the tree only (keys, sort, topology, moments) for the current
 * positions; used for topology checks. */
bh_hip_result_t bh_hip_build_tree(bh_hip_handle_t* h);

/* Number of nodes of the last built tree, reference leaf depth and the
 * deepest level; any pointer may be NULL. */
bh_hip_result_t bh_hip_tree_info(bh_hip_handle_t* h, uint64_t* node_count, int32_t* leaf_depth,
                                 int32_t* max_level);

/* Dump the last built tree. Arrays sized n (particles) or cap_nodes (nodes);
 * any pointer may be NULL.
 *   sorted_keys/sorted_ids : keys in sorted order and the caller index at each
 *                            sorted position
 *   node_prefix/level/start/count/first_child/nchild : breadth-first nodes
 *                            (first_child = -1 for leaves)
 *   node_moments           : 16 floats per node: com offset xyz, mass,
 *                            qxx qxy qxz qyy qyz (traceless, about the COM),
 *                            bmax, cell size, trs = sum m |x - com|^2,
 *                            lattice centre xyz (int32), meta (uint32)
 *   geom                   : 12 doubles: FP32 frame lo[3], inv[3]; lattice
 *                            origin[3], unit[3] */
bh_hip_result_t bh_hip_dump_tree(bh_hip_handle_t* h, uint64_t* sorted_keys, uint32_t* sorted_ids,
                                 uint64_t* node_prefix, int32_t* node_level,
                                 uint32_t* node_start, uint32_t* node_count,
                                 int32_t* node_first_child, uint32_t* node_nchild,
                                 float* node_moments, double* geom, uint64_t cap_nodes);

/* FP64 direct sum on sampled targets (caller indices), against all n
 * sources, with the configured softening and min_distance2
 * (engine/physics/treepm/model/TpmCpu.cpp:78-116). acc3 is nt x 3. */
bh_hip_result_t bh_hip_direct_forces_fp64(bh_hip_handle_t* h, const uint32_t* targets,
                                          uint32_t nt, double* acc3);

/* FP64 diagnostics (engine/physics/fmm/metrics/FmmMetrics.cpp:47-83):
 * out[0] kinetic energy, out[1] potential energy (exact when n <= sample_limit,
 * otherwise estimated from sample_limit evenly spaced particles), out[2..4]
 * linear momentum, out[5..7] angular momentum. */
bh_hip_result_t bh_hip_diagnostics(bh_hip_handle_t* h, uint32_t sample_limit, double* out8);

/* Timings: enable (1) or disable (0) per-phase event timing; read the last
 * values. */
bh_hip_result_t bh_hip_set_timing(bh_hip_handle_t* h, int32_t enabled);
bh_hip_result_t bh_hip_get_timings(bh_hip_handle_t* h, bh_hip_timings_t* out);

/* Last error message of the handle (FfiApi.hpp:83-100 style). */
bh_hip_result_t bh_hip_get_last_error(bh_hip_handle_t* h, char* buf, size_t cap);

/* ---- Learned far-field correction --------------------------------------
 * No counterpart in the reference. A small MLP (32 -> 64 -> 64 -> 3, SiLU,
 * BF16 MFMA inference with FP32 accumulation) predicts, for an accepted node
 * of at least min_count particles close to a target (size / distance >=
 * band), the residual of the second-order expansion in units of M / r^2 from
 * the node's moments up to third order and the target-node geometry; the
 * corrected walk runs at a larger opening angle. Training tuples come from this solver's own walks
 * (exact residual = direct sum over the node's particles) and the network is
 * trained on the GPU by this library's forward/backward kernels (Adam). */
typedef struct bh_hip_lff_config {
    float theta;       /* opening angle of the corrected walk (e.g. 0.9) */
    float band;        /* near nodes: size / distance >= band */
    int32_t list_cap;  /* nodes recorded per target group (0 = default 192) */
    int32_t enabled;   /* 1: force evaluations and steps use the corrected walk */
    int32_t min_count; /* near nodes with fewer particles are opened (summed at
                          the angle `band`) instead of corrected (default 32) */
} bh_hip_lff_config_t;

#define BH_HIP_LFF_INPUTS 32
#define BH_HIP_LFF_PARAMS 6467 /* W1[32][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3] */

bh_hip_result_t bh_hip_lff_configure(bh_hip_handle_t* h, const bh_hip_lff_config_t* cfg);
/* Build the tree at the current positions, walk at cfg->theta recording the
 * near nodes of every group_stride-th group, and append their (features,
 * normalised residual) tuples to the device training set (holding at most
 * max_tuples; allocated on the first call). *total receives the tuple count. */
bh_hip_result_t bh_hip_lff_collect(bh_hip_handle_t* h, int32_t group_stride, uint64_t max_tuples,
                                   uint64_t* total);
/* Train on the device set: epochs x (n / batch) Adam steps, learning rate lr
 * decayed by a cosine to lr / 10; loss_per_epoch (may be NULL) receives the
 * mean squared error per tuple in units of the target RMS. */
bh_hip_result_t bh_hip_lff_train(bh_hip_handle_t* h, int32_t epochs, int32_t batch, float lr,
                                 uint64_t seed, float* loss_per_epoch);
bh_hip_result_t bh_hip_lff_get_params(bh_hip_handle_t* h, float* params);
bh_hip_result_t bh_hip_lff_set_params(bh_hip_handle_t* h, const float* params);
/* Gradient of the mean over n samples of |MLP(x) - y|^2 with the given
 * parameters (FP32 training kernels); X is n x 32, Y is n x 3. */
bh_hip_result_t bh_hip_lff_gradient(bh_hip_handle_t* h, const float* X, const float* Y, uint32_t n,
                                    const float* params, float* grad, float* loss);
/* MLP outputs for n feature rows through the BF16 MFMA inference kernel. */
bh_hip_result_t bh_hip_lff_forward(bh_hip_handle_t* h, const float* X, uint32_t n, float* Y);
/* Copy up to cap training tuples to the host; *n receives the count copied.
 * With X and Y both NULL, *n receives the number of tuples held. */
bh_hip_result_t bh_hip_lff_get_tuples(bh_hip_handle_t* h, float* X, float* Y, uint64_t cap,
                                      uint64_t* n);

/* Long-range (mesh) part of the periodic TreePM force at the current
 * positions, caller order, n x 3 (needs pm_mesh > 0). No counterpart with the
 * split in the reference (its periodic PM runs with r_s = 0,
 * engine/physics/treepm/cuda/cosmology/TpmCosmology.inl:52). */
bh_hip_result_t bh_hip_pm_accelerations(bh_hip_handle_t* h, float* acc3);

/* Per-particle block time-step level (caller order) of the adaptive
 * schedule: the particle steps with dt / 2^level. Zero before the first
 * adaptive step. */
bh_hip_result_t bh_hip_adaptive_levels(bh_hip_handle_t* h, int32_t* levels);

/* Standalone stable radix sort of (key, value) pairs on the device, the
 * building block behind thrust::sort_by_key at OctBuild.inl:96. Sorts by the
 * low key_bits bits (1..64); host arrays in and out. */
bh_hip_result_t bh_hip_sort_pairs(const uint64_t* keys, const uint32_t* vals, uint64_t n,
                                  int32_t key_bits, uint64_t* keys_out, uint32_t* vals_out,
                                  char* err, size_t err_cap);

/* Allocation-free variant of bh_hip_sort_pairs. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * keys / vals / keys_out / vals_out are DEVICE pointers of the current device;
 * each output array is either its input array (in place) or disjoint from
 * every other array. stream is a hipStream_t (NULL: the null stream); the
 * call waits for it before returning. Same ordering as bh_hip_sort_pairs
 * (stable, by the low key_bits bits). Scratch lives in a fixed device array
 * (about 24.6 MiB per GPU, held until the process ends): n above 2,097,152
 * pairs is rejected with BH_HIP_INVALID_ARGUMENT. Calls that overlap on one
 * GPU share that array and race. */
bh_hip_result_t bh_hip_sort_pairs_noalloc(const uint64_t* keys, const uint32_t* vals, uint64_t n,
                                          int32_t key_bits, uint64_t* keys_out,
                                          uint32_t* vals_out, void* stream, char* err,
                                          size_t err_cap);

/* Library version string. */
const char* bh_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* BH_HIP_H_ */
