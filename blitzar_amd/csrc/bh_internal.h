// Internal declarations shared by the octree_hip HIP translation units.
// Not part of the C ABI (include/bh_hip.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <cstddef>
#include <cstdint>

namespace bh {

constexpr int kMaxLevel = 21;          // 21 bits per axis, 63-bit keys
constexpr int kLevels = kMaxLevel + 1;  // levels 0..21
constexpr uint32_t kQMax = 2097151u;    // 2^21 - 1
constexpr uint32_t kGroupMax = 64;      // targets per walk group (one wavefront)
constexpr int kInfoTotal = 2 * kLevels;      // level_info slot: total nodes
constexpr int kInfoMaxLevel = 2 * kLevels + 1;  // deepest populated level
constexpr int kInfoGroups = 2 * kLevels + 2;    // number of walk groups
constexpr int kInfoBig = 2 * kLevels + 3;       // large nodes of the current level (tree build)
// AABB: per-block partial extrema (6 doubles per block) in Device::diag.
constexpr int kAabbBlocks = 1024;

// Walk tickets: one counter per XCD (workgroup i runs on XCD i mod 8), each
// on its own 256-byte line.
constexpr int kTicketQueues = 8;
constexpr int kTicketStride = 64;  // uint32 words between counters
constexpr int kTicketWords = kTicketQueues * kTicketStride;

// Learned far-field MLP: 32 inputs (29 features + 3 pads) -> 64 -> 64 -> 3.
constexpr int kLffInputs = 32;
constexpr int kLffHidden = 64;
constexpr int kLffOutputs = 3;
constexpr int kLffParams = kLffInputs * kLffHidden + kLffHidden + kLffHidden * kLffHidden +
                           kLffHidden + kLffHidden * kLffOutputs + kLffOutputs;  // 6467

// Root geometry, written on the device once per build by geometry_kernel.
//   FP32 part: the quantisation frame of the key kernel (AABB of the
//   float-cast positions and its per-axis reciprocal extent).
//   FP64 part: lattice origin and unit. A cell (level L, index i along an
//   axis) has its centre at lattice coordinate c = (2i+1)*2^(21-L) - 1, i.e.
//   at position lo + c * unit. Lattice coordinates are integers in
//   [-1, 2^22 - 1), so differences of cell centres are exact in int32 and in
//   FP32, and only the root geometry needs FP64.
struct RootGeom {
    float lo[3];
    float inv[3];
    float unit_f[3];   // FP32 copy of unit, used by the walk
    float width_f[3];  // finest cell width = 2 * unit (FP32)
    double lo_d[3];
    double unit_d[3];
};

// One tree node as the walk reads it: 64 bytes, 16 words.
//   com/mass          : centre of mass as an FP32 offset from the node's
//                       lattice centre, total mass
//   q                 : traceless quadrupole about the COM (xx, xy, xz, yy, yz;
//                       zz = -xx - yy)
//   bmax              : radius about the COM enclosing every particle
//   trs               : trace of the second moment about the COM,
//                       sum m |x - com|^2; with Plummer softening eps the
//                       second-order expansion carries -5/2 eps^2 trs d / r^7
//   first             : first child node (internal) or first sorted particle (leaf)
// The cell size used by the opening criterion follows from the level
// (cell_size below).
//   cx, cy, cz        : lattice centre
//   meta              : nchild (bits 0-3) | level (bits 4-8) | count (bits 9-31,
//                       particles in the subtree, saturated at 2^23 - 1)
struct alignas(16) WalkNode {
    float comx, comy, comz, mass;
    float qxx, qxy, qxz, qyy;
    float qyz, bmax, trs;
    int first;
    int cx, cy, cz;
    uint32_t meta;
};
static_assert(sizeof(WalkNode) == 64, "WalkNode must be 64 bytes");

// Position relative to a reference lattice centre: the lattice part
// dc * unit is rounded on its own before the offset is added (no fused
// multiply-add), so every kernel forms the same bits for the same particle and
// a target meets itself as a source at exactly zero separation.
__host__ __device__ inline float rel_pos(int dc, float unit, float off)
{
#pragma clang fp contract(off)
    return static_cast<float>(dc) * unit + off;
}

// Largest cell width at `level`: wmax = max finest cell width over the axes.
__host__ __device__ inline float cell_size(float wmax, int level)
{
    return wmax * ldexpf(1.0f, kMaxLevel - level);
}

__host__ __device__ inline uint32_t meta_nchild(uint32_t m) { return m & 15u; }
__host__ __device__ inline uint32_t meta_level(uint32_t m) { return (m >> 4) & 31u; }
__host__ __device__ inline uint32_t meta_count(uint32_t m) { return m >> 9; }
__host__ __device__ inline uint32_t make_meta(uint32_t nchild, uint32_t level, uint32_t count)
{
    return nchild | (level << 4) | (count << 9);
}
constexpr uint32_t kMaxLeafCount = (1u << 23) - 1u;

// Error bits raised on the device; the host reports them after a sync.
enum DeviceError : int {
    kErrNodeOverflow = 1,
    kErrStackOverflow = 2,
    kErrLeafTooLarge = 4,
    kErrKeyMode = 8,
    kErrScan = 16,  // a scan block gave up waiting for its predecessors
    kErrTopology = 32,  // a tree link the build could not resolve
};

// One-pass tree topology (k_tree.hip): tiles of kTopoTile sorted keys; it
// handles buckets of up to kTopoHalo particles (wider ones build level by
// level). Its per-tile, per-level node counts need topo_count_words(n) words.
constexpr int kTopoTile = 4096;
constexpr int kTopoHalo = 64;
inline uint32_t topo_tiles(size_t n) { return static_cast<uint32_t>((n + kTopoTile - 1) / kTopoTile); }
inline size_t topo_count_words(size_t n) { return static_cast<size_t>(kLevels) * (topo_tiles(n) + 1); }

// Parameters the walk needs (by value into the kernel).
struct WalkParams {
    float theta2;
    float eps2;
    float min_d2;
    float max_acc;      // <= 0: no clamp
    int criterion;      // 0 = centre of mass, 1 = bounds, 2 = extent, 3 = relative
    int order;          // 1 = monopole, 2 = quadrupole
    float rs;           // TreePM split scale (0 = Newton)
    float rcut2;        // TreePM cutoff squared (0 = none)
    float box = 0.0f;   // periodic cube side for the TreePM walk (0 = open)
    float far2 = 0.0f;  // group far field: (bmax + group radius)^2 < far2 * d^2 (0 = off)
    float ralpha = 0.0f;  // relative criterion tolerance
    float rmono = 0.0f;   // relative criterion: point-mass tolerance as a fraction of ralpha
    // |a| per particle (store order): read by the relative criterion, written
    // by every walk for the targets it evaluates (nullptr: neither)
    float* amag = nullptr;
    // Block time steps: when alevel is set, only groups holding a particle
    // active at tick `atick` are walked (active: atick mod 2^(alevels -
    // level) == 0); the other groups keep their previous acc.
    const uint32_t* alevel = nullptr;
    uint32_t atick = 0;
    int alevels = 0;
};

// Block time-step settings of one adaptive slice.
struct AdaptiveParams {
    int levels;       // L = min(max_level, 12)
    double eta;
    double soft;      // max(softening, min_softening)
    double dt;        // global step
    double quantum;   // dt / 2^L
    int velocity;     // also bound by eta * soft / |v|
    double max_speed;
    double box;       // periodic cube side (0 = open): positions wrapped after updates
};

// Device buffers of one solver instance. The particle store is kept in the
// order of the last sort (Morton order); `id` maps a store slot to the
// caller's particle index and `rank` maps back.
struct Device {
    size_t n = 0;
    size_t node_cap = 0;
    hipStream_t stream = nullptr;

    double* pos[3] = {nullptr, nullptr, nullptr};
    double* vel[3] = {nullptr, nullptr, nullptr};
    float* mass = nullptr;
    uint32_t* id = nullptr;
    double* pos_b[3] = {nullptr, nullptr, nullptr};  // permutation targets
    double* vel_b[3] = {nullptr, nullptr, nullptr};
    float* mass_b = nullptr;
    uint32_t* id_b = nullptr;
    uint32_t* rank = nullptr;  // caller index -> store slot
    float* acc[3] = {nullptr, nullptr, nullptr};
    float* amag = nullptr;    // |a| of the last force evaluation (relative criterion)
    float* amag_b = nullptr;
    bool amag_live = false;   // amag is in use: the store permutation carries it
    // > 0: the last drift left this many per-block AABB partials of the
    // current positions in diag (the next build skips its AABB pass); every
    // other writer of positions or diag clears it
    int aabb_parts = 0;

    uint64_t* keys = nullptr;
    uint64_t* keys_b = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* vals_b = nullptr;
    uint32_t* sort_hist = nullptr;  // per-tile digit counts
    size_t sort_hist_cap = 0;

    float4* spart = nullptr;   // sorted particle: offset from leaf lattice centre, mass
    uint32_t* sleaf = nullptr; // sorted particle -> leaf node

    // tree (breadth-first; level L occupies [level_base[L], level_base[L]+level_count[L]))
    uint64_t* nprefix = nullptr;
    uint32_t* nstart = nullptr;
    uint32_t* ncount = nullptr;
    uint32_t* nchild = nullptr;   // child count (0 = leaf)
    uint32_t* nfirst = nullptr;   // first child index
    uint8_t* nlevel = nullptr;
    uint32_t* nparent = nullptr;  // parent index (root: 0xffffffff)
    WalkNode* wnode = nullptr;
    float* next = nullptr;            // per node: particle extent about the lattice centre
    float* noct = nullptr;            // per node: 10 third moments about the COM (learned
                                      // correction only; null otherwise)
    uint32_t* gscan = nullptr;        // per-node group counts -> offsets
    uint32_t* gscan_blocks = nullptr;
    uint4* groups = nullptr;          // walk groups: first particle, count, node

    uint32_t* scan_tmp = nullptr;  // per-level scan scratch
    uint32_t* bnd_tmp = nullptr;   // per-level child bounds (8 per node) from count to emit
    uint32_t* big_list = nullptr;  // per-level indices of nodes searched by a whole wavefront
    uint32_t* topo_cnt = nullptr;  // one-pass topology: node counts per tile and level
    uint32_t* scan_blocks = nullptr;
    int* level_info = nullptr;  // [0..22) base, [22..44) count, [44] total, [45] max level, [46] groups

    RootGeom* geom = nullptr;
    unsigned long long* aabb_bits = nullptr;  // 6 ordered-encoded doubles
    int* err = nullptr;
    unsigned int* work_counter = nullptr;
    int* stack = nullptr;  // per-resident-wave traversal stacks
    size_t stack_waves = 0;
    size_t stack_cap = 0;

    // block time steps (allocated on the first adaptive step; permuted with
    // the store): acceleration of the last force evaluation, level, tick of
    // the last force evaluation
    float* aacc[3] = {nullptr, nullptr, nullptr};
    uint32_t* alevel = nullptr;
    uint32_t* atick = nullptr;
    float* aacc_b[3] = {nullptr, nullptr, nullptr};
    uint32_t* alevel_b = nullptr;
    uint32_t* atick_b = nullptr;

    double* stage = nullptr;  // 3n doubles host<->device staging
    double* diag = nullptr;   // reduction scratch
    size_t diag_cap = 0;
};

// Learned far-field state of one solver instance (device pointers).
struct LffState {
    uint32_t* lists = nullptr;    // per group: recorded node ids, list_cap each
    uint32_t* list_n = nullptr;   // per group: number of recorded nodes
    int list_cap = 0;
    size_t list_groups = 0;       // groups the lists can hold
    float* params = nullptr;      // kLffParams FP32
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;        // kLffParams
    float* partial = nullptr;     // partial_blocks x kLffParams
    float* loss_partial = nullptr;
    float* loss_acc = nullptr;
    int partial_blocks = 0;
    void* frag = nullptr;         // 14 x 64 BF16x8 MFMA fragments
    float* bias = nullptr;        // b1[64], b2[64], b3[16]
    float* X = nullptr;           // training features, tuple_cap x kLffInputs
    float* Y = nullptr;           // training targets, tuple_cap x 3
    unsigned int* tuple_counter = nullptr;  // [0] tuples held, [1] scratch count
    size_t tuple_cap = 0;
    uint32_t collects = 0;                  // collect calls so far (hash seed)
};

// Particle-mesh state of one solver instance (k_pm.hip).
struct PmState {
    int N = 0;
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
    float* rho = nullptr;                    // N^3 density
    float* rho_k = nullptr;                  // N^2 (N/2+1) complex
    float* tmp_k = nullptr;                  // N^2 (N/2+1) complex
    float* field[3] = {nullptr, nullptr, nullptr};  // N^3 acceleration components (x N^3)
};

// ---- launchers (defined in the kernel translation units) -------------------
// k_keys.hip
void launch_aabb(Device& d);
void launch_keys(Device& d, int key_mode);
// k_sort.hip: stable LSD radix sort of (keys, vals) pairs, 8-bit digits over
// bits [0, key_bits). Result lands back in d.keys / d.vals.
void launch_radix_sort(Device& d, size_t n, int key_bits);
size_t radix_hist_words(size_t n);
// The same sort on explicit buffers: (keys, vals) in, (keys_alt, vals_alt) the
// second buffer pair, hist_ws radix_hist_words(n) words whose scan part
// starts zeroed once. Returns true when the result ends in the alternate
// buffers (an odd number of passes). Nothing is allocated.
int radix_sort_passes(int key_bits);
bool radix_sort_buffers(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                        uint32_t* hist_ws, size_t n, int key_bits, int* err, hipStream_t stream);
// The same sort reading (keys_in, vals_in), which it never writes: pass 0
// writes buffer pair a, pass 1 pair b, and so on alternately. Returns true
// when the result ends in pair a (an odd number of passes), false when it
// ends in pair b. Nothing is allocated.
bool radix_sort_from(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_a,
                     uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b, uint32_t* hist_ws,
                     size_t n, int key_bits, int* err, hipStream_t stream);
// Exclusive scan of `data` in place (one launch). The element count is *n_dev
// when n_dev is non-null (then max_n bounds it), else n_host. The total is
// written to total_out / total_out_i when non-null. `block_sums` is the scan's
// workspace of scan_block_sums_words(max_n) words, zeroed once when allocated.
void launch_scan_exclusive(uint32_t* data, size_t max_n, const int* n_dev, size_t n_host,
                           uint32_t* block_sums, uint32_t* total_out, int* total_out_i, int* err,
                           hipStream_t s);
size_t scan_block_sums_words(size_t max_n);
void launch_tie_fix(Device& d);
// sorted order for the store: positions and masses (what the moments read),
// then velocities, ids, |a| and the block time-step arrays, each on its stream
void launch_permute_store(Device& d, hipStream_t s);
void launch_permute_rest(Device& d, hipStream_t s);
void launch_rank(Device& d);  // rebuild rank from id
// k_tree.hip
void launch_tree_build(Device& d, int leaf_depth, int bucket, hipStream_t s);
// Moments, leaf packing and the walk's node records of the current
// positions. refit: the topology is that of an earlier build (positions have
// moved since); each node's level in the walk record is lowered until its
// cell covers the node's particles.
void launch_tree_moments(Device& d, bool refit = false);
// Walk groups: a partition of the sorted particles into spatially compact
// runs of at most gmax <= kGroupMax particles (consecutive small children of
// one node packed together; see node_groups in k_tree.hip).
void launch_groups(Device& d, uint32_t gmax, hipStream_t s);
// k_walk.hip
// stats (optional, device): [0] particle-particle, [1] particle-cell interactions,
// [3] node-group far-field interactions, [4] far-field target evaluations
void launch_walk(Device& d, const WalkParams& p, unsigned long long* stats);
// Walk of the learned correction. A node that passes the opening test and is
// near the group, size^2 >= band^2 * (distance from its COM to the group
// box)^2, is opened when it holds fewer than min_count particles (so small
// near nodes are summed at the finer angle `band`), else accepted and
// recorded in L.lists for the MLP pass.
void launch_walk_record(Device& d, const WalkParams& p, unsigned long long* stats,
                        const LffState& L, float band, int min_count);
// k_lff.hip
// Training tuples of every group_stride-th group; a pair is kept when its
// hash is <= keep. count_only: only count the kept pairs into *counter.
void launch_lff_tuples(Device& d, const LffState& L, int group_stride, float band, float eps2,
                       int min_count, uint32_t keep, uint32_t seed, bool count_only,
                       unsigned int* counter, unsigned int max_tuples);
void launch_lff_grad(Device& d, const LffState& L, const float* X, const float* Y,
                     const uint32_t* order, size_t first, uint32_t nbatch, float yscale,
                     const float* P, int nblocks);
void launch_lff_adam(Device& d, const LffState& L, int nblocks, float* P, float* grad, float lr,
                     int step);
void launch_lff_pack(Device& d, const LffState& L);
void launch_lff_forward(Device& d, const LffState& L, const float* X, uint32_t n, float* Y);
void launch_lff_apply(Device& d, const LffState& L, float band, float eps2, int min_count,
                      float max_acc, unsigned long long* stats);
// k_integrate.hip
void launch_kick(Device& d, double dt_kick, double max_speed);
// box > 0: positions wrapped into [0, box) after the drift
void launch_kick_drift(Device& d, double dt_kick, double dt_drift, double max_speed,
                       double box = 0.0);
void launch_wrap(Device& d, double box);
void launch_upload_scatter(Device& d, int which);   // 0 = pos, 1 = vel
void launch_download_gather(Device& d, int which);  // 0 = pos, 1 = vel, 2 = acc(float->double)
void launch_init_ids(Device& d);
void launch_direct_fp64(Device& d, const uint32_t* targets_dev, uint32_t nt, double eps,
                        double min_d2, double* out_dev);
void launch_energy(Device& d, double* out_dev);
void launch_potential(Device& d, const uint32_t* slots_dev, uint32_t ns, double eps,
                      double min_d2, double* out_dev);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// k_pm.hip: plans and meshes for an N^3 mesh (host allocation); the force
// pass adds the long-range acceleration to acc and applies the clamp.
bool pm_setup(PmState& pm, int N, hipStream_t stream);
void pm_release(PmState& pm);
bool launch_pm_accel(Device& d, PmState& pm, double L, float rs, float max_acc);
// block time steps
void launch_adaptive_init(Device& d, const AdaptiveParams& a, uint32_t tick);
// Predictor over `span` (k slices of constant acceleration: x += v s + a s^2/2,
// v += a s, the same as k single-slice predictors).
void launch_adaptive_drift(Device& d, const AdaptiveParams& a, double span);
// *out = the first tick after t at which some particle is due.
void launch_adaptive_next_due(Device& d, const AdaptiveParams& a, uint32_t t, uint32_t* out);
// count (optional, device): += number of particles corrected
void launch_adaptive_correct(Device& d, const AdaptiveParams& a, uint32_t tick,
                             unsigned long long* count);
void launch_gather_u32(Device& d, const uint32_t* src_by_slot, uint32_t* out_by_caller);

}  // namespace bh
