// C ABI of octree_hip (include/bh_hip.h): handle lifecycle, buffer
// management and the per-step orchestration of the HIP kernels.
//
// Step structure (leapfrog), after the first force evaluation:
//   kick_drift(dt/2, dt) -> AABB -> keys -> radix sort -> tie fix -> permute
//   -> tree levels -> moments -> walk -> kick(dt/2)
// All of it is queued on the handle's stream without host synchronisation;
// the host synchronises once per API call to report device-side errors.
#include "../../include/bh_hip.h"
#include "bh_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace {

constexpr int kDefaultWavesPerCu = 24;
constexpr float kDefaultRelativeAlpha = 0.002f;
constexpr size_t kStackCap = 16384;

// b0 build start, s0 sort start (after AABB + keys), s1 sort end, p1 after
// the tie fix and permutation, t1 after the tree levels, b1 build end (after
// moments and groups), w1 walk end.
struct StepEvents {
    hipEvent_t b0 = nullptr, s0 = nullptr, s1 = nullptr, p1 = nullptr, t1 = nullptr,
               j1 = nullptr, b1 = nullptr, w1 = nullptr;
    hipEvent_t l0 = nullptr;  // learned correction start (after its walk)
    bool l0_used = false;
};

// Start and end of one kick / kick-drift launch (read at the end of the call,
// so timing adds no host wait inside the step loop).
struct IntegrationEvents {
    hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace

struct bh_hip_handle {
    bh::Device d;
    bh_hip_params_t p{};
    int leaf_depth = 1;
    int cu_count = 256;
    bool primed = false;      // acc holds a(x) of the current positions
    bool amag_valid = false;  // amag holds |a| of a force evaluation of the current state
    bh::PmState pm;           // particle mesh of the periodic split (pm_mesh > 0)
    bool pm_failed = false;   // a rocFFT execution failed in the last call
    bool ad_primed = false;   // block time-step schedule initialised
    double ad_quantum = 0.0;  // slice length of the schedule
    uint64_t ad_tick = 0;     // slices since the schedule started
    bool have_tree = false;   // the last build matches the current positions
    bool rank_valid = true;   // rank (caller index -> slot) matches the store order
    bool timing = false;
    std::string last_error;
    // A failed device allocation after creation leaves the handle unusable:
    // every later call returns an error with this message.
    std::string broken;
    // Error word of the device copied (pinned) after each build, and the event
    // that marks the copy; read while the walk queued behind the build runs.
    int* err_host = nullptr;
    hipEvent_t built = nullptr;
    // The topology and the walk groups of a build run on a second stream
    // (default priority), beside the store permutation and the moments:
    // sorted -> topology (side), topo -> moments (main), groups -> walk (main).
    hipStream_t side = nullptr;
    hipEvent_t ev_sorted = nullptr, ev_topo = nullptr, ev_groups = nullptr;
    int err_carry = 0;      // error bits of this call kept across a node regrow
    int steps_done = -1;    // bh_hip_step: steps completed before an error stopped it
    // Test switch (environment BH_HIP_TEST_FAIL_NODE_GROWTH at creation): a
    // node regrow asks for an impossible size, so the failed-allocation path
    // runs for real.
    bool test_fail_growth = false;
    // device counters: pp, pc, block-step particle steps, far node-group,
    // far evaluations, point-mass cells, learned-correction pairs and tiles
    unsigned long long* stats = nullptr;
    std::vector<StepEvents> events;
    size_t events_used = 0;
    std::vector<IntegrationEvents> int_events;
    size_t int_used = 0;
    bh_hip_timings_t t{};
    std::vector<void*> allocations;
    // learned far-field correction
    bh::LffState lff;
    bh_hip_lff_config_t lff_cfg{0.9f, 0.5f, 192, 0, 32};
    float* lff_train_params = nullptr;  // normalised-output copy used while training
    uint32_t* lff_order = nullptr;
    size_t lff_order_cap = 0;
};

namespace {

void set_err(char* err, size_t cap, const std::string& msg)
{
    if (err && cap > 0) {
        std::snprintf(err, cap, "%s", msg.c_str());
    }
}

bool check(bh_hip_handle* h, hipError_t e, const char* what)
{
    if (e == hipSuccess)
        return true;
    h->last_error = std::string("[hip] ") + what + " failed: " + hipGetErrorString(e);
    return false;
}

template <typename T>
bool dalloc(bh_hip_handle* h, T** p, size_t count, const char* what)
{
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (!check(h, hipMalloc(&q, bytes), what)) {
        // HIP keeps the failure as the thread's last error until it is read;
        // read it here so that the next launch check of this or another
        // handle does not report it again
        (void)hipGetLastError();
        return false;
    }
    h->allocations.push_back(q);
    *p = static_cast<T*>(q);
    return true;
}

// dalloc, then zero the buffer in stream order (for workspaces whose state
// must start at zero, such as the single-pass scans' status words)
template <typename T>
bool dalloc_zero(bh_hip_handle* h, T** p, size_t count, const char* what, hipStream_t s)
{
    return dalloc(h, p, count, what) &&
           check(h, hipMemsetAsync(*p, 0, std::max<size_t>(count, 1) * sizeof(T), s), what);
}

void dfree(bh_hip_handle* h, void* p)
{
    if (!p)
        return;
    auto it = std::find(h->allocations.begin(), h->allocations.end(), p);
    if (it != h->allocations.end())
        h->allocations.erase(it);
    (void)hipFree(p);
}

// PhyForceLawPolicy.cpp:22-32 and the octree setters (OctConfiguration.cpp:26-39).
bh_hip_params_t resolve(const bh_hip_params_t& in)
{
    bh_hip_params_t p = in;
    const float min_soft = std::max(0.0f, in.min_softening);
    const float min_theta = std::clamp(in.min_theta, 0.05f, 4.0f);
    p.min_softening = min_soft;
    p.min_distance2 = std::max(0.0f, in.min_distance2);
    p.min_theta = min_theta;
    p.softening = std::max(in.softening, min_soft);
    p.theta = std::clamp(in.theta, min_theta, 4.0f);
    p.bucket_capacity = std::max(1, in.bucket_capacity);
    p.multipole_order = in.multipole_order >= 2 ? 2 : 1;
    p.criterion = in.criterion >= 1 && in.criterion <= 3 ? in.criterion : 0;
    p.relative_alpha = in.relative_alpha > 0.0f ? in.relative_alpha : kDefaultRelativeAlpha;
    p.relative_mono = std::max(0.0f, in.relative_mono);
    p.key_mode = in.key_mode == 1 ? 1 : 0;
    p.integrator = in.integrator == 1 ? 1 : 0;
    // CfgScenarioRuntime.cpp:63-68: max level <= 12, eta in [0.01, 1]
    p.adaptive_max_level = std::clamp(in.adaptive_max_level, 0, 12);
    p.adaptive_eta = std::clamp(in.adaptive_eta, 0.01f, 1.0f);
    p.adaptive_flags = in.adaptive_flags & (BH_HIP_ADAPTIVE_VELOCITY | BH_HIP_ADAPTIVE_REBUILD);
    p.periodic_box = std::max(0.0f, in.periodic_box);
    p.pm_mesh = std::max(0, in.pm_mesh);
    p.far_alpha = std::clamp(in.far_alpha, 0.0f, 1.0f);
    if (p.waves_per_cu <= 0)
        p.waves_per_cu = kDefaultWavesPerCu;
    return p;
}

// Parameter combinations the solver does not run; nullptr when valid.
const char* invalid_combination(const bh_hip_params_t& p)
{
    if (p.periodic_box > 0.0f) {
        if (!(p.treepm_split_scale > 0.0f) || !(p.treepm_cutoff > 0.0f))
            return "periodic_box needs the TreePM short-range walk (split scale and cutoff > 0)";
        if (p.treepm_cutoff * 4.0f > p.periodic_box)
            return "periodic_box needs treepm_cutoff <= periodic_box / 4 (nearest image)";
    }
    if (p.pm_mesh > 0) {
        if (!(p.periodic_box > 0.0f))
            return "pm_mesh needs periodic_box > 0 and the short-range split";
        if (p.pm_mesh < 8 || p.pm_mesh > 1024 || (p.pm_mesh & 1))
            return "pm_mesh must be even and within [8, 1024]";
    }
    return nullptr;
}

// OctBuild.inl:28-39
int reference_leaf_depth(uint64_t n, int leaf_capacity)
{
    const int cap = std::max(16, leaf_capacity);
    int d = 1;
    while (d < 21) {
        const double avg = static_cast<double>(n) / static_cast<double>(1ull << (3 * d));
        if (avg <= static_cast<double>(cap))
            break;
        ++d;
    }
    return d;
}

bool alloc_nodes(bh_hip_handle* h, size_t cap)
{
    bh::Device& d = h->d;
    dfree(h, d.nprefix);
    dfree(h, d.nstart);
    dfree(h, d.ncount);
    dfree(h, d.nchild);
    dfree(h, d.nfirst);
    dfree(h, d.nlevel);
    dfree(h, d.wnode);
    dfree(h, d.next);
    dfree(h, d.scan_tmp);
    dfree(h, d.bnd_tmp);
    dfree(h, d.big_list);
    dfree(h, d.topo_cnt);
    dfree(h, d.scan_blocks);
    dfree(h, d.nparent);
    dfree(h, d.gscan);
    dfree(h, d.gscan_blocks);
    const bool oct = d.noct != nullptr;
    dfree(h, d.noct);
    d.noct = nullptr;
    d.nprefix = nullptr;
    d.nstart = d.ncount = d.nchild = d.nfirst = d.scan_tmp = d.bnd_tmp = d.big_list = nullptr;
    d.topo_cnt = nullptr;
    d.scan_blocks = d.nparent = d.gscan = d.gscan_blocks = nullptr;
    d.nlevel = nullptr;
    d.wnode = nullptr;
    d.next = nullptr;
    d.node_cap = cap;
    const size_t scan_n = std::min(d.n, cap);
    return dalloc(h, &d.nprefix, cap, "node prefix") && dalloc(h, &d.nstart, cap, "node start") &&
           dalloc(h, &d.ncount, cap, "node count") && dalloc(h, &d.nchild, cap, "node nchild") &&
           dalloc(h, &d.nfirst, cap, "node first") && dalloc(h, &d.nlevel, cap, "node level") &&
           dalloc(h, &d.wnode, cap, "walk nodes") && dalloc(h, &d.next, cap, "node extent") &&
           dalloc(h, &d.scan_tmp, scan_n, "scan tmp") &&
           dalloc(h, &d.bnd_tmp, 8 * scan_n, "child bounds") &&
           dalloc(h, &d.big_list, scan_n, "large-node list") &&
           dalloc(h, &d.topo_cnt, bh::topo_count_words(d.n), "tile node counts") &&
           dalloc_zero(h, &d.scan_blocks, bh::scan_block_sums_words(scan_n), "scan blocks",
                       d.stream) &&
           dalloc(h, &d.nparent, cap, "node parent") && dalloc(h, &d.gscan, cap, "group scan") &&
           dalloc_zero(h, &d.gscan_blocks, bh::scan_block_sums_words(cap), "group scan blocks",
                       d.stream) &&
           (!oct || dalloc(h, &d.noct, 10 * cap, "node third moments"));
}

bool alloc_all(bh_hip_handle* h, uint64_t n)
{
    bh::Device& d = h->d;
    d.n = n;
    for (int a = 0; a < 3; ++a) {
        if (!dalloc(h, &d.pos[a], n, "pos") || !dalloc(h, &d.vel[a], n, "vel") ||
            !dalloc(h, &d.pos_b[a], n, "pos_b") || !dalloc(h, &d.vel_b[a], n, "vel_b") ||
            !dalloc(h, &d.acc[a], n, "acc"))
            return false;
    }
    if (!dalloc(h, &d.amag, n, "amag") || !dalloc(h, &d.amag_b, n, "amag_b"))
        return false;
    if (!dalloc(h, &d.mass, n, "mass") || !dalloc(h, &d.mass_b, n, "mass_b") ||
        !dalloc(h, &d.id, n, "id") || !dalloc(h, &d.id_b, n, "id_b") ||
        !dalloc(h, &d.rank, n, "rank") || !dalloc(h, &d.keys, n, "keys") ||
        !dalloc(h, &d.keys_b, n, "keys_b") || !dalloc(h, &d.vals, n, "vals") ||
        !dalloc(h, &d.vals_b, n, "vals_b") || !dalloc(h, &d.spart, n, "spart") ||
        !dalloc(h, &d.sleaf, n, "sleaf") || !dalloc(h, &d.stage, 3 * n, "stage") ||
        !dalloc(h, &d.groups, n, "groups"))
        return false;
    d.sort_hist_cap = bh::radix_hist_words(n);
    if (!dalloc_zero(h, &d.sort_hist, d.sort_hist_cap, "sort hist", d.stream))
        return false;
    if (!dalloc(h, &d.level_info, 2 * bh::kLevels + 8, "level info") ||
        !dalloc(h, &d.geom, 1, "geometry") || !dalloc(h, &d.aabb_bits, 8, "aabb") ||
        !dalloc(h, &d.err, 4, "error flags") || !dalloc(h, &d.work_counter, bh::kTicketWords, "work counter") ||
        !dalloc(h, &h->stats, 8, "walk stats"))
        return false;
    d.diag_cap = 6 * 1024 + 8 * 256 + 16;  // AABB partials, energy partials
    if (!dalloc(h, &d.diag, d.diag_cap, "diagnostics"))
        return false;
    if (!alloc_nodes(h, 2 * n + 4096))
        return false;
    // resident walk waves: one per group of the smallest cap (16, see
    // group_max) up to the residency limit
    const size_t groups = (n + 15) / 16;
    d.stack_waves = std::max<size_t>(
        1, std::min(groups, static_cast<size_t>(h->cu_count) * static_cast<size_t>(h->p.waves_per_cu)));
    d.stack_cap = std::min(kStackCap, d.node_cap + 64);
    return dalloc(h, &d.stack, d.stack_waves * d.stack_cap, "walk stacks");
}

bool zero_err(bh_hip_handle* h)
{
    h->err_carry = 0;
    bh::launch_fill_u32(reinterpret_cast<uint32_t*>(h->d.err), 0u, 1, h->d.stream);
    return check(h, hipGetLastError(), "error flag reset");
}

// Device errors that no retry repairs (a node-capacity overflow is repaired
// inside the force evaluation, see check_build).
constexpr int kFatalBits =
    bh::kErrStackOverflow | bh::kErrLeafTooLarge | bh::kErrScan | bh::kErrTopology;

// Synchronise and turn device error bits into a message.
bh_hip_result_t finish(bh_hip_handle* h, const char* what)
{
    if (!check(h, hipGetLastError(), what) || !check(h, hipStreamSynchronize(h->d.stream), what))
        return BH_HIP_INTERNAL_ERROR;
    if (h->pm_failed) {
        h->pm_failed = false;
        h->last_error = std::string("[octree_hip] ") + what + ": rocFFT execution failed";
        return BH_HIP_INTERNAL_ERROR;
    }
    int flags = 0;
    if (!check(h, hipMemcpy(&flags, h->d.err, sizeof(int), hipMemcpyDeviceToHost), "error read"))
        return BH_HIP_INTERNAL_ERROR;
    flags |= h->err_carry;
    h->err_carry = 0;
    if (flags != 0) {
        std::string m = std::string("[octree_hip] ") + what + ":";
        if (flags & bh::kErrNodeOverflow)
            m += " node capacity exceeded (tree truncated);";
        if (flags & bh::kErrStackOverflow)
            m += " traversal stack overflow;";
        if (flags & bh::kErrLeafTooLarge)
            m += " leaf larger than 2^23 particles;";
        if (flags & bh::kErrScan)
            m += " device scan did not complete (results of this call invalid);";
        if (flags & bh::kErrTopology)
            m += " tree links inconsistent (results of this call invalid);";
        if (h->steps_done >= 0)
            m += " stopped after " + std::to_string(h->steps_done) +
                 " completed steps (the state holds them and part of the next)";
        h->last_error = m;
        return BH_HIP_INTERNAL_ERROR;
    }
    return BH_HIP_OK;
}

bool usable(bh_hip_handle* h)
{
    if (h->broken.empty())
        return true;
    h->last_error = h->broken;
    return false;
}

// Double the node capacity. A failed allocation leaves the handle unusable
// (alloc_nodes has released the node arrays and nulled their pointers).
bool grow_nodes(bh_hip_handle* h)
{
    const size_t cap = h->d.node_cap;
    h->have_tree = false;
    const size_t want = h->test_fail_growth ? (static_cast<size_t>(1) << 46) : cap * 2;
    if (alloc_nodes(h, want))
        return true;
    h->d.node_cap = 0;
    h->broken = "[octree_hip] handle unusable: growing the node arrays to " +
                std::to_string(want) + " nodes failed (" + h->last_error + ")";
    h->last_error = h->broken;
    return false;
}

enum class Built { kOk, kRetry, kFatal };

// Called once the work that uses a build (the walk) has been queued behind
// it: waits for the build alone (the walk keeps the GPU busy meanwhile) and
// reads the error word the build left. On a node-capacity overflow the walk
// ran on a truncated tree: the call waits for it, keeps the error bits raised
// before this build, grows the capacity and asks the caller to queue the
// build and its walk again (positions are unchanged, so the repeat is the
// evaluation a larger capacity would have given).
Built check_build(bh_hip_handle* h)
{
    if (!check(h, hipEventSynchronize(h->built), "build"))
        return Built::kFatal;
    const int bits = *h->err_host;
    if (!(bits & bh::kErrNodeOverflow))
        return Built::kOk;
    if (!check(h, hipStreamSynchronize(h->d.stream), "build"))
        return Built::kFatal;
    h->err_carry |= bits & ~bh::kErrNodeOverflow;
    bh::launch_fill_u32(reinterpret_cast<uint32_t*>(h->d.err), 0u, 1, h->d.stream);
    if (!grow_nodes(h))
        return Built::kFatal;
    h->amag_valid = false;  // the truncated walk wrote |a|
    return Built::kRetry;
}

// Bits of the device error word at the last build (and before it).
int bits_at_build(const bh_hip_handle* h) { return *h->err_host | h->err_carry; }

StepEvents* next_events(bh_hip_handle* h)
{
    if (!h->timing)
        return nullptr;
    if (h->events_used == h->events.size()) {
        StepEvents e;
        for (hipEvent_t* x : {&e.b0, &e.s0, &e.s1, &e.p1, &e.t1, &e.j1, &e.b1, &e.w1, &e.l0})
            if (hipEventCreate(x) != hipSuccess)
                return nullptr;
        h->events.push_back(e);
    }
    StepEvents* ev = &h->events[h->events_used++];
    ev->l0_used = false;
    return ev;
}

IntegrationEvents* next_integration_events(bh_hip_handle* h)
{
    if (!h->timing)
        return nullptr;
    if (h->int_used == h->int_events.size()) {
        IntegrationEvents e;
        if (hipEventCreate(&e.a) != hipSuccess)
            return nullptr;
        if (hipEventCreate(&e.b) != hipSuccess) {
            (void)hipEventDestroy(e.a);
            return nullptr;
        }
        h->int_events.push_back(e);
    }
    return &h->int_events[h->int_used++];
}

void rec(hipEvent_t e, hipStream_t s)
{
    if (e)
        (void)hipEventRecord(e, s);
}

// Walk-group size. A walk keeps about cu_count * 24 waves resident and
// hands each group to one wave, so with fewer groups than that it lasts as
// long as its heaviest group. Small workloads therefore cut smaller groups:
// the cap halves from 64 while n < 2 * cap * cu_count * 24, down to 16, and
// the lists of a small group split over 2 or 4 lanes per target. The cap
// depends on n and the GPU only, never on the requested residency, so
// schedules do not change results. It holds for the COM and relative
// criteria without the group far field. The bounds and extent criteria keep
// 64: with groups of 16 the bounds walk on the reference's 2,000-particle
// cloud (theta 0.5) erred more than the reference's own per-target walk,
// which groups of 64 do not. The far field expands about the group centre
// and is built for groups of 64.
uint32_t group_max(const bh_hip_handle* h)
{
    if (h->p.criterion == BH_HIP_CRITERION_BOUNDS || h->p.criterion == BH_HIP_CRITERION_EXTENT ||
        h->p.far_alpha > 0.0f)
        return bh::kGroupMax;
    const size_t waves = static_cast<size_t>(h->cu_count) * static_cast<size_t>(kDefaultWavesPerCu);
    uint32_t g = bh::kGroupMax;
    while (g > 16 && h->d.n < 2 * static_cast<size_t>(g) * waves)
        g /= 2;
    return g;
}

// Queue one tree build for the current positions.
void queue_build(bh_hip_handle* h, StepEvents* ev)
{
    bh::Device& d = h->d;
    rec(ev ? ev->b0 : nullptr, d.stream);
    bh::launch_aabb(d);
    bh::launch_keys(d, h->p.key_mode);
    rec(ev ? ev->s0 : nullptr, d.stream);
    bh::launch_radix_sort(d, d.n, h->p.key_mode == 1 ? 30 : 63);
    rec(ev ? ev->s1 : nullptr, d.stream);
    bh::launch_tie_fix(d);
    // The topology needs only the sorted keys: it runs on the side stream
    // while the main stream permutes the positions and masses. The rest of
    // the store (velocities, ids, |a|) and the walk groups follow the
    // topology there, beside the moments; the walk waits for both streams.
    (void)hipEventRecord(h->ev_sorted, d.stream);
    (void)hipStreamWaitEvent(h->side, h->ev_sorted, 0);
    bh::launch_tree_build(d, h->leaf_depth, h->p.bucket_capacity, h->side);
    rec(ev ? ev->t1 : nullptr, h->side);
    (void)hipEventRecord(h->ev_topo, h->side);
    bh::launch_permute_rest(d, h->side);
    if (h->p.key_mode == 0)
        bh::launch_groups(d, group_max(h), h->side);
    (void)hipEventRecord(h->ev_groups, h->side);
    bh::launch_permute_store(d, d.stream);
    rec(ev ? ev->p1 : nullptr, d.stream);
    h->rank_valid = false;
    (void)hipStreamWaitEvent(d.stream, h->ev_topo, 0);
    rec(ev ? ev->j1 : nullptr, d.stream);
    if (h->p.key_mode == 0)
        bh::launch_tree_moments(d);
    (void)hipStreamWaitEvent(d.stream, h->ev_groups, 0);
    rec(ev ? ev->b1 : nullptr, d.stream);
    (void)hipMemcpyAsync(h->err_host, d.err, sizeof(int), hipMemcpyDeviceToHost, d.stream);
    (void)hipEventRecord(h->built, d.stream);
    h->have_tree = true;
}

// Moments of the current positions on the topology of the last build (block
// time-step slices between rebuilds): no sort, no permutation, no new nodes.
void queue_refit(bh_hip_handle* h, StepEvents* ev)
{
    bh::Device& d = h->d;
    for (hipEvent_t e : {ev ? ev->b0 : nullptr, ev ? ev->s0 : nullptr, ev ? ev->s1 : nullptr,
                         ev ? ev->p1 : nullptr, ev ? ev->t1 : nullptr, ev ? ev->j1 : nullptr})
        rec(e, d.stream);
    bh::launch_tree_moments(d, true);
    rec(ev ? ev->b1 : nullptr, d.stream);
    (void)hipMemcpyAsync(h->err_host, d.err, sizeof(int), hipMemcpyDeviceToHost, d.stream);
    (void)hipEventRecord(h->built, d.stream);
}

// A build with nothing queued behind it (topology dumps), regrown and
// repeated on a node-capacity overflow.
bool build_checked(bh_hip_handle* h, StepEvents* ev)
{
    for (int attempt = 0; attempt < 8; ++attempt) {
        queue_build(h, ev);
        const Built b = check_build(h);
        if (b == Built::kOk)
            return true;
        if (b == Built::kFatal)
            return false;
    }
    h->last_error = "[octree_hip] build: node capacity still exceeded after 8 doublings";
    return false;
}

void ensure_rank(bh_hip_handle* h)
{
    if (!h->rank_valid) {
        bh::launch_rank(h->d);
        h->rank_valid = true;
    }
}



bh::WalkParams walk_params(const bh_hip_params_t& p)
{
    bh::WalkParams w{};
    w.theta2 = p.theta * p.theta;
    w.eps2 = p.softening * p.softening;
    w.min_d2 = p.min_distance2;
    w.max_acc = p.max_acceleration;
    w.criterion = p.criterion;
    w.order = p.multipole_order;
    w.rs = p.treepm_split_scale;
    w.rcut2 = p.treepm_cutoff > 0.0f ? p.treepm_cutoff * p.treepm_cutoff : 0.0f;
    w.box = p.periodic_box;
    w.far2 = p.far_alpha > 0.0f ? p.far_alpha * p.far_alpha : 0.0f;
    w.ralpha = p.relative_alpha;
    w.rmono = p.relative_mono;
    return w;
}

// A force walk. Under the relative criterion every walk records |a| per
// target (the store permutation carries it), and the criterion reads it; the
// first such walk on a new state is preceded by a COM walk at min(theta, 0.7)
// over all groups. Other criteria neither read nor keep |a|.
void launch_force_walk(bh_hip_handle* h, bh::WalkParams wp)
{
    if (wp.criterion != BH_HIP_CRITERION_RELATIVE) {
        h->amag_valid = false;
        h->d.amag_live = false;
        bh::launch_walk(h->d, wp, h->stats);
        return;
    }
    wp.amag = h->d.amag;
    if (!h->amag_valid) {
        bh::WalkParams w0 = wp;
        w0.criterion = BH_HIP_CRITERION_COM;
        w0.theta2 = std::min(wp.theta2, 0.49f);
        w0.alevel = nullptr;
        bh::launch_walk(h->d, w0, h->stats);
    }
    bh::launch_walk(h->d, wp, h->stats);
    h->amag_valid = true;
    h->d.amag_live = true;
}

bool lff_ready(bh_hip_handle* h) { return h->lff_cfg.enabled && h->lff.lists && h->lff.frag; }

// The corrected walk exists for the quadrupole, non-periodic, 63-bit-key path.
const char* lff_incompatible(const bh_hip_params_t& p)
{
    if (p.multipole_order != 2)
        return "learned far-field correction needs multipole_order = 2";
    if (p.treepm_split_scale > 0.0f)
        return "learned far-field correction is not available with the TreePM short-range walk";
    if (p.key_mode != 0)
        return "learned far-field correction needs the 63-bit keys";
    if (p.criterion == BH_HIP_CRITERION_RELATIVE)
        return "learned far-field correction runs the geometric criteria (com, bounds, extent), "
               "not the relative criterion";
    return nullptr;
}

// The walk(s) of one force evaluation, queued behind a build.
void queue_walks(bh_hip_handle* h, StepEvents* ev)
{
    if (lff_ready(h)) {
        bh::WalkParams wp = walk_params(h->p);
        wp.theta2 = h->lff_cfg.theta * h->lff_cfg.theta;
        wp.max_acc = 0.0f;  // clamped after the correction
        bh::launch_walk_record(h->d, wp, h->stats, h->lff, h->lff_cfg.band, h->lff_cfg.min_count);
        if (ev) {
            rec(ev->l0, h->d.stream);
            ev->l0_used = true;
        }
        bh::launch_lff_apply(h->d, h->lff, h->lff_cfg.band, wp.eps2, h->lff_cfg.min_count,
                             h->p.max_acceleration, h->stats);
    } else if (h->pm.N > 0) {
        bh::WalkParams wp = walk_params(h->p);
        wp.max_acc = 0.0f;  // clamped after the mesh force
        launch_force_walk(h, wp);
        if (!bh::launch_pm_accel(h->d, h->pm, h->p.periodic_box, h->p.treepm_split_scale,
                                 h->p.max_acceleration))
            h->pm_failed = true;
    } else {
        launch_force_walk(h, walk_params(h->p));
    }
}

// Queue build + walk (a force evaluation); with the learned correction
// enabled the walk runs at the correction's opening angle, records the near
// nodes of each group and the MLP pass adds its corrections. Returns false
// on an error that stops the call (message in last_error).
bool queue_force(bh_hip_handle* h)
{
    StepEvents* ev = next_events(h);
    for (int attempt = 0;; ++attempt) {
        queue_build(h, ev);
        queue_walks(h, ev);
        const Built b = check_build(h);
        if (b == Built::kOk)
            break;
        if (b == Built::kFatal)
            return false;
        if (attempt == 7) {
            h->last_error = "[octree_hip] node capacity still exceeded after 8 doublings";
            return false;
        }
    }
    rec(ev ? ev->w1 : nullptr, h->d.stream);
    h->primed = true;
    return true;
}


void begin_call(bh_hip_handle* h)
{
    h->events_used = 0;
    h->int_used = 0;
    bh::launch_fill_u32(reinterpret_cast<uint32_t*>(h->stats), 0u, 16, h->d.stream);
}

// After a synchronised call: fold event times and counters into h->t.
void end_call(bh_hip_handle* h)
{
    bh_hip_timings_t t{};
    if (h->timing) {
        for (size_t k = 0; k < h->events_used; ++k) {
            const StepEvents& e = h->events[k];
            auto span = [](hipEvent_t x, hipEvent_t y, double& acc) {
                float v = 0.0f;
                if (hipEventElapsedTime(&v, x, y) == hipSuccess)
                    acc += v;
            };
            span(e.b0, e.b1, t.build_ms);
            span(e.b0, e.s0, t.keys_ms);
            span(e.s0, e.s1, t.sort_ms);
            span(e.s1, e.p1, t.permute_ms);
            span(e.s1, e.t1, t.tree_ms);
            span(e.j1, e.b1, t.moments_ms);
            span(e.b1, e.w1, t.walk_ms);
            if (e.l0_used)
                span(e.l0, e.w1, t.lff_ms);
        }
        for (size_t k = 0; k < h->int_used; ++k) {
            float a = 0.0f;
            if (hipEventElapsedTime(&a, h->int_events[k].a, h->int_events[k].b) == hipSuccess)
                t.integrate_ms += a;
        }
    }
    unsigned long long st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpy(st, h->stats, sizeof(st), hipMemcpyDeviceToHost) == hipSuccess) {
        t.pp_interactions = st[0];
        t.pc_interactions = st[1];
        t.particle_steps = st[2];
        t.far_interactions = st[3];
        t.far_evaluations = st[4];
        t.mono_interactions = st[5];
        t.lff_pairs = st[6];
        t.lff_tiles = st[7];
    }
    int info[2 * bh::kLevels + 4];
    if (hipMemcpy(info, h->d.level_info, sizeof(info), hipMemcpyDeviceToHost) == hipSuccess) {
        t.node_count = static_cast<uint64_t>(info[bh::kInfoTotal]);
        t.max_level = info[bh::kInfoMaxLevel];
        t.groups = info[bh::kInfoGroups];
    }
    t.leaf_depth = h->leaf_depth;
    t.walks = static_cast<int32_t>(h->events_used);
    h->t = t;
}

// Block time steps (engine/physics/core/cuda/integration/CudOctreeGpuAdaptive.inl:
// 80-117 for the schedule, CudAdaptive.inl for the per-particle rules): each
// global step of dt runs 2^L slices of q = dt / 2^L; every slice drifts all
// particles with their last force, rebuilds the tree at the predicted
// positions, walks the groups that hold a particle due a force, and corrects
// those particles. (The reference GPU path keeps the tree of the start of the
// global step for all slices and corrects a particle while other threads may
// read it; here the tree is rebuilt per slice and the correction runs after
// the walk.)
bh::AdaptiveParams adaptive_params(const bh_hip_handle* h, double dt)
{
    bh::AdaptiveParams a;
    a.levels = std::min(h->p.adaptive_max_level, 12);
    a.eta = h->p.adaptive_eta;
    a.soft = std::max(h->p.softening, h->p.min_softening);
    a.dt = dt;
    a.quantum = dt / static_cast<double>(1u << a.levels);
    a.velocity = (h->p.adaptive_flags & BH_HIP_ADAPTIVE_VELOCITY) != 0;
    a.max_speed = h->p.max_speed;
    a.box = h->p.periodic_box;
    return a;
}

bool alloc_adaptive(bh_hip_handle* h)
{
    bh::Device& d = h->d;
    if (d.alevel)
        return true;
    for (int a = 0; a < 3; ++a)
        if (!dalloc(h, &d.aacc[a], d.n, "adaptive acc") ||
            !dalloc(h, &d.aacc_b[a], d.n, "adaptive acc"))
            return false;
    uint32_t *lv = nullptr, *lt = nullptr, *lvb = nullptr, *ltb = nullptr;
    if (!dalloc(h, &lv, d.n, "adaptive level") || !dalloc(h, &lt, d.n, "adaptive tick") ||
        !dalloc(h, &lvb, d.n, "adaptive level") || !dalloc(h, &ltb, d.n, "adaptive tick"))
        return false;
    bh::launch_fill_u32(lv, 0u, d.n, d.stream);
    bh::launch_fill_u32(lt, 0u, d.n, d.stream);
    d.atick = lt;
    d.alevel_b = lvb;
    d.atick_b = ltb;
    d.alevel = lv;  // last: its presence switches the permutation on
    return true;
}

// Build (or refit) + walk of one block-step force evaluation, regrown and
// repeated on a node-capacity overflow (a refit makes no nodes).
bool build_and_walk(bh_hip_handle* h, StepEvents* ev, const bh::WalkParams& wp,
                    bool refit = false)
{
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (refit)
            queue_refit(h, ev);
        else
            queue_build(h, ev);
        // A build permutes the block time-step levels with the store (the
        // buffer pointers swap): the walk must read the permuted ones.
        bh::WalkParams w = wp;
        if (w.alevel)
            w.alevel = h->d.alevel;
        launch_force_walk(h, w);
        const Built b = check_build(h);
        if (b == Built::kOk) {
            rec(ev ? ev->w1 : nullptr, h->d.stream);
            return true;
        }
        if (b == Built::kFatal)
            return false;
    }
    h->last_error = "[octree_hip] node capacity still exceeded after 8 doublings";
    return false;
}

bh_hip_result_t step_adaptive(bh_hip_handle* h, double dt, int32_t nsteps)
{
    if (lff_ready(h)) {
        h->last_error = "step: block time steps do not combine with the learned correction";
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (!alloc_adaptive(h))
        return BH_HIP_INTERNAL_ERROR;
    begin_call(h);
    bh::Device& d = h->d;
    const bh::AdaptiveParams a = adaptive_params(h, dt);
    const bh::WalkParams wp0 = walk_params(h->p);
    auto fail = [&]() {
        (void)hipStreamSynchronize(d.stream);
        h->ad_primed = false;
        h->primed = false;
        end_call(h);
        return BH_HIP_INTERNAL_ERROR;
    };
    if (!h->ad_primed || std::fabs(h->ad_quantum - a.quantum) > 1e-12 * a.quantum) {
        // ticks start at 0 so that every cadence 2^(L - level) divides the
        // start of each global step
        h->ad_tick = 0;
        if (!build_and_walk(h, next_events(h), wp0))
            return fail();
        bh::launch_adaptive_init(d, a, 0u);
        h->ad_primed = true;
        h->ad_quantum = a.quantum;
    }
    // Event-driven slices: from tick t the next tick at which any particle is
    // due is found on the device; the slices in between have no force
    // evaluation, so one predictor spans them (constant acceleration makes k
    // single-slice predictors one predictor over k q).
    uint32_t* due = reinterpret_cast<uint32_t*>(d.diag);
    const uint64_t slices = 1ull << a.levels;
    // The tree is built at the first force evaluation of each global step
    // and refitted (moments of the predicted positions on the same topology)
    // at the later ones, as the reference keeps one tree per global step
    // (CudOctreeGpuAdaptive.inl:89-113,147); BH_HIP_ADAPTIVE_REBUILD builds
    // at every one.
    const bool rebuild_each = (h->p.adaptive_flags & BH_HIP_ADAPTIVE_REBUILD) != 0;
    h->steps_done = 0;
    bool stop = false;
    for (int k = 0; k < nsteps && !stop; ++k) {
        const uint64_t end = h->ad_tick + slices;
        bool first_slice = true;
        while (h->ad_tick < end) {
            const uint32_t t0 = static_cast<uint32_t>(h->ad_tick);
            bh::launch_adaptive_next_due(d, a, t0, due);
            uint32_t next = 0;
            if (!check(h, hipMemcpyAsync(&next, due, sizeof(next), hipMemcpyDeviceToHost,
                                         d.stream),
                       "adaptive schedule") ||
                !check(h, hipStreamSynchronize(d.stream), "adaptive schedule"))
                return fail();
            const uint64_t t = std::min<uint64_t>(end, h->ad_tick + (next - t0));
            bh::launch_adaptive_drift(d, a, static_cast<double>(t - h->ad_tick) * a.quantum);
            h->ad_tick = t;
            const uint32_t tt = static_cast<uint32_t>(t);
            bh::WalkParams wp = wp0;
            wp.alevel = d.alevel;
            wp.atick = tt;
            wp.alevels = a.levels;
            if (!build_and_walk(h, next_events(h), wp, !first_slice && !rebuild_each))
                return fail();
            first_slice = false;
            bh::launch_adaptive_correct(d, a, tt, h->stats + 2);
            if (bits_at_build(h) & kFatalBits) {
                stop = true;
                break;
            }
        }
        if (stop)
            break;
        // Every particle was corrected at the last tick of the global step
        // (each cadence divides 2^L), so every last-force tick equals `end`:
        // ticks restart at 0 and stay below 2^L + 1 however long the run.
        bh::launch_fill_u32(d.atick, 0u, d.n, d.stream);
        h->ad_tick = 0;
        h->steps_done = k + 1;
    }
    h->primed = false;
    const bh_hip_result_t r = finish(h, "step (block time steps)");
    h->steps_done = -1;
    if (r != BH_HIP_OK)
        h->ad_primed = false;
    end_call(h);
    return r;
}

bool stage_to_device(bh_hip_handle* h, const double* src3)
{
    return check(h,
                 hipMemcpyAsync(h->d.stage, src3, 3 * h->d.n * sizeof(double),
                                hipMemcpyHostToDevice, h->d.stream),
                 "stage upload");
}

}  // namespace

extern "C" {

void bh_hip_default_params(bh_hip_params_t* out)
{
    if (!out)
        return;
    std::memset(out, 0, sizeof(*out));
    out->theta = 0.5f;
    out->softening = 0.01f;
    out->min_softening = 1.0e-4f;   // FndConstants.hpp:22
    out->min_distance2 = 1.0e-12f;  // FndConstants.hpp:23
    out->min_theta = 0.05f;         // FndConstants.hpp:24
    out->max_acceleration = 0.0f;
    out->max_speed = 0.0f;
    out->criterion = BH_HIP_CRITERION_COM;
    out->leaf_capacity = 256;  // CfgConfig.hpp:63
    out->bucket_capacity = 16;
    out->multipole_order = 2;
    out->key_mode = BH_HIP_KEYS_MORTON63;
    out->integrator = BH_HIP_INTEGRATOR_LEAPFROG;
    out->waves_per_cu = kDefaultWavesPerCu;
    out->adaptive_max_level = 0;  // off (SrvSimulationServer.hpp:412-414: max level 4 when on)
    out->adaptive_eta = 0.25f;
    out->adaptive_flags = BH_HIP_ADAPTIVE_VELOCITY;
    out->relative_alpha = kDefaultRelativeAlpha;
}

const char* bh_hip_version(void) { return "octree_hip 0.2 (gfx950)"; }

void bh_hip_resolve_params(const bh_hip_params_t* in, bh_hip_params_t* out)
{
    if (in && out)
        *out = resolve(*in);
}

bh_hip_result_t bh_hip_get_params(bh_hip_handle_t* h, bh_hip_params_t* out)
{
    if (!h || !out)
        return BH_HIP_INVALID_ARGUMENT;
    *out = h->p;
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_create(const bh_hip_params_t* params, uint64_t n, bh_hip_handle_t** out,
                              char* err, size_t err_cap)
{
    if (out)
        *out = nullptr;
    if (!params || !out || n == 0 || n > 0x7fffffffull) {
        set_err(err, err_cap, "bh_hip_create: invalid argument (params/out null or n out of range)");
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (const char* why = invalid_combination(resolve(*params))) {
        set_err(err, err_cap, (std::string("bh_hip_create: ") + why).c_str());
        return BH_HIP_INVALID_ARGUMENT;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_err(err, err_cap, "bh_hip_create: no HIP device available");
        return BH_HIP_NOT_READY;
    }
    auto* h = new (std::nothrow) bh_hip_handle();
    if (!h) {
        set_err(err, err_cap, "bh_hip_create: out of host memory");
        return BH_HIP_INTERNAL_ERROR;
    }
    h->p = resolve(*params);
    h->test_fail_growth = std::getenv("BH_HIP_TEST_FAIL_NODE_GROWTH") != nullptr;
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
        h->cu_count = prop.multiProcessorCount;
    h->leaf_depth = reference_leaf_depth(n, h->p.leaf_capacity);
    bool ok = check(h, hipStreamCreateWithFlags(&h->d.stream, hipStreamNonBlocking), "stream") &&
              check(h, hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking), "stream");
    for (hipEvent_t* x : {&h->ev_sorted, &h->ev_topo, &h->ev_groups})
        ok = ok && check(h, hipEventCreateWithFlags(x, hipEventDisableTiming), "event");
    ok = ok && alloc_all(h, n);
    ok = ok && check(h, hipEventCreateWithFlags(&h->built, hipEventDisableTiming), "event") &&
         check(h, hipHostMalloc(reinterpret_cast<void**>(&h->err_host), sizeof(int),
                                hipHostMallocDefault),
               "pinned error word");
    if (ok)
        *h->err_host = 0;
    if (ok && h->p.pm_mesh > 0 && !bh::pm_setup(h->pm, h->p.pm_mesh, h->d.stream)) {
        h->last_error = "bh_hip_create: particle-mesh setup (rocFFT plans, meshes) failed";
        ok = false;
    }
    if (ok) {
        bh::launch_init_ids(h->d);
        ok = zero_err(h) && finish(h, "create") == BH_HIP_OK;
    }
    if (!ok) {
        set_err(err, err_cap, h->last_error);
        bh_hip_destroy(h);
        return BH_HIP_INTERNAL_ERROR;
    }
    *out = h;
    return BH_HIP_OK;
}

void bh_hip_destroy(bh_hip_handle_t* h)
{
    if (!h)
        return;
    if (h->d.stream)
        (void)hipStreamSynchronize(h->d.stream);
    if (h->side)
        (void)hipStreamSynchronize(h->side);
    bh::pm_release(h->pm);
    for (void* p : h->allocations)
        (void)hipFree(p);
    h->allocations.clear();
    for (auto& e : h->events) {
        for (hipEvent_t x : {e.b0, e.s0, e.s1, e.p1, e.t1, e.j1, e.b1, e.w1, e.l0})
            if (x)
                (void)hipEventDestroy(x);
    }
    for (auto& e : h->int_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (h->built)
        (void)hipEventDestroy(h->built);
    for (hipEvent_t x : {h->ev_sorted, h->ev_topo, h->ev_groups})
        if (x)
            (void)hipEventDestroy(x);
    if (h->err_host)
        (void)hipHostFree(h->err_host);
    if (h->d.stream)
        (void)hipStreamDestroy(h->d.stream);
    if (h->side)
        (void)hipStreamDestroy(h->side);
    delete h;
}

bh_hip_result_t bh_hip_set_params(bh_hip_handle_t* h, const bh_hip_params_t* params)
{
    if (!h || !params)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    const bh_hip_params_t old = h->p;
    const bh_hip_params_t next = resolve(*params);
    if (const char* why = invalid_combination(next)) {
        h->last_error = std::string("set_params: ") + why;
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (h->lff_cfg.enabled) {
        if (const char* why = lff_incompatible(next)) {
            h->last_error = std::string("set_params: ") + why;
            return BH_HIP_INVALID_ARGUMENT;
        }
    }
    if (next.adaptive_max_level != old.adaptive_max_level ||
        next.adaptive_eta != old.adaptive_eta || next.adaptive_flags != old.adaptive_flags)
        h->ad_primed = false;
    if (next.pm_mesh != h->pm.N) {
        if (next.pm_mesh > 0) {
            if (!bh::pm_setup(h->pm, next.pm_mesh, h->d.stream)) {
                bh::pm_release(h->pm);
                h->last_error = "set_params: particle-mesh setup (rocFFT plans, meshes) failed";
                return BH_HIP_INTERNAL_ERROR;
            }
        } else {
            (void)hipStreamSynchronize(h->d.stream);
            bh::pm_release(h->pm);
        }
    }
    h->p = next;
    h->p.waves_per_cu = old.waves_per_cu;  // residency is fixed at creation
    h->leaf_depth = reference_leaf_depth(h->d.n, h->p.leaf_capacity);
    h->primed = false;
    h->have_tree = false;
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_upload(bh_hip_handle_t* h, const double* pos3, const double* vel3,
                              const float* mass)
{
    if (!h || !pos3 || !mass)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::Device& d = h->d;
    // The store keeps its current order; rank maps caller index -> slot.
    std::vector<float> m_sorted;
    ensure_rank(h);
    bool ok = zero_err(h) && stage_to_device(h, pos3);
    if (ok) {
        bh::launch_upload_scatter(d, 0);
        bh::launch_wrap(d, h->p.periodic_box);
        ok = check(h, hipStreamSynchronize(d.stream), "upload pos");
    }
    if (ok) {
        if (vel3) {
            ok = stage_to_device(h, vel3);
        } else {
            ok = check(h, hipMemsetAsync(d.stage, 0, 3 * d.n * sizeof(double), d.stream),
                       "zero velocities");
        }
    }
    if (ok) {
        bh::launch_upload_scatter(d, 1);
        ok = check(h, hipStreamSynchronize(d.stream), "upload vel");
    }
    if (ok) {
        // Masses: reorder on the host by the store's id map.
        std::vector<uint32_t> id(d.n);
        ok = check(h, hipMemcpy(id.data(), d.id, d.n * sizeof(uint32_t), hipMemcpyDeviceToHost),
                   "id read");
        if (ok) {
            m_sorted.resize(d.n);
            for (size_t s = 0; s < d.n; ++s)
                m_sorted[s] = mass[id[s]];
            ok = check(h, hipMemcpy(d.mass, m_sorted.data(), d.n * sizeof(float),
                                    hipMemcpyHostToDevice),
                       "mass upload");
        }
    }
    h->primed = false;
    h->ad_primed = false;
    h->have_tree = false;
    h->amag_valid = false;
    if (!ok)
        return BH_HIP_INTERNAL_ERROR;
    return finish(h, "upload");
}

bh_hip_result_t bh_hip_download(bh_hip_handle_t* h, double* pos3, double* vel3)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::Device& d = h->d;
    ensure_rank(h);
    for (int which = 0; which < 2; ++which) {
        double* dst = which == 0 ? pos3 : vel3;
        if (!dst)
            continue;
        bh::launch_download_gather(d, which);
        if (!check(h, hipMemcpyAsync(dst, d.stage, 3 * d.n * sizeof(double),
                                     hipMemcpyDeviceToHost, d.stream),
                   "download") ||
            !check(h, hipStreamSynchronize(d.stream), "download"))
            return BH_HIP_INTERNAL_ERROR;
    }
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_build_tree(bh_hip_handle_t* h)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h) || !zero_err(h))
        return BH_HIP_INTERNAL_ERROR;
    begin_call(h);
    if (!build_checked(h, next_events(h))) {
        (void)hipStreamSynchronize(h->d.stream);
        end_call(h);
        return BH_HIP_INTERNAL_ERROR;
    }
    const bh_hip_result_t r = finish(h, "build_tree");
    end_call(h);
    return r;
}

bh_hip_result_t bh_hip_compute_accelerations(bh_hip_handle_t* h, float* acc3)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    if (h->p.key_mode != 0) {
        h->last_error = "compute_accelerations: key_mode REFERENCE30 builds topology only";
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (!usable(h) || !zero_err(h))
        return BH_HIP_INTERNAL_ERROR;
    begin_call(h);
    if (!queue_force(h)) {
        (void)hipStreamSynchronize(h->d.stream);
        h->primed = false;
        end_call(h);
        return BH_HIP_INTERNAL_ERROR;
    }
    bh_hip_result_t r = finish(h, "compute_accelerations");
    end_call(h);
    if (r != BH_HIP_OK || !acc3)
        return r;
    bh::Device& d = h->d;
    ensure_rank(h);
    bh::launch_download_gather(d, 2);
    std::vector<double> tmp(3 * d.n);
    if (!check(h, hipMemcpyAsync(tmp.data(), d.stage, tmp.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, d.stream),
               "acc download") ||
        !check(h, hipStreamSynchronize(d.stream), "acc download"))
        return BH_HIP_INTERNAL_ERROR;
    for (size_t i = 0; i < tmp.size(); ++i)
        acc3[i] = static_cast<float>(tmp[i]);
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_step(bh_hip_handle_t* h, double dt, int32_t nsteps)
{
    if (!h || !(dt > 0.0) || nsteps < 0)
        return BH_HIP_INVALID_ARGUMENT;
    if (h->p.key_mode != 0) {
        h->last_error = "step: key_mode REFERENCE30 builds topology only";
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (!usable(h) || !zero_err(h))
        return BH_HIP_INTERNAL_ERROR;
    if (h->p.adaptive_max_level > 0)
        return step_adaptive(h, dt, nsteps);
    begin_call(h);
    bh::Device& d = h->d;
    const double vmax = h->p.max_speed;
    auto fail = [&]() {
        (void)hipStreamSynchronize(d.stream);
        h->primed = false;
        h->steps_done = -1;
        end_call(h);
        return BH_HIP_INTERNAL_ERROR;
    };
    // Leapfrog: the closing half kick of a step and the opening half kick of
    // the next use the same acceleration, so without a velocity clamp they
    // are applied as one kick of dt (v + a dt instead of (v + a dt/2) + a dt/2,
    // which differ by rounding only). The last step of the call keeps its
    // closing half kick: velocities leave the call at whole steps.
    // A device error that no retry repairs (read at each build) stops the
    // loop; the message then says how many steps completed.
    bool half_kick_due = false;
    h->steps_done = 0;
    for (int k = 0; k < nsteps; ++k) {
        if (h->p.integrator == BH_HIP_INTEGRATOR_EULER) {
            if (!queue_force(h))
                return fail();
            if (bits_at_build(h) & kFatalBits)
                break;
            IntegrationEvents* ie = next_integration_events(h);
            rec(ie ? ie->a : nullptr, d.stream);
            bh::launch_kick_drift(d, dt, dt, vmax, h->p.periodic_box);
            rec(ie ? ie->b : nullptr, d.stream);
            h->primed = false;
        } else {
            if (!h->primed && !queue_force(h))
                return fail();
            IntegrationEvents* ie = next_integration_events(h);
            rec(ie ? ie->a : nullptr, d.stream);
            bh::launch_kick_drift(d, half_kick_due ? dt : 0.5 * dt, dt, vmax, h->p.periodic_box);
            rec(ie ? ie->b : nullptr, d.stream);
            if (!queue_force(h))
                return fail();
            const bool fatal = (bits_at_build(h) & kFatalBits) != 0;
            half_kick_due = vmax <= 0.0 && k + 1 < nsteps && !fatal;
            if (!half_kick_due) {
                ie = next_integration_events(h);
                rec(ie ? ie->a : nullptr, d.stream);
                bh::launch_kick(d, 0.5 * dt, vmax);
                rec(ie ? ie->b : nullptr, d.stream);
            }
            if (fatal)
                break;
        }
        h->steps_done = k + 1;
    }
    const bh_hip_result_t r = finish(h, "step");
    h->steps_done = -1;
    end_call(h);
    return r;
}

bh_hip_result_t bh_hip_tree_info(bh_hip_handle_t* h, uint64_t* node_count, int32_t* leaf_depth,
                                 int32_t* max_level)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    int info[2 * bh::kLevels + 4];
    if (!check(h, hipStreamSynchronize(h->d.stream), "tree_info") ||
        !check(h, hipMemcpy(info, h->d.level_info, sizeof(info), hipMemcpyDeviceToHost),
               "tree_info"))
        return BH_HIP_INTERNAL_ERROR;
    if (node_count)
        *node_count = static_cast<uint64_t>(info[2 * bh::kLevels]);
    if (leaf_depth)
        *leaf_depth = h->leaf_depth;
    if (max_level)
        *max_level = info[2 * bh::kLevels + 1];
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_dump_tree(bh_hip_handle_t* h, uint64_t* sorted_keys, uint32_t* sorted_ids,
                                 uint64_t* node_prefix, int32_t* node_level,
                                 uint32_t* node_start, uint32_t* node_count,
                                 int32_t* node_first_child, uint32_t* node_nchild,
                                 float* node_moments, double* geom, uint64_t cap_nodes)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    if (!h->have_tree) {
        h->last_error = "dump_tree: no tree built for the current positions";
        return BH_HIP_NOT_READY;
    }
    bh::Device& d = h->d;
    uint64_t nn = 0;
    if (bh_hip_tree_info(h, &nn, nullptr, nullptr) != BH_HIP_OK)
        return BH_HIP_INTERNAL_ERROR;
    if ((node_prefix || node_level || node_start || node_count || node_first_child ||
         node_nchild || node_moments) &&
        cap_nodes < nn) {
        h->last_error = "dump_tree: cap_nodes smaller than the node count";
        return BH_HIP_INVALID_ARGUMENT;
    }
    auto cp = [&](void* dst, const void* src, size_t bytes, const char* what) {
        return !dst || check(h, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), what);
    };
    bool ok = cp(sorted_keys, d.keys, d.n * 8, "keys") && cp(sorted_ids, d.id, d.n * 4, "ids") &&
              cp(node_prefix, d.nprefix, nn * 8, "prefix") &&
              cp(node_start, d.nstart, nn * 4, "start") && cp(node_count, d.ncount, nn * 4, "count") &&
              cp(node_nchild, d.nchild, nn * 4, "nchild") &&
              cp(node_moments, d.wnode, nn * sizeof(bh::WalkNode), "moments");
    if (ok && (node_level || node_moments)) {
        std::vector<uint8_t> lv(nn);
        bh::RootGeom g{};
        ok = cp(lv.data(), d.nlevel, nn, "level") && cp(&g, d.geom, sizeof(g), "geometry");
        for (uint64_t i = 0; ok && node_level && i < nn; ++i)
            node_level[i] = lv[i];
        // record slots 10, 11: the cell size (from the level) and trs
        const float wmax = std::max(std::max(g.width_f[0], g.width_f[1]), g.width_f[2]);
        for (uint64_t i = 0; ok && node_moments && i < nn; ++i) {
            float* rec = node_moments + 16 * i;
            rec[11] = rec[10];
            rec[10] = bh::cell_size(wmax, lv[i]);
        }
    }
    if (ok && node_first_child) {
        std::vector<uint32_t> f(nn), c(nn);
        ok = cp(f.data(), d.nfirst, nn * 4, "first") && cp(c.data(), d.nchild, nn * 4, "nchild");
        for (uint64_t i = 0; ok && i < nn; ++i)
            node_first_child[i] = c[i] ? static_cast<int32_t>(f[i]) : -1;
    }
    if (ok && geom) {
        bh::RootGeom g{};
        ok = cp(&g, d.geom, sizeof(g), "geometry");
        for (int a = 0; a < 3; ++a) {
            geom[a] = g.lo[a];
            geom[3 + a] = g.inv[a];
            geom[6 + a] = g.lo_d[a];
            geom[9 + a] = g.unit_d[a];
        }
    }
    return ok ? BH_HIP_OK : BH_HIP_INTERNAL_ERROR;
}

bh_hip_result_t bh_hip_direct_forces_fp64(bh_hip_handle_t* h, const uint32_t* targets,
                                          uint32_t nt, double* acc3)
{
    if (!h || (!targets && nt) || (!acc3 && nt))
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    if (nt == 0)
        return BH_HIP_OK;
    for (uint32_t i = 0; i < nt; ++i)
        if (targets[i] >= h->d.n) {
            h->last_error = "direct_forces_fp64: target index out of range";
            return BH_HIP_INVALID_ARGUMENT;
        }
    uint32_t* dt = nullptr;
    double* dout = nullptr;
    bool ok = dalloc(h, &dt, nt, "targets") && dalloc(h, &dout, 3 * static_cast<size_t>(nt), "out");
    ok = ok && check(h, hipMemcpy(dt, targets, nt * 4, hipMemcpyHostToDevice), "targets");
    if (ok) {
        ensure_rank(h);
        bh::launch_direct_fp64(h->d, dt, nt, h->p.softening, h->p.min_distance2, dout);
        ok = check(h, hipGetLastError(), "direct launch") &&
             check(h, hipStreamSynchronize(h->d.stream), "direct") &&
             check(h, hipMemcpy(acc3, dout, 3 * static_cast<size_t>(nt) * 8, hipMemcpyDeviceToHost),
                   "direct out");
    }
    dfree(h, dt);
    dfree(h, dout);
    return ok ? BH_HIP_OK : BH_HIP_INTERNAL_ERROR;
}

bh_hip_result_t bh_hip_diagnostics(bh_hip_handle_t* h, uint32_t sample_limit, double* out8)
{
    if (!h || !out8)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::Device& d = h->d;
    double* dres = nullptr;
    uint32_t* dslots = nullptr;
    double* dphi = nullptr;
    const size_t ns = (sample_limit == 0 || d.n <= sample_limit) ? d.n : sample_limit;
    std::vector<uint32_t> slots(ns);
    for (size_t k = 0; k < ns; ++k)
        slots[k] = static_cast<uint32_t>(ns == d.n ? k : (k * d.n) / ns);
    bool ok = dalloc(h, &dres, 8, "diag out") && dalloc(h, &dslots, ns, "diag slots") &&
              dalloc(h, &dphi, ns, "diag phi");
    ok = ok && check(h, hipMemcpy(dslots, slots.data(), ns * 4, hipMemcpyHostToDevice), "slots");
    std::vector<double> phi(ns);
    double base[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ok) {
        bh::launch_energy(d, dres);
        bh::launch_potential(d, dslots, static_cast<uint32_t>(ns), h->p.softening,
                             h->p.min_distance2, dphi);
        ok = check(h, hipGetLastError(), "diagnostics") &&
             check(h, hipStreamSynchronize(d.stream), "diagnostics") &&
             check(h, hipMemcpy(base, dres, sizeof(base), hipMemcpyDeviceToHost), "diag read") &&
             check(h, hipMemcpy(phi.data(), dphi, ns * 8, hipMemcpyDeviceToHost), "phi read");
    }
    dfree(h, dres);
    dfree(h, dslots);
    dfree(h, dphi);
    if (!ok)
        return BH_HIP_INTERNAL_ERROR;
    double pe = 0.0;
    for (double v : phi)
        pe += v;
    pe *= 0.5 * static_cast<double>(d.n) / static_cast<double>(ns);
    out8[0] = base[0];
    out8[1] = pe;
    for (int a = 0; a < 3; ++a) {
        out8[2 + a] = base[1 + a];
        out8[5 + a] = base[4 + a];
    }
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_set_timing(bh_hip_handle_t* h, int32_t enabled)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    h->timing = enabled != 0;
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_get_timings(bh_hip_handle_t* h, bh_hip_timings_t* out)
{
    if (!h || !out)
        return BH_HIP_INVALID_ARGUMENT;
    *out = h->t;
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_get_last_error(bh_hip_handle_t* h, char* buf, size_t cap)
{
    if (!h || !buf || cap == 0)
        return BH_HIP_INVALID_ARGUMENT;
    std::snprintf(buf, cap, "%s", h->last_error.c_str());
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_sort_pairs(const uint64_t* keys, const uint32_t* vals, uint64_t n,
                                  int32_t key_bits, uint64_t* keys_out, uint32_t* vals_out,
                                  char* err, size_t err_cap)
{
    if ((n && (!keys || !vals || !keys_out || !vals_out)) || key_bits < 1 || key_bits > 64 ||
        n > 0x7fffffffull) {
        set_err(err, err_cap, "bh_hip_sort_pairs: invalid argument");
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (n == 0)
        return BH_HIP_OK;
    bh_hip_handle tmp;
    bh::Device& d = tmp.d;
    d.n = n;
    bool ok = check(&tmp, hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking), "stream");
    d.sort_hist_cap = bh::radix_hist_words(n);
    ok = ok && dalloc(&tmp, &d.keys, n, "keys") && dalloc(&tmp, &d.keys_b, n, "keys_b") &&
         dalloc(&tmp, &d.vals, n, "vals") && dalloc(&tmp, &d.vals_b, n, "vals_b") &&
         dalloc_zero(&tmp, &d.sort_hist, d.sort_hist_cap, "hist", d.stream) &&
         dalloc_zero(&tmp, &d.err, 4, "error flags", d.stream) &&
         check(&tmp, hipStreamSynchronize(d.stream), "hist");
    ok = ok && check(&tmp, hipMemcpy(d.keys, keys, n * 8, hipMemcpyHostToDevice), "keys in") &&
         check(&tmp, hipMemcpy(d.vals, vals, n * 4, hipMemcpyHostToDevice), "vals in");
    if (ok) {
        bh::launch_radix_sort(d, n, key_bits);
        int flags = 0;
        ok = check(&tmp, hipGetLastError(), "sort") &&
             check(&tmp, hipStreamSynchronize(d.stream), "sort") &&
             check(&tmp, hipMemcpy(&flags, d.err, sizeof(int), hipMemcpyDeviceToHost), "sort") &&
             check(&tmp, hipMemcpy(keys_out, d.keys, n * 8, hipMemcpyDeviceToHost), "keys out") &&
             check(&tmp, hipMemcpy(vals_out, d.vals, n * 4, hipMemcpyDeviceToHost), "vals out");
        if (ok && flags != 0) {
            tmp.last_error = "[octree_hip] sort: scan did not complete";
            ok = false;
        }
    }
    for (void* p : tmp.allocations)
        (void)hipFree(p);
    tmp.allocations.clear();
    if (d.stream)
        (void)hipStreamDestroy(d.stream);
    if (!ok) {
        set_err(err, err_cap, tmp.last_error);
        return BH_HIP_INTERNAL_ERROR;
    }
    return BH_HIP_OK;
}

}  // extern "C"

// ---- learned far-field correction ------------------------------------------
namespace {

bool lff_alloc_core(bh_hip_handle* h)
{
    bh::LffState& L = h->lff;
    if (L.params)
        return true;
    L.partial_blocks = 256;
    if (!dalloc(h, &L.params, bh::kLffParams, "lff params") ||
        !dalloc(h, &L.adam_m, bh::kLffParams, "lff adam m") ||
        !dalloc(h, &L.adam_v, bh::kLffParams, "lff adam v") ||
        !dalloc(h, &L.grad, bh::kLffParams, "lff grad") ||
        !dalloc(h, &h->lff_train_params, bh::kLffParams, "lff train params") ||
        !dalloc(h, &L.partial, static_cast<size_t>(L.partial_blocks) * bh::kLffParams,
                "lff partial") ||
        !dalloc(h, &L.loss_partial, L.partial_blocks, "lff loss partial") ||
        !dalloc(h, &L.loss_acc, 4, "lff loss") || !dalloc(h, &L.bias, 144, "lff bias") ||
        !dalloc(h, &L.tuple_counter, 2, "lff counter"))
        return false;
    short* frag = nullptr;
    if (!dalloc(h, &frag, 14 * 64 * 8, "lff fragments"))  // 14 x 64 lanes x 8 BF16
        return false;
    L.frag = frag;
    // the features carry the nodes' third moments: computed by every build
    // from now on (the current tree has none)
    if (!h->d.noct) {
        if (!dalloc(h, &h->d.noct, 10 * h->d.node_cap, "node third moments"))
            return false;
        h->have_tree = false;
    }
    // He-normal hidden weights, small output layer, zero biases (seeded).
    std::vector<float> p(bh::kLffParams, 0.0f);
    std::mt19937_64 rng(20261202ull);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    const int NF = bh::kLffInputs, NH = bh::kLffHidden, NO = bh::kLffOutputs;
    for (int i = 0; i < NF * NH; ++i)
        p[i] = nd(rng) * std::sqrt(2.0f / 29.0f);
    for (int i = 0; i < NH * NH; ++i)
        p[NF * NH + NH + i] = nd(rng) * std::sqrt(2.0f / NH);
    for (int i = 0; i < NH * NO; ++i)
        p[NF * NH + NH + NH * NH + NH + i] = nd(rng) * 0.01f;
    for (int f : {11, 26, 27})  // the pad inputs carry no weight
        for (int k = 0; k < NH; ++k)
            p[f * NH + k] = 0.0f;
    if (!check(h, hipMemcpy(L.params, p.data(), p.size() * 4, hipMemcpyHostToDevice), "lff init"))
        return false;
    bh::launch_lff_pack(h->d, L);
    return check(h, hipStreamSynchronize(h->d.stream), "lff pack");
}

bool lff_alloc_lists(bh_hip_handle* h)
{
    bh::LffState& L = h->lff;
    const int cap = h->lff_cfg.list_cap > 0 ? h->lff_cfg.list_cap : 192;
    const size_t groups = h->d.n / 8 + 1024;
    if (L.lists && L.list_cap == cap && L.list_groups == groups)
        return true;
    dfree(h, L.lists);
    dfree(h, L.list_n);
    L.lists = nullptr;
    L.list_n = nullptr;
    L.list_cap = cap;
    L.list_groups = groups;
    return dalloc(h, &L.lists, groups * static_cast<size_t>(cap), "lff lists") &&
           dalloc(h, &L.list_n, groups, "lff list counts");
}

}  // namespace

extern "C" {

bh_hip_result_t bh_hip_lff_configure(bh_hip_handle_t* h, const bh_hip_lff_config_t* cfg)
{
    if (!h || !cfg || !(cfg->theta > 0.0f) || !(cfg->band > 0.0f) || cfg->list_cap < 0 ||
        cfg->min_count < 0)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    if (cfg->enabled) {
        if (const char* why = lff_incompatible(h->p)) {
            h->last_error = std::string("lff_configure: ") + why;
            return BH_HIP_INVALID_ARGUMENT;
        }
    }
    h->lff_cfg = *cfg;
    if (h->lff_cfg.list_cap == 0)
        h->lff_cfg.list_cap = 192;
    if (!lff_alloc_core(h) || !lff_alloc_lists(h))
        return BH_HIP_INTERNAL_ERROR;
    h->primed = false;
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_lff_collect(bh_hip_handle_t* h, int32_t group_stride, uint64_t max_tuples,
                                   uint64_t* total)
{
    if (!h || group_stride < 1 || max_tuples == 0 || max_tuples > 0xffffffffull)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    if (const char* why = lff_incompatible(h->p)) {
        h->last_error = std::string("lff_collect: ") + why;
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (!lff_alloc_core(h) || !lff_alloc_lists(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::LffState& L = h->lff;
    if (!L.X) {
        if (!dalloc(h, &L.X, max_tuples * bh::kLffInputs, "lff X") ||
            !dalloc(h, &L.Y, max_tuples * bh::kLffOutputs, "lff Y"))
            return BH_HIP_INTERNAL_ERROR;
        L.tuple_cap = max_tuples;
        bh::launch_fill_u32(L.tuple_counter, 0u, 1, h->d.stream);
    }
    if (!zero_err(h))
        return BH_HIP_INTERNAL_ERROR;
    begin_call(h);
    bh::WalkParams wp = walk_params(h->p);
    wp.theta2 = h->lff_cfg.theta * h->lff_cfg.theta;
    for (int attempt = 0;; ++attempt) {
        queue_build(h, nullptr);
        bh::launch_walk_record(h->d, wp, h->stats, L, h->lff_cfg.band, h->lff_cfg.min_count);
        const Built b = check_build(h);
        if (b == Built::kOk)
            break;
        if (b == Built::kFatal || attempt == 7) {
            (void)hipStreamSynchronize(h->d.stream);
            end_call(h);
            return BH_HIP_INTERNAL_ERROR;
        }
    }
    // pass 1: count the candidate pairs
    bh::launch_fill_u32(L.tuple_counter + 1, 0u, 1, h->d.stream);
    bh::launch_lff_tuples(h->d, L, group_stride, h->lff_cfg.band, wp.eps2, h->lff_cfg.min_count,
                          0xffffffffu, L.collects, true, L.tuple_counter + 1, 0u);
    bh_hip_result_t r = finish(h, "lff_collect");
    unsigned int cnt[2] = {0, 0};
    if (r == BH_HIP_OK &&
        !check(h, hipMemcpy(cnt, L.tuple_counter, 8, hipMemcpyDeviceToHost), "lff count"))
        r = BH_HIP_INTERNAL_ERROR;
    if (r == BH_HIP_OK && cnt[0] < L.tuple_cap && cnt[1] > 0) {
        // pass 2: a uniform hash subsample that fills (most of) the free room
        const double room = static_cast<double>(L.tuple_cap - cnt[0]);
        const double frac = std::min(1.0, 0.97 * room / static_cast<double>(cnt[1]));
        const uint32_t keep =
            frac >= 1.0 ? 0xffffffffu : static_cast<uint32_t>(frac * 4294967295.0);
        bh::launch_lff_tuples(h->d, L, group_stride, h->lff_cfg.band, wp.eps2,
                              h->lff_cfg.min_count, keep, L.collects, false, L.tuple_counter,
                              static_cast<unsigned int>(L.tuple_cap));
        r = finish(h, "lff_collect");
    }
    ++L.collects;
    end_call(h);
    h->primed = false;  // acc holds the uncorrected coarse walk
    if (r != BH_HIP_OK)
        return r;
    if (!check(h, hipMemcpy(cnt, L.tuple_counter, 4, hipMemcpyDeviceToHost), "lff count"))
        return BH_HIP_INTERNAL_ERROR;
    if (total)
        *total = std::min<uint64_t>(cnt[0], L.tuple_cap);
    return BH_HIP_OK;
}

bh_hip_result_t bh_hip_lff_train(bh_hip_handle_t* h, int32_t epochs, int32_t batch, float lr,
                                 uint64_t seed, float* loss_per_epoch)
{
    if (!h || epochs < 1 || batch < 64 || !(lr > 0.0f))
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::LffState& L = h->lff;
    if (!L.X || !L.params) {
        h->last_error = "lff_train: no training tuples (call bh_hip_lff_collect first)";
        return BH_HIP_NOT_READY;
    }
    unsigned int cnt = 0;
    if (!check(h, hipMemcpy(&cnt, L.tuple_counter, 4, hipMemcpyDeviceToHost), "lff count"))
        return BH_HIP_INTERNAL_ERROR;
    const size_t n = std::min<size_t>(cnt, L.tuple_cap);
    if (n < static_cast<size_t>(batch)) {
        h->last_error = "lff_train: fewer tuples than one batch";
        return BH_HIP_NOT_READY;
    }
    // Output scale: RMS of the targets.
    std::vector<float> y(3 * n);
    if (!check(h, hipMemcpy(y.data(), L.Y, y.size() * 4, hipMemcpyDeviceToHost), "lff Y"))
        return BH_HIP_INTERNAL_ERROR;
    double ss = 0.0;
    for (float v : y)
        ss += static_cast<double>(v) * v;
    const float sigma = static_cast<float>(std::sqrt(ss / static_cast<double>(y.size()))) + 1e-30f;
    // Training copy in normalised output units.
    std::vector<float> p(bh::kLffParams);
    if (!check(h, hipMemcpy(p.data(), L.params, p.size() * 4, hipMemcpyDeviceToHost), "lff p"))
        return BH_HIP_INTERNAL_ERROR;
    const int NF = bh::kLffInputs, NH = bh::kLffHidden, NO = bh::kLffOutputs;
    const size_t off_w3 = NF * NH + NH + NH * NH + NH;
    for (size_t i = off_w3; i < p.size(); ++i)
        p[i] /= sigma;
    if (!check(h, hipMemcpy(h->lff_train_params, p.data(), p.size() * 4, hipMemcpyHostToDevice),
               "lff p") ||
        !check(h, hipMemset(L.adam_m, 0, bh::kLffParams * 4), "adam") ||
        !check(h, hipMemset(L.adam_v, 0, bh::kLffParams * 4), "adam"))
        return BH_HIP_INTERNAL_ERROR;
    if (h->lff_order_cap < n) {
        dfree(h, h->lff_order);
        h->lff_order = nullptr;
        if (!dalloc(h, &h->lff_order, n, "lff order"))
            return BH_HIP_INTERNAL_ERROR;
        h->lff_order_cap = n;
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::mt19937_64 rng(seed);
    const size_t steps_per_epoch = n / static_cast<size_t>(batch);
    const int nblocks = std::min(L.partial_blocks, (batch + 63) / 64);
    int step = 0;
    const size_t total_steps = steps_per_epoch * static_cast<size_t>(epochs);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), rng);
        if (!check(h, hipMemcpyAsync(h->lff_order, order.data(), n * 4, hipMemcpyHostToDevice,
                                     h->d.stream),
                   "lff order") ||
            !check(h, hipMemsetAsync(L.loss_acc, 0, 4, h->d.stream), "lff loss"))
            return BH_HIP_INTERNAL_ERROR;
        for (size_t b = 0; b < steps_per_epoch; ++b) {
            ++step;
            const double frac = static_cast<double>(step) / static_cast<double>(total_steps);
            const float lr_t = static_cast<float>(lr * (0.1 + 0.9 * 0.5 * (1.0 + std::cos(M_PI * frac))));
            bh::launch_lff_grad(h->d, L, L.X, L.Y, h->lff_order, b * batch,
                                static_cast<uint32_t>(batch), 1.0f / sigma, h->lff_train_params,
                                nblocks);
            bh::launch_lff_adam(h->d, L, nblocks, h->lff_train_params, nullptr, lr_t, step);
        }
        float lsum = 0.0f;
        if (!check(h, hipGetLastError(), "lff train") ||
            !check(h, hipStreamSynchronize(h->d.stream), "lff train") ||
            !check(h, hipMemcpy(&lsum, L.loss_acc, 4, hipMemcpyDeviceToHost), "lff loss"))
            return BH_HIP_INTERNAL_ERROR;
        if (loss_per_epoch)
            loss_per_epoch[e] = lsum / static_cast<float>(steps_per_epoch * batch);
    }
    // Deploy: fold the output scale back into W3/b3 and repack the fragments.
    if (!check(h, hipMemcpy(p.data(), h->lff_train_params, p.size() * 4, hipMemcpyDeviceToHost),
               "lff p"))
        return BH_HIP_INTERNAL_ERROR;
    for (size_t i = off_w3; i < p.size(); ++i)
        p[i] *= sigma;
    (void)NO;
    return bh_hip_lff_set_params(h, p.data());
}

bh_hip_result_t bh_hip_lff_get_params(bh_hip_handle_t* h, float* params)
{
    if (!h || !params)
        return BH_HIP_INVALID_ARGUMENT;
    if (!lff_alloc_core(h))
        return BH_HIP_INTERNAL_ERROR;
    return check(h, hipMemcpy(params, h->lff.params, bh::kLffParams * 4, hipMemcpyDeviceToHost),
                 "lff get params")
               ? BH_HIP_OK
               : BH_HIP_INTERNAL_ERROR;
}

bh_hip_result_t bh_hip_lff_set_params(bh_hip_handle_t* h, const float* params)
{
    if (!h || !params)
        return BH_HIP_INVALID_ARGUMENT;
    if (!lff_alloc_core(h))
        return BH_HIP_INTERNAL_ERROR;
    if (!check(h, hipMemcpy(h->lff.params, params, bh::kLffParams * 4, hipMemcpyHostToDevice),
               "lff set params"))
        return BH_HIP_INTERNAL_ERROR;
    bh::launch_lff_pack(h->d, h->lff);
    h->primed = false;
    return check(h, hipStreamSynchronize(h->d.stream), "lff pack") ? BH_HIP_OK
                                                                    : BH_HIP_INTERNAL_ERROR;
}

bh_hip_result_t bh_hip_lff_gradient(bh_hip_handle_t* h, const float* X, const float* Y, uint32_t n,
                                    const float* params, float* grad, float* loss)
{
    if (!h || !X || !Y || !params || !grad || n == 0)
        return BH_HIP_INVALID_ARGUMENT;
    if (!lff_alloc_core(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::LffState& L = h->lff;
    float *dX = nullptr, *dY = nullptr, *dP = nullptr;
    uint32_t* dO = nullptr;
    std::vector<uint32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0u);
    const int nblocks = std::min<int>(L.partial_blocks, static_cast<int>((n + 63) / 64));
    bool ok = dalloc(h, &dX, static_cast<size_t>(n) * bh::kLffInputs, "grad X") &&
              dalloc(h, &dY, static_cast<size_t>(n) * bh::kLffOutputs, "grad Y") &&
              dalloc(h, &dP, bh::kLffParams, "grad P") && dalloc(h, &dO, n, "grad order");
    ok = ok && check(h, hipMemcpy(dX, X, static_cast<size_t>(n) * bh::kLffInputs * 4,
                                  hipMemcpyHostToDevice), "grad X") &&
         check(h, hipMemcpy(dY, Y, static_cast<size_t>(n) * bh::kLffOutputs * 4,
                            hipMemcpyHostToDevice), "grad Y") &&
         check(h, hipMemcpy(dP, params, bh::kLffParams * 4, hipMemcpyHostToDevice), "grad P") &&
         check(h, hipMemcpy(dO, ord.data(), n * 4, hipMemcpyHostToDevice), "grad order") &&
         check(h, hipMemset(L.loss_acc, 0, 4), "grad loss");
    if (ok) {
        bh::launch_lff_grad(h->d, L, dX, dY, dO, 0, n, 1.0f, dP, nblocks);
        bh::launch_lff_adam(h->d, L, nblocks, dP, L.grad, 0.0f, 1);
        float lsum = 0.0f;
        ok = check(h, hipGetLastError(), "grad") && check(h, hipStreamSynchronize(h->d.stream), "grad") &&
             check(h, hipMemcpy(grad, L.grad, bh::kLffParams * 4, hipMemcpyDeviceToHost), "grad") &&
             check(h, hipMemcpy(&lsum, L.loss_acc, 4, hipMemcpyDeviceToHost), "grad loss");
        if (loss)
            *loss = lsum / static_cast<float>(n);
    }
    dfree(h, dX);
    dfree(h, dY);
    dfree(h, dP);
    dfree(h, dO);
    return ok ? BH_HIP_OK : BH_HIP_INTERNAL_ERROR;
}

bh_hip_result_t bh_hip_lff_forward(bh_hip_handle_t* h, const float* X, uint32_t n, float* Y)
{
    if (!h || !X || !Y || n == 0)
        return BH_HIP_INVALID_ARGUMENT;
    if (!lff_alloc_core(h))
        return BH_HIP_INTERNAL_ERROR;
    float *dX = nullptr, *dY = nullptr;
    bool ok = dalloc(h, &dX, static_cast<size_t>(n) * bh::kLffInputs, "fwd X") &&
              dalloc(h, &dY, static_cast<size_t>(n) * bh::kLffOutputs, "fwd Y");
    ok = ok && check(h, hipMemcpy(dX, X, static_cast<size_t>(n) * bh::kLffInputs * 4,
                                  hipMemcpyHostToDevice), "fwd X");
    if (ok) {
        bh::launch_lff_forward(h->d, h->lff, dX, n, dY);
        ok = check(h, hipGetLastError(), "fwd") && check(h, hipStreamSynchronize(h->d.stream), "fwd") &&
             check(h, hipMemcpy(Y, dY, static_cast<size_t>(n) * bh::kLffOutputs * 4,
                                hipMemcpyDeviceToHost), "fwd Y");
    }
    dfree(h, dX);
    dfree(h, dY);
    return ok ? BH_HIP_OK : BH_HIP_INTERNAL_ERROR;
}

bh_hip_result_t bh_hip_lff_get_tuples(bh_hip_handle_t* h, float* X, float* Y, uint64_t cap,
                                      uint64_t* n)
{
    if (!h)
        return BH_HIP_INVALID_ARGUMENT;
    unsigned int cnt = 0;
    if (h->lff.X && !check(h, hipMemcpy(&cnt, h->lff.tuple_counter, 4, hipMemcpyDeviceToHost),
                           "lff count"))
        return BH_HIP_INTERNAL_ERROR;
    const uint64_t avail = std::min<uint64_t>(cnt, h->lff.tuple_cap);
    if (!X && !Y) {  // count query
        if (n)
            *n = avail;
        return BH_HIP_OK;
    }
    const uint64_t m = std::min<uint64_t>(avail, cap);
    if (m && X && !check(h, hipMemcpy(X, h->lff.X, m * bh::kLffInputs * 4, hipMemcpyDeviceToHost), "X"))
        return BH_HIP_INTERNAL_ERROR;
    if (m && Y && !check(h, hipMemcpy(Y, h->lff.Y, m * bh::kLffOutputs * 4, hipMemcpyDeviceToHost), "Y"))
        return BH_HIP_INTERNAL_ERROR;
    if (n)
        *n = m;
    return BH_HIP_OK;
}

}  // extern "C"

extern "C" bh_hip_result_t bh_hip_adaptive_levels(bh_hip_handle_t* h, int32_t* levels)
{
    if (!h || !levels)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    bh::Device& d = h->d;
    if (!d.alevel) {
        std::memset(levels, 0, d.n * sizeof(int32_t));
        return BH_HIP_OK;
    }
    ensure_rank(h);
    uint32_t* tmp = reinterpret_cast<uint32_t*>(d.stage);
    bh::launch_gather_u32(d, d.alevel, tmp);
    if (!check(h, hipGetLastError(), "adaptive_levels") ||
        !check(h, hipStreamSynchronize(d.stream), "adaptive_levels") ||
        !check(h, hipMemcpy(levels, tmp, d.n * sizeof(uint32_t), hipMemcpyDeviceToHost),
               "adaptive_levels"))
        return BH_HIP_INTERNAL_ERROR;
    return BH_HIP_OK;
}

extern "C" bh_hip_result_t bh_hip_pm_accelerations(bh_hip_handle_t* h, float* acc3)
{
    if (!h || !acc3)
        return BH_HIP_INVALID_ARGUMENT;
    if (!usable(h))
        return BH_HIP_INTERNAL_ERROR;
    if (h->pm.N == 0) {
        h->last_error = "pm_accelerations: no particle mesh (pm_mesh = 0)";
        return BH_HIP_NOT_READY;
    }
    bh::Device& d = h->d;
    if (!zero_err(h))
        return BH_HIP_INTERNAL_ERROR;
    for (int a = 0; a < 3; ++a)
        bh::launch_fill_u32(reinterpret_cast<uint32_t*>(d.acc[a]), 0u, d.n, d.stream);
    if (!bh::launch_pm_accel(d, h->pm, h->p.periodic_box, h->p.treepm_split_scale, 0.0f))
        h->pm_failed = true;
    h->primed = false;  // acc now holds the mesh part only
    bh_hip_result_t r = finish(h, "pm_accelerations");
    if (r != BH_HIP_OK)
        return r;
    ensure_rank(h);
    bh::launch_download_gather(d, 2);
    std::vector<double> tmp(3 * d.n);
    if (!check(h, hipMemcpyAsync(tmp.data(), d.stage, tmp.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, d.stream),
               "pm download") ||
        !check(h, hipStreamSynchronize(d.stream), "pm download"))
        return BH_HIP_INTERNAL_ERROR;
    for (size_t i = 0; i < tmp.size(); ++i)
        acc3[i] = static_cast<float>(tmp[i]);
    return BH_HIP_OK;
}
