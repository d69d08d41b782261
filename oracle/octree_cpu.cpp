// TEST INFRASTRUCTURE ONLY -- the oracle. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this code, and only as the checker
// (or the timed CPU baseline). The product (libbh_hip.so) never links it.
//
// CPU restatement of the reference `octree_cpu` solver (CPU-only build):
//   * force law policy       engine/physics/core/force/PhyForceLawPolicy.cpp:22-32
//   * tree build              engine/physics/octree/build/OctreeBuild.cpp:13-17,54-86,147-216
//   * LIFO-stack walk          engine/physics/octree/traversal/OctreeTraversal.cpp:24-32,129-227
//   * pair acceleration        engine/physics/octree/force/OctreeForce.cpp:16-61 (SSE rsqrt + 1 NR)
//   * acceleration clamp       engine/physics/octree/model/OctHostMath.cpp:50-57
//   * force pass + OMP gate    engine/physics/octree/force/OctForces.cpp:19-31,118-135
//   * semi-implicit Euler      engine/physics/octree/model/OctParticleSystemHost.cpp:41-80
//
// Every floating-point operation is performed in FP32 in the same order as the
// reference, so on the same CPU the forces are bit-identical to the reference
// compiled from its own sources (checked by tests/test_oracle_ref.py).
//
// Parity status per function (how each is pinned):
//   oc_resolve_policy  : closed form (clamps), tests/test_oracle_octree_cpu.py
//   oc_build / walk    : bit-exact vs oracle/_ref (reference sources), the 9-body
//                        criterion fixture (TST_UNT_PHYS_019), theta->0 limit ==
//                        pairwise sum, two_body_rest.xyz closed form
//   oc_pair_accel      : closed form two-body (|a| = m / (d^2+eps^2)^1.5)
//   oc_euler_update    : closed form one step on two_body_rest.xyz
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>
#ifdef __SSE__
#include <xmmintrin.h>
#endif

extern "C" {
struct oc_params {
    float theta;          // requested opening angle
    float softening;      // requested Plummer softening
    float min_softening;  // physics(min_softening)
    float min_distance2;  // physics(min_distance2)
    float min_theta;      // physics(min_theta)
    float max_acceleration;  // physics(max_acceleration); <= 0 disables the clamp
    float cutoff2;        // TreePM cutoff squared (0 = none)
    float short_range_scale;  // TreePM split scale (0 = Newton)
    int criterion;        // 0 = centre of mass, 1 = bounds
    int deterministic;    // 1 = serial walk (reference deterministic_mode)
};
}

namespace {
struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }

constexpr float kPhysicsMinTheta = 0.05f;  // FndConstants.hpp:24
constexpr float kPhysicsMaxTheta = 4.0f;   // FndConstants.hpp:25

struct Policy {
    float theta, softening, minSoftening, minDistance2, minTheta, shortRange;
};

// PhyForceLawPolicy.cpp:22-32
Policy resolvePolicy(const oc_params& p)
{
    Policy q;
    q.minSoftening = std::max(0.0f, p.min_softening);
    q.minDistance2 = std::max(0.0f, p.min_distance2);
    q.minTheta = std::clamp(p.min_theta, kPhysicsMinTheta, kPhysicsMaxTheta);
    q.softening = std::max(p.softening, q.minSoftening);
    q.theta = std::clamp(p.theta, q.minTheta, kPhysicsMaxTheta);
    q.shortRange = p.short_range_scale;
    return q;
}

// OctreeForce.cpp:23-57. FP32 throughout; 1/sqrt from the SSE estimate plus
// one Newton step where SSE is compiled in (as the reference does).
inline float dist2Softened(V3 d, const Policy& q)
{
    return (d.x * d.x + d.y * d.y + d.z * d.z) + q.softening * q.softening;
}

inline V3 pairAccel(V3 self, V3 src, float m, const Policy& q)
{
    const V3 d = sub(src, self);
    const float r2 = dist2Softened(d, q);
    if (r2 <= q.minDistance2)
        return V3{0.0f, 0.0f, 0.0f};
    float inv = 1.0f / std::sqrt(r2);
#ifdef __SSE__
    inv = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(r2)));
    inv = inv * (1.5f - 0.5f * r2 * inv * inv);
#endif
    const float inv3 = inv * inv * inv;
    float w = 1.0f;
    if (q.shortRange > 0.0f) {
        const float r = 1.0f / inv;
        const float arg = 0.5f * r / q.shortRange;
        w = std::erfc(arg) + r * 0.5641895835477563f / q.shortRange * std::exp(-arg * arg);
    }
    return mul(d, 1.0f * m * inv3 * w);
}

// OctHostMath.cpp:50-57
inline V3 clampVec(V3 v, float limit)
{
    const float s = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (limit <= 0.0f || s <= limit || s <= 1e-6f)
        return v;
    return mul(v, limit / s);
}

// Pre-order pointer tree, OctreeBuild.cpp:13-17,147-216.
struct Tree {
    struct Node {
        V3 center;
        float half;
        float mass;
        V3 com;
        int child[8];
        unsigned char mask;
        int first;  // into leafIdx (leaves only)
        int count;
    };
    std::vector<Node> nodes;
    std::vector<int> leafIdx;  // concatenated leaf particle indices
    int root = -1;

    static constexpr size_t kLeafCap = 16u;
    static constexpr int kMaxDepth = 24;
    static constexpr float kMinHalf = 0.01f;

    static int octant(V3 p, V3 c)
    {
        return (p.x >= c.x ? 1 : 0) | (p.y >= c.y ? 2 : 0) | (p.z >= c.z ? 4 : 0);
    }

    int grow(const float* xyzm, const std::vector<int>& ids, V3 c, float h, int depth)
    {
        Node nd{};
        nd.center = c;
        nd.half = h;
        V3 wc{0.0f, 0.0f, 0.0f};
        float mt = 0.0f;
        for (int id : ids) {
            const float m = xyzm[4 * id + 3];
            mt += m;
            wc = add(wc, mul(v3(xyzm[4 * id], xyzm[4 * id + 1], xyzm[4 * id + 2]), m));
        }
        nd.mass = mt;
        nd.com = mt > 0.0f ? divs(wc, mt) : c;
        for (int& ch : nd.child)
            ch = -1;
        nd.mask = 0;
        nd.first = 0;
        nd.count = 0;
        const int self = static_cast<int>(nodes.size());
        nodes.push_back(nd);
        auto makeLeaf = [&]() {
            nodes[self].first = static_cast<int>(leafIdx.size());
            nodes[self].count = static_cast<int>(ids.size());
            leafIdx.insert(leafIdx.end(), ids.begin(), ids.end());
            return self;
        };
        if (ids.size() <= kLeafCap || h <= kMinHalf || depth >= kMaxDepth)
            return makeLeaf();
        std::array<std::vector<int>, 8> bucket;
        int used = 0;
        for (int id : ids) {
            const int o = octant(v3(xyzm[4 * id], xyzm[4 * id + 1], xyzm[4 * id + 2]), c);
            if (bucket[o].empty())
                ++used;
            bucket[o].push_back(id);
        }
        if (used <= 1)
            return makeLeaf();
        const float ch = h * 0.5f;
        for (int o = 0; o < 8; ++o) {
            if (bucket[o].empty())
                continue;
            const V3 cc = v3(c.x + ((o & 1) ? ch : -ch), c.y + ((o & 2) ? ch : -ch),
                             c.z + ((o & 4) ? ch : -ch));
            const int k = grow(xyzm, bucket[o], cc, ch, depth + 1);
            nodes[self].child[o] = k;
            nodes[self].mask |= static_cast<unsigned char>(1u << o);
        }
        return self;
    }

    // OctreeBuild.cpp:54-86
    void build(const float* xyzm, int n)
    {
        nodes.clear();
        leafIdx.clear();
        root = -1;
        if (n <= 0)
            return;
        nodes.reserve(static_cast<size_t>(n) * 2u);
        V3 lo = v3(xyzm[0], xyzm[1], xyzm[2]);
        V3 hi = lo;
        for (int i = 1; i < n; ++i) {
            lo.x = std::min(lo.x, xyzm[4 * i]);
            lo.y = std::min(lo.y, xyzm[4 * i + 1]);
            lo.z = std::min(lo.z, xyzm[4 * i + 2]);
            hi.x = std::max(hi.x, xyzm[4 * i]);
            hi.y = std::max(hi.y, xyzm[4 * i + 1]);
            hi.z = std::max(hi.z, xyzm[4 * i + 2]);
        }
        const V3 c = v3((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
        const float sx = hi.x - lo.x, sy = hi.y - lo.y, sz = hi.z - lo.z;
        const float h = std::max(0.5f * std::max(sx, std::max(sy, sz)), kMinHalf) + 0.001f;
        std::vector<int> all(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            all[static_cast<size_t>(i)] = i;
        root = grow(xyzm, all, c, h, 0);
    }

    // OctreeTraversal.cpp:129-222: LIFO stack, children pushed 0..7.
    V3 accelOn(const float* xyzm, int self, const Policy& q, int criterion, float cutoff2,
               std::vector<int>& stack) const
    {
        V3 acc{0.0f, 0.0f, 0.0f};
        if (root < 0)
            return acc;
        const V3 p = v3(xyzm[4 * self], xyzm[4 * self + 1], xyzm[4 * self + 2]);
        stack.clear();
        stack.push_back(root);
        while (!stack.empty()) {
            const int ni = stack.back();
            stack.pop_back();
            if (ni < 0)
                continue;
            const Node& nd = nodes[static_cast<size_t>(ni)];
            if (nd.mass <= 0.0f)
                continue;
            if (cutoff2 > 0.0f) {
                const float dx = std::max(std::fabs(p.x - nd.center.x) - nd.half, 0.0f);
                const float dy = std::max(std::fabs(p.y - nd.center.y) - nd.half, 0.0f);
                const float dz = std::max(std::fabs(p.z - nd.center.z) - nd.half, 0.0f);
                if (dx * dx + dy * dy + dz * dz > cutoff2)
                    continue;
            }
            if (nd.mask == 0u) {
                for (int k = 0; k < nd.count; ++k) {
                    const int j = leafIdx[static_cast<size_t>(nd.first + k)];
                    if (j == self)
                        continue;
                    const V3 s = v3(xyzm[4 * j], xyzm[4 * j + 1], xyzm[4 * j + 2]);
                    if (cutoff2 > 0.0f) {
                        const V3 d = sub(s, p);
                        if (d.x * d.x + d.y * d.y + d.z * d.z > cutoff2)
                            continue;
                    }
                    acc = add(acc, pairAccel(p, s, xyzm[4 * j + 3], q));
                }
                continue;
            }
            const float size = nd.half * 2.0f;
            const bool inside = std::fabs(p.x - nd.center.x) <= nd.half &&
                                std::fabs(p.y - nd.center.y) <= nd.half &&
                                std::fabs(p.z - nd.center.z) <= nd.half;
            const float d2 = dist2Softened(sub(nd.com, p), q);
            float cd = std::max(std::sqrt(d2), 1.0e-6f);
            if (criterion == 1) {
                const float dx = std::max(std::fabs(p.x - nd.center.x) - nd.half, 0.0f);
                const float dy = std::max(std::fabs(p.y - nd.center.y) - nd.half, 0.0f);
                const float dz = std::max(std::fabs(p.z - nd.center.z) - nd.half, 0.0f);
                cd = std::max(std::sqrt(dx * dx + dy * dy + dz * dz), 1.0e-6f);
            }
            const float ex = std::fabs(p.x - nd.center.x) + nd.half;
            const float ey = std::fabs(p.y - nd.center.y) + nd.half;
            const float ez = std::fabs(p.z - nd.center.z) + nd.half;
            const bool within = cutoff2 <= 0.0f || ex * ex + ey * ey + ez * ez <= cutoff2;
            if (within && !inside && (size / cd) < q.theta) {
                acc = add(acc, pairAccel(p, nd.com, nd.mass, q));
                continue;
            }
            for (int o = 0; o < 8; ++o)
                if (nd.mask & (1u << o))
                    stack.push_back(nd.child[o]);
        }
        return acc;
    }
};
}  // namespace

extern "C" {

// Resolved policy values (theta, softening, minDistance2) as the walk sees them.
void oc_resolve_policy(const oc_params* p, float* theta, float* softening, float* min_distance2)
{
    const Policy q = resolvePolicy(*p);
    *theta = q.theta;
    *softening = q.softening;
    *min_distance2 = q.minDistance2;
}

// Single pair acceleration (OctreeForce.cpp:23-57) for closed-form checks.
void oc_pair_accel(const float* self3, const float* src3, float mass, const oc_params* p,
                   float* out3)
{
    const Policy q = resolvePolicy(*p);
    const V3 a = pairAccel(v3(self3[0], self3[1], self3[2]), v3(src3[0], src3[1], src3[2]),
                           mass, q);
    out3[0] = a.x;
    out3[1] = a.y;
    out3[2] = a.z;
}

// Octree force pass over hot data xyzm (x, y, z, m per particle), then the
// acceleration clamp (OctForces.cpp:118-133). nthreads <= 0: OpenMP default.
// Returns the node count of the tree that was built.
int oc_octree_forces(int n, const float* xyzm, const oc_params* p, float* acc3, int nthreads)
{
    const Policy q = resolvePolicy(*p);
    Tree tree;
    tree.build(xyzm, n);
    const int threads = p->deterministic ? 1 : (nthreads > 0 ? nthreads : omp_get_max_threads());
#pragma omp parallel num_threads(threads)
    {
        std::vector<int> stack;
        stack.reserve(64);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3 a = tree.accelOn(xyzm, i, q, p->criterion, p->cutoff2, stack);
            a = clampVec(a, p->max_acceleration);
            acc3[3 * i] = a.x;
            acc3[3 * i + 1] = a.y;
            acc3[3 * i + 2] = a.z;
        }
    }
    return static_cast<int>(tree.nodes.size());
}

// Tree shape export for topology checks: per node centre/half/mass/com,
// child mask, leaf range; plus the concatenated leaf index list.
int oc_octree_export(int n, const float* xyzm, int cap_nodes, float* node8, int* child8,
                     int* mask_first_count, int* leaf_idx)
{
    Tree tree;
    tree.build(xyzm, n);
    const int nn = static_cast<int>(tree.nodes.size());
    if (cap_nodes < nn)
        return -nn;
    for (int i = 0; i < nn; ++i) {
        const auto& nd = tree.nodes[static_cast<size_t>(i)];
        float* o = node8 + 8 * i;
        o[0] = nd.center.x;
        o[1] = nd.center.y;
        o[2] = nd.center.z;
        o[3] = nd.half;
        o[4] = nd.mass;
        o[5] = nd.com.x;
        o[6] = nd.com.y;
        o[7] = nd.com.z;
        for (int c = 0; c < 8; ++c)
            child8[8 * i + c] = nd.child[c];
        mask_first_count[3 * i] = nd.mask;
        mask_first_count[3 * i + 1] = nd.first;
        mask_first_count[3 * i + 2] = nd.count;
    }
    if (leaf_idx)
        std::memcpy(leaf_idx, tree.leafIdx.data(), tree.leafIdx.size() * sizeof(int));
    return nn;
}

// One ParticleSystem::update(dt) of the CPU-only build
// (OctParticleSystemHost.cpp:41-80): forces, then per particle
// v = clamp(v + a dt, max_speed); x += v dt. FP32 state, in place.
int oc_euler_update(int n, float* xyzm, float* vel3, float dt, float max_speed,
                    const oc_params* p, float* acc3, int nthreads)
{
    if (n <= 0 || dt <= 0.0f)
        return 0;
    oc_octree_forces(n, xyzm, p, acc3, nthreads);
    for (int i = 0; i < n; ++i) {
        V3 v = v3(vel3[3 * i] + acc3[3 * i] * dt, vel3[3 * i + 1] + acc3[3 * i + 1] * dt,
                  vel3[3 * i + 2] + acc3[3 * i + 2] * dt);
        v = clampVec(v, max_speed);
        vel3[3 * i] = v.x;
        vel3[3 * i + 1] = v.y;
        vel3[3 * i + 2] = v.z;
        xyzm[4 * i] = xyzm[4 * i] + v.x * dt;
        xyzm[4 * i + 1] = xyzm[4 * i + 1] + v.y * dt;
        xyzm[4 * i + 2] = xyzm[4 * i + 2] + v.z * dt;
    }
    return 1;
}

// Kick-drift-kick leapfrog driven by the same force pass (the corrected
// `integrator=leapfrog` mapping; the reference headless path runs Euler).
// acc3 must hold a(x) on entry (call oc_octree_forces once to prime) and
// holds a(x_new) on exit.
int oc_kdk_step(int n, float* xyzm, float* vel3, float dt, float max_speed, const oc_params* p,
                float* acc3, int nthreads)
{
    if (n <= 0 || dt <= 0.0f)
        return 0;
    const float h = 0.5f * dt;
    for (int i = 0; i < n; ++i) {
        V3 v = v3(vel3[3 * i] + acc3[3 * i] * h, vel3[3 * i + 1] + acc3[3 * i + 1] * h,
                  vel3[3 * i + 2] + acc3[3 * i + 2] * h);
        v = clampVec(v, max_speed);
        vel3[3 * i] = v.x;
        vel3[3 * i + 1] = v.y;
        vel3[3 * i + 2] = v.z;
        xyzm[4 * i] += v.x * dt;
        xyzm[4 * i + 1] += v.y * dt;
        xyzm[4 * i + 2] += v.z * dt;
    }
    oc_octree_forces(n, xyzm, p, acc3, nthreads);
    for (int i = 0; i < n; ++i) {
        V3 v = v3(vel3[3 * i] + acc3[3 * i] * h, vel3[3 * i + 1] + acc3[3 * i + 1] * h,
                  vel3[3 * i + 2] + acc3[3 * i + 2] * h);
        v = clampVec(v, max_speed);
        vel3[3 * i] = v.x;
        vel3[3 * i + 1] = v.y;
        vel3[3 * i + 2] = v.z;
    }
    return 1;
}

// Subset walk: forces on the listed targets only (for sampled comparisons at
// large N where a full CPU pass would be slow).
int oc_octree_forces_subset(int n, const float* xyzm, const oc_params* p, int nt,
                            const int* targets, float* acc3, int nthreads)
{
    const Policy q = resolvePolicy(*p);
    Tree tree;
    tree.build(xyzm, n);
    const int threads = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(threads)
    {
        std::vector<int> stack;
        stack.reserve(64);
#pragma omp for schedule(dynamic, 16)
        for (int t = 0; t < nt; ++t) {
            V3 a = tree.accelOn(xyzm, targets[t], q, p->criterion, p->cutoff2, stack);
            a = clampVec(a, p->max_acceleration);
            acc3[3 * t] = a.x;
            acc3[3 * t + 1] = a.y;
            acc3[3 * t + 2] = a.z;
        }
    }
    return static_cast<int>(tree.nodes.size());
}
}
