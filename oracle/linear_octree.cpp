// TEST INFRASTRUCTURE ONLY -- the oracle. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this code, and only as the checker.
//
// CPU restatement of the reference linear octree (`octree_gpu`, CUDA build),
// which is the topology oracle, plus the production tree rule of octree_hip.
//
//   AABB                      engine/physics/octree/cuda/linear/OctBuild.inl:58-72
//   key generation            engine/physics/octree/cuda/linear/OctKeyKernels.inl:22-46
//   bit expansion             engine/physics/octree/morton/OctMortonCodes.inl:7-22
//   stable sort_by_key        engine/physics/octree/cuda/linear/OctBuild.inl:96
//   leaf depth rule           engine/physics/octree/cuda/linear/OctBuild.inl:28-39
//   leaf prefixes/reduce      engine/physics/octree/cuda/linear/OctBuild.inl:109-124
//   leaf nodes                engine/physics/octree/cuda/linear/OctLeafKernels.inl:7-81
//   parent levels             engine/physics/octree/cuda/linear/OctBuild.inl:156-199,
//                             engine/physics/octree/cuda/linear/OctParentKernels.inl:7-103
//   next links                engine/physics/octree/cuda/linear/OctLinkKernels.inl:7-70
//
// Reference quirk reproduced in key mode 1: the reference's "21-bit" expansion
// runs on a 32-bit unsigned with the 10-bit interleave masks, so only the low
// 10 bits of each 21-bit quantised coordinate reach the key (a 30-bit key).
// Key mode 0 is the same quantisation with a full 21-bit interleave (63 bits);
// it is the key octree_hip builds its tree from.
//
// Parity status: keys pinned by closed form (bit interleave, decode(encode)),
// brute force on small grids and the quantisation formula; sort by
// std::stable_sort; topology by integer invariants (tests/test_oracle_linear.py).
// The thrust/CUB arithmetic boundary itself is not vendored in the reference
// (CUDA Toolkit 12.6): parity unpinned there beyond the stable-sort contract.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {
// OctMortonCodes.inl:9-15 as the reference compiles it (32-bit lanes).
inline uint64_t expand_ref32(uint32_t v)
{
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return static_cast<uint64_t>(v);
}

// Full 21-bit interleave spread (bit k -> bit 3k), written as a loop so it is
// independent of the mask-based GPU version it checks.
inline uint64_t expand21_loop(uint32_t v)
{
    uint64_t r = 0;
    for (int k = 0; k < 21; ++k)
        r |= static_cast<uint64_t>((v >> k) & 1u) << (3 * k);
    return r;
}

inline uint32_t quantise(float p, float lo, float inv)
{
    const float t = std::fmin(std::fmax((p - lo) * inv, 0.0f), 1.0f);
    // nvcc contracts `t * 2097151.0f + 0.5f` into one FMA (default --fmad=true).
    const float s = std::fmaf(t, 2097151.0f, 0.5f);
    const uint32_t q = static_cast<uint32_t>(s);
    return std::min(2097151u, q);
}
}  // namespace

extern "C" {

// x bit at 3k+2, y at 3k+1, z at 3k.
uint64_t lo_encode(uint32_t qx, uint32_t qy, uint32_t qz, int key_mode)
{
    if (key_mode == 1)
        return (expand_ref32(qx) << 2) | (expand_ref32(qy) << 1) | expand_ref32(qz);
    return (expand21_loop(qx) << 2) | (expand21_loop(qy) << 1) | expand21_loop(qz);
}

// FP32 AABB over positions given as doubles cast to float (min over floats of
// the cast equals the cast of the FP64 min: rounding is monotonic).
void lo_aabb(int n, const double* pos3, float* lo3, float* hi3)
{
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float v = static_cast<float>(pos3[3 * i + a]);
            lo[a] = std::fmin(lo[a], v);
            hi[a] = std::fmax(hi[a], v);
        }
    for (int a = 0; a < 3; ++a) {
        lo3[a] = lo[a];
        hi3[a] = hi[a];
    }
}

// Quantised coordinates and keys; index output is the identity (OctKeyKernels.inl:22-46).
void lo_keys(int n, const double* pos3, int key_mode, uint64_t* keys, uint32_t* q3)
{
    float lo[3], hi[3];
    lo_aabb(n, pos3, lo, hi);
    float inv[3];
    for (int a = 0; a < 3; ++a)
        inv[a] = 1.0f / std::fmax(hi[a] - lo[a], 1.0e-12f);
    for (int i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a)
            q[a] = quantise(static_cast<float>(pos3[3 * i + a]), lo[a], inv[a]);
        keys[i] = lo_encode(q[0], q[1], q[2], key_mode);
        if (q3) {
            q3[3 * i] = q[0];
            q3[3 * i + 1] = q[1];
            q3[3 * i + 2] = q[2];
        }
    }
}

// Stable sort of (key, identity index) -- thrust::sort_by_key semantics.
void lo_sort(int n, const uint64_t* keys, uint64_t* sorted_keys, uint32_t* order)
{
    std::vector<uint32_t> idx(static_cast<size_t>(n));
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(),
                     [keys](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    for (int i = 0; i < n; ++i) {
        order[i] = idx[static_cast<size_t>(i)];
        sorted_keys[i] = keys[idx[static_cast<size_t>(i)]];
    }
}

// OctBuild.inl:28-39
int lo_leaf_depth(long long n, int leaf_capacity)
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

// Reference linear octree over already-sorted keys. Node order: leaves
// (unique prefixes at leafDepth, ascending) then each parent level bottom-up;
// root last. Outputs per node: level, prefix, particle start/count (subtree
// range in sorted order), first child, child count, parent, next (skip link),
// FP64 mass and mass-weighted position sums over the subtree in sorted order.
// Returns the node count, or -needed when cap is too small.
int lo_reference_tree(int n, const uint64_t* sorted_keys, const uint32_t* order,
                      const double* pos3, const float* mass, int leaf_depth, int cap,
                      int* level, uint64_t* prefix, int* start, int* count, int* child8,
                      int* parent, int* next, double* msum, double* mxyz)
{
    struct Node {
        int level;
        uint64_t prefix;
        int start, count;
        int child[8];
        int parent;
        double m, mx, my, mz;
    };
    std::vector<Node> nodes;
    const int shift = 3 * (21 - leaf_depth);
    // Leaves: runs of equal prefix (reduce_by_key + exclusive_scan).
    std::vector<int> cur;  // node ids of the current level
    for (int i = 0; i < n;) {
        const uint64_t p = sorted_keys[i] >> shift;
        int j = i;
        Node nd{};
        nd.level = leaf_depth;
        nd.prefix = p;
        nd.start = i;
        nd.parent = -1;
        for (int c = 0; c < 8; ++c)
            nd.child[c] = -1;
        while (j < n && (sorted_keys[j] >> shift) == p) {
            const uint32_t id = order[j];
            const double m = static_cast<double>(mass[id]);
            nd.m += m;
            nd.mx += static_cast<double>(static_cast<float>(pos3[3 * id])) * m;
            nd.my += static_cast<double>(static_cast<float>(pos3[3 * id + 1])) * m;
            nd.mz += static_cast<double>(static_cast<float>(pos3[3 * id + 2])) * m;
            ++j;
        }
        nd.count = j - i;
        cur.push_back(static_cast<int>(nodes.size()));
        nodes.push_back(nd);
        i = j;
    }
    int lvl = leaf_depth;
    while (cur.size() > 1) {
        std::vector<int> up;
        --lvl;
        for (size_t a = 0; a < cur.size();) {
            const uint64_t pp = nodes[static_cast<size_t>(cur[a])].prefix >> 3;
            Node nd{};
            nd.level = lvl;
            nd.prefix = pp;
            nd.parent = -1;
            for (int c = 0; c < 8; ++c)
                nd.child[c] = -1;
            nd.start = nodes[static_cast<size_t>(cur[a])].start;
            nd.count = 0;
            const int self = static_cast<int>(nodes.size());
            size_t b = a;
            std::vector<int> kids;
            while (b < cur.size() && (nodes[static_cast<size_t>(cur[b])].prefix >> 3) == pp) {
                kids.push_back(cur[b]);
                ++b;
            }
            for (int k : kids) {
                Node& ch = nodes[static_cast<size_t>(k)];
                int o = static_cast<int>(ch.prefix & 7u);
                if (nd.child[o] >= 0) {  // collision fallback, OctParentKernels.inl:48-58
                    o = 0;
                    while (o < 8 && nd.child[o] >= 0)
                        ++o;
                    if (o >= 8)
                        continue;
                }
                nd.child[o] = k;
                ch.parent = self;
                nd.count += ch.count;
                nd.m += ch.m;
                nd.mx += ch.mx;
                nd.my += ch.my;
                nd.mz += ch.mz;
            }
            up.push_back(self);
            nodes.push_back(nd);
            a = b;
        }
        cur.swap(up);
    }
    const int nn = static_cast<int>(nodes.size());
    if (nn > cap)
        return -nn;
    // Skip links: next existing sibling in octant order, else climb.
    for (int i = 0; i < nn; ++i) {
        int nx = -1, curi = i, par = nodes[static_cast<size_t>(i)].parent;
        while (par >= 0) {
            const Node& pn = nodes[static_cast<size_t>(par)];
            int slot = -1;
            for (int c = 0; c < 8; ++c)
                if (pn.child[c] == curi) {
                    slot = c;
                    break;
                }
            if (slot >= 0)
                for (int c = slot + 1; c < 8; ++c)
                    if (pn.child[c] >= 0) {
                        nx = pn.child[c];
                        break;
                    }
            if (nx >= 0)
                break;
            curi = par;
            par = nodes[static_cast<size_t>(curi)].parent;
        }
        next[i] = nx;
    }
    for (int i = 0; i < nn; ++i) {
        const Node& nd = nodes[static_cast<size_t>(i)];
        level[i] = nd.level;
        prefix[i] = nd.prefix;
        start[i] = nd.start;
        count[i] = nd.count;
        for (int c = 0; c < 8; ++c)
            child8[8 * i + c] = nd.child[c];
        parent[i] = nd.parent;
        msum[i] = nd.m;
        mxyz[3 * i] = nd.mx;
        mxyz[3 * i + 1] = nd.my;
        mxyz[3 * i + 2] = nd.mz;
    }
    return nn;
}

// octree_hip's production tree over sorted 63-bit keys, breadth-first.
// Rule: a node at level L < 21 is split when L < leaf_depth (the reference
// levels) or when it holds more than `bucket` particles. Children are the
// non-empty octants in ascending order; every level is stored contiguously
// and children of one parent are contiguous, so node order is
// (level, prefix)-ascending. Outputs level, prefix, start, count, first child
// (-1 for leaves) and child count. Returns node count (or -needed when cap is
// too small).
int lo_hip_tree(int n, const uint64_t* sorted_keys, int leaf_depth, int bucket, int cap,
                int* level, uint64_t* prefix, int* start, int* count, int* first_child,
                int* nchild)
{
    struct Node {
        int level;
        uint64_t prefix;
        int start, count, first, nc;
    };
    std::vector<Node> nodes;
    if (n <= 0)
        return 0;
    nodes.push_back(Node{0, 0u, 0, n, -1, 0});
    size_t lvl_begin = 0, lvl_end = 1;
    for (int L = 0; L < 21 && lvl_begin < lvl_end; ++L) {
        const int shift = 3 * (21 - (L + 1));
        for (size_t k = lvl_begin; k < lvl_end; ++k) {
            Node nd = nodes[k];
            const bool split = (L < leaf_depth) || (nd.count > bucket);
            if (!split)
                continue;
            nodes[k].first = static_cast<int>(nodes.size());
            int nc = 0;
            int i = nd.start;
            const int end = nd.start + nd.count;
            while (i < end) {
                const uint64_t p = sorted_keys[i] >> shift;
                int j = i;
                while (j < end && (sorted_keys[j] >> shift) == p)
                    ++j;
                nodes.push_back(Node{L + 1, p, i, j - i, -1, 0});
                ++nc;
                i = j;
            }
            nodes[k].nc = nc;
        }
        lvl_begin = lvl_end;
        lvl_end = nodes.size();
    }
    const int nn = static_cast<int>(nodes.size());
    if (nn > cap)
        return -nn;
    for (int i = 0; i < nn; ++i) {
        level[i] = nodes[static_cast<size_t>(i)].level;
        prefix[i] = nodes[static_cast<size_t>(i)].prefix;
        start[i] = nodes[static_cast<size_t>(i)].start;
        count[i] = nodes[static_cast<size_t>(i)].count;
        first_child[i] = nodes[static_cast<size_t>(i)].first;
        nchild[i] = nodes[static_cast<size_t>(i)].nc;
    }
    return nn;
}
}
