// Linear octree build and upward moment pass for octree_hip.
//
// Topology contract. The reference linear octree keeps every distinct key
// prefix down to a fixed leaf depth D chosen from the average density
// (engine/physics/octree/cuda/linear/OctBuild.inl:28-39, 109-199). octree_hip
// keeps exactly those nodes (levels 0..D) and additionally splits any node
// holding more than `bucket` particles, continuing the same prefix
// subdivision down to level 21. So the reference tree is a truncation of this
// tree at level D (checked bit-exactly by the tests).
//
// Layout: breadth-first. Level L occupies node indices
// [base[L], base[L] + count[L]); children of one node are contiguous and in
// octant order, so node order equals (level, prefix) order. Level counts live
// on the device and every level kernel reads them there: the build issues no
// host synchronisation.
//
// Moments (per node, FP32, relative to the node's exact lattice centre):
// mass, centre-of-mass offset, traceless quadrupole about the COM, and bmax
// (radius about the COM that encloses every particle of the node). Leaves
// read their particles' FP64 positions and store FP32 offsets from the leaf's
// FP64 lattice centre (the only FP64 arithmetic here). Parents combine
// children with the parallel-axis theorem, using exact integer differences of
// lattice centres.
#include "bh_internal.h"
#include "scan_lookback.h"

namespace bh {
namespace {

__device__ inline uint32_t compact3(unsigned long long x)
{
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x00000000001fffffull;
    return static_cast<uint32_t>(x);
}

// Lattice centre of the cell with `prefix` at `level`: (2i+1)*2^(21-L) - 1.
__device__ inline void lattice_centre(unsigned long long prefix, int level, int* c)
{
    const uint32_t ix = compact3(prefix >> 2), iy = compact3(prefix >> 1), iz = compact3(prefix);
    const int sh = kMaxLevel - level;
    c[0] = static_cast<int>(((2u * ix + 1u) << sh) - 1u);
    c[1] = static_cast<int>(((2u * iy + 1u) << sh) - 1u);
    c[2] = static_cast<int>(((2u * iz + 1u) << sh) - 1u);
}

__device__ inline int* level_base(int* info) { return info; }
__device__ inline int* level_count(int* info) { return info + kLevels; }

__global__ void tree_root_kernel(uint64_t* nprefix, uint8_t* nlevel, uint32_t* nstart,
                                 uint32_t* ncount, uint32_t* nchild, uint32_t* nfirst,
                                 uint32_t* nparent, int* info, uint32_t n)
{
    const int t = threadIdx.x;
    if (t < kLevels) {
        info[t] = 0;
        info[kLevels + t] = 0;
    }
    if (t == 0) {
        nprefix[0] = 0;
        nlevel[0] = 0;
        nstart[0] = 0;
        ncount[0] = n;
        nchild[0] = 0;
        nfirst[0] = 0;
        nparent[0] = 0xffffffffu;
        info[kLevels] = 1;  // count[0]
        info[2 * kLevels] = 1;
        info[2 * kLevels + 1] = 0;
        info[kInfoBig] = 0;
    }
}

// Boundaries of the 8 child octants of a node spanning sorted keys [a, b):
// bnd[o] = first index whose level-(L+1) digit is >= o; bnd[8] = b. Keys are
// sorted, so bnd[o] = a + #{keys in [a, b) with digit < o}: nodes of up to 16
// keys count with independent loads; larger nodes run the seven binary
// searches interleaved (seven independent loads per step, one dependent step
// per halving instead of seven).
__device__ inline void child_bounds(const unsigned long long* __restrict__ keys, uint32_t a,
                                    uint32_t b, int shift, uint32_t* bnd)
{
    bnd[0] = a;
    bnd[8] = b;
    if (b - a <= 16) {
        uint32_t below[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
            if (a + k < b) {
                const uint32_t dg = static_cast<uint32_t>((keys[a + k] >> shift) & 7u);
#pragma unroll
                for (uint32_t o = 1; o < 8; ++o)
                    below[o] += dg < o ? 1u : 0u;
            }
        }
#pragma unroll
        for (int o = 1; o < 8; ++o)
            bnd[o] = a + below[o];
        return;
    }
    uint32_t lo[8], hi[8];
#pragma unroll
    for (int o = 1; o < 8; ++o) {
        lo[o] = a;
        hi[o] = b;
    }
    for (uint32_t len = b - a; len > 0; len >>= 1) {  // ceil(log2(b - a + 1)) rounds
#pragma unroll
        for (int o = 1; o < 8; ++o) {
            if (lo[o] < hi[o]) {
                const uint32_t mid = lo[o] + ((hi[o] - lo[o]) >> 1);
                if (static_cast<int>((keys[mid] >> shift) & 7u) < o)
                    lo[o] = mid + 1;
                else
                    hi[o] = mid;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 8; ++o) {
        // a search still open after the rounds cannot happen: each round at
        // least halves hi - lo, and b - a halvings reach zero
        bnd[o] = lo[o];
    }
}

// child_bounds of one node [a, a + c) by a whole wavefront: lanes 8o..8o+7
// (o = 1..7) bracket boundary o with 8 probes per step, so each step shrinks
// the bracket ninefold (about 8 dependent key loads at 8M particles instead of
// 7 binary searches of 23 loads each). Every lane returns all 9 bounds. Used
// for levels with few nodes, where one thread per node would leave the
// search a long chain of dependent loads.
constexpr int kWavePerNodeMax = 4096;

__device__ inline void wave_child_bounds(const unsigned long long* __restrict__ keys, uint32_t a,
                                         uint32_t c, int shift, uint32_t* bnd)
{
    const int lane = threadIdx.x & 63;
    const int o = lane >> 3, j = lane & 7;
    uint32_t lo = a, hi = a + c;  // boundary o lies in [lo, hi]
    for (;;) {
        const bool open = o > 0 && lo < hi;
        if (!__any(open))
            break;
        const uint32_t len = hi - lo;
        const uint32_t pj = lo + static_cast<uint32_t>((static_cast<uint64_t>(len) * (j + 1)) / 9u);
        const bool pred = open && static_cast<int>((keys[pj] >> shift) & 7u) >= o;
        const unsigned long long m = __ballot(pred);
        if (open) {
            const uint32_t gm = static_cast<uint32_t>(m >> (8 * o)) & 0xffu;
            const int t = gm ? static_cast<int>(__ffs(gm)) - 1 : 8;
            const uint32_t p_t =
                lo + static_cast<uint32_t>((static_cast<uint64_t>(len) * (t + 1)) / 9u);
            const uint32_t p_prev =
                t > 0 ? lo + static_cast<uint32_t>((static_cast<uint64_t>(len) * t) / 9u) : 0u;
            const uint32_t nlo = t > 0 ? p_prev + 1u : lo;
            const uint32_t nhi = t < 8 ? p_t : hi;
            lo = nlo;
            hi = nhi;
        }
    }
    bnd[0] = a;
    bnd[8] = a + c;
    for (int q = 1; q < 8; ++q)
        bnd[q] = __shfl(lo, 8 * q, 64);  // lane 8q holds boundary q
}

// Nodes of more than this many keys are searched by a whole wavefront even
// on levels run thread-per-node: a dense clump can hold 10^5..10^6 keys in
// one node, and one thread's 20 dependent search rounds would set the time
// of the whole level. tree_count_kernel lists them; tree_count_big_kernel
// spreads them over wavefronts (the ninefold bracketing takes about 6 rounds).
constexpr uint32_t kWaveSearchMin = 2048;

__device__ inline void store_count(uint32_t* __restrict__ nchild, uint32_t* __restrict__ nfirst,
                                   uint32_t* __restrict__ scan_tmp, uint32_t* __restrict__ bnd_tmp,
                                   int node, int k, uint32_t a, bool split, const uint32_t* bnd)
{
    uint32_t nc = 0;
    if (split) {
        for (int o = 0; o < 8; ++o)
            nc += bnd[o + 1] > bnd[o] ? 1u : 0u;
        // bounds 1..8 for the emit pass (bound 0 is the node start)
        uint4* out = reinterpret_cast<uint4*>(bnd_tmp + static_cast<size_t>(k) * 8);
        out[0] = make_uint4(bnd[1], bnd[2], bnd[3], bnd[4]);
        out[1] = make_uint4(bnd[5], bnd[6], bnd[7], bnd[8]);
    }
    nchild[node] = nc;
    nfirst[node] = a;  // leaves keep their particle start; parents are overwritten
    scan_tmp[k] = nc;
}

__global__ __launch_bounds__(256) void tree_count_kernel(
    const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, uint32_t* __restrict__ nchild,
    uint32_t* __restrict__ nfirst, uint32_t* __restrict__ scan_tmp,
    uint32_t* __restrict__ bnd_tmp, uint32_t* __restrict__ big_list, int* __restrict__ info,
    int level, int leaf_depth, uint32_t bucket)
{
    const int base = level_base(info)[level];
    const int cnt = level_count(info)[level];
    const int shift = 3 * (kMaxLevel - level - 1);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (cnt <= kWavePerNodeMax) {  // uniform across the grid: a wavefront per node
        const int nwaves = static_cast<int>((gridDim.x * blockDim.x) >> 6);
        for (int k = tid >> 6; k < cnt; k += nwaves) {
            const int node = base + k;
            const uint32_t a = nstart[node], c = ncount[node];
            const bool split = level < kMaxLevel && (level < leaf_depth || c > bucket);
            uint32_t bnd[9];
            if (split)  // uniform across the wave
                wave_child_bounds(keys, a, c, shift, bnd);
            if ((threadIdx.x & 63) == 0)
                store_count(nchild, nfirst, scan_tmp, bnd_tmp, node, k, a, split, bnd);
        }
        return;
    }
    for (int k = tid; k < cnt; k += gridDim.x * blockDim.x) {
        const int node = base + k;
        const uint32_t a = nstart[node], c = ncount[node];
        const bool split = level < kMaxLevel && (level < leaf_depth || c > bucket);
        if (split && c > kWaveSearchMin) {
            big_list[atomicAdd(&info[kInfoBig], 1)] = static_cast<uint32_t>(k);
            continue;
        }
        uint32_t bnd[9];
        if (split)
            child_bounds(keys, a, a + c, shift, bnd);
        store_count(nchild, nfirst, scan_tmp, bnd_tmp, node, k, a, split, bnd);
    }
}

// The large nodes tree_count_kernel listed: a wavefront per node.
__global__ __launch_bounds__(256) void tree_count_big_kernel(
    const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, uint32_t* __restrict__ nchild,
    uint32_t* __restrict__ nfirst, uint32_t* __restrict__ scan_tmp,
    uint32_t* __restrict__ bnd_tmp, const uint32_t* __restrict__ big_list,
    const int* __restrict__ info, int level)
{
    const int base = info[level];  // level_base
    const int nbig = info[kInfoBig];
    const int shift = 3 * (kMaxLevel - level - 1);
    const int nwaves = static_cast<int>((gridDim.x * blockDim.x) >> 6);
    for (int i = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6); i < nbig;
         i += nwaves) {
        const int k = static_cast<int>(big_list[i]);
        const int node = base + k;
        const uint32_t a = nstart[node], c = ncount[node];
        uint32_t bnd[9];
        wave_child_bounds(keys, a, c, shift, bnd);
        if ((threadIdx.x & 63) == 0)
            store_count(nchild, nfirst, scan_tmp, bnd_tmp, node, k, a, true, bnd);
    }
}

// After the scan of level `level` (whose total is the node count of level+1):
// place level+1, clamp to capacity.
__device__ inline void level_fix(int* info, int level, int next_count, int node_cap, int* err)
{
    int* base = level_base(info);
    int* count = level_count(info);
    const int nb = base[level] + count[level];
    base[level + 1] = nb;
    int c = next_count;
    if (nb + c > node_cap) {
        // Out of node slots: level+1 is dropped and every node of `level` that
        // wanted to split stays a leaf (see tree_emit_kernel). Forces remain
        // correct; the host grows the capacity for the next build.
        atomicOr(err, kErrNodeOverflow);
        c = 0;
    }
    count[level + 1] = c;
    if (c > 0)
        info[2 * kLevels + 1] = level + 1;  // deepest populated level
    info[2 * kLevels] = nb + c;             // total nodes
    info[kInfoBig] = 0;                     // large-node list of the next level
}

// Exclusive scan of the child counts of `level` (scan_tmp, one word per node
// of the level) in one pass, then level_fix in the block that ends the scan.
__global__ __launch_bounds__(kScanBlock) void tree_scan_fix_kernel(uint32_t* __restrict__ scan_tmp,
                                                                   uint32_t* __restrict__ ws,
                                                                   int* info, int level,
                                                                   int node_cap, int* err)
{
    const int cnt = level_count(info)[level];
    size_t n = static_cast<size_t>(cnt > 0 ? cnt : 0);
    const size_t cap = static_cast<size_t>(gridDim.x) * kScanChunk;
    n = n < cap ? n : cap;
    const ScanBlockResult r = scan_lookback_block(scan_tmp, n, ws, err, kErrScan);
    if (r.last && threadIdx.x == 0)
        level_fix(info, level, static_cast<int>(r.total), node_cap, err);
}

// Children of the nodes of `level` from the bounds tree_count_kernel left in
// bnd_tmp (one thread per node: no key reads).
__global__ __launch_bounds__(256) void tree_emit_kernel(
    uint64_t* __restrict__ nprefix, uint8_t* __restrict__ nlevel, uint32_t* __restrict__ nstart,
    uint32_t* __restrict__ ncount, uint32_t* __restrict__ nchild, uint32_t* __restrict__ nfirst,
    uint32_t* __restrict__ nparent, const uint32_t* __restrict__ scan_tmp,
    const uint32_t* __restrict__ bnd_tmp, const int* __restrict__ info, int level)
{
    const int base = info[level];
    const int cnt = info[kLevels + level];
    const int next_base = info[level + 1];
    const bool dropped = info[kLevels + level + 1] == 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int node = base + k;
        const uint32_t nc = nchild[node];
        if (nc == 0)
            continue;
        if (dropped) {
            nchild[node] = 0;
            continue;
        }
        const uint32_t first = static_cast<uint32_t>(next_base) + scan_tmp[k];
        const uint4* in = reinterpret_cast<const uint4*>(bnd_tmp + static_cast<size_t>(k) * 8);
        const uint4 b0 = in[0], b1 = in[1];
        const uint32_t bnd[9] = {nstart[node], b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        nfirst[node] = first;
        const unsigned long long p = nprefix[node];
        uint32_t j = 0;
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            if (bnd[o + 1] <= bnd[o])
                continue;
            const uint32_t child = first + j++;
            nprefix[child] = (p << 3) | static_cast<unsigned long long>(o);
            nlevel[child] = static_cast<uint8_t>(level + 1);
            nstart[child] = bnd[o];
            ncount[child] = bnd[o + 1] - bnd[o];
            nchild[child] = 0;
            nfirst[child] = bnd[o];
            nparent[child] = static_cast<uint32_t>(node);
        }
    }
}

// One-pass topology. Every node of the tree is a run of equal level-L key
// prefixes, so it is named by its level L and its first particle s. With d_p
// the number of leading 3-bit digits that sorted keys p - 1 and p share
// (0..21), particle p starts a run at level L exactly when d_p < L (particle
// 0 starts every run). A run at level l holding more than B particles
// contains B + 1 consecutive particles whose B inner boundaries all share at
// least l digits. So with
//     w_j = min(d_{j+1}, ..., d_{j+B})   over windows inside [0, n)
//     W_i = max(w_j : i - B <= j <= i)   (-1 when no window holds i)
// the run at level l that holds particle i has more than B particles exactly
// when W_i >= l. A node exists when every ancestor split; a node at level L
// splits when L < 21 and (L < D or it holds more than B particles). Hence the
// nodes that start at particle s are those of levels
//     lo_s = d_s + 1 (0 for s = 0)  ...  hi_s = min(21, max(D, W_s + 1)),
// the one at hi_s a leaf and each other the parent of the next. Level L's
// nodes in breadth-first order are the particles with lo_s <= L <= hi_s in
// increasing order, so a node's index is its level's base plus its rank
// among them. Passes: per 4,096-particle tile, each particle's [lo, hi] (one
// word) and the tile's node count per level; a scan of the counts per level;
// per tile, the nodes in place; then, per node, what a tile cannot see
// (parents of later siblings, child counts, counts of internal nodes and of
// leaves longer than B, each particle's leaf).
constexpr int kTopoThreads = 256;
constexpr int kTopoPer = kTopoTile / kTopoThreads;  // particles per thread (consecutive in the emit pass)
constexpr uint32_t kPending = 0xfffffffeu;
constexpr unsigned long long kKeyMask = 0x7fffffffffffffffull;

__device__ inline int shared_digits(unsigned long long a, unsigned long long b)
{
    const unsigned long long x = (a ^ b) & kKeyMask;
    return x == 0ull ? kMaxLevel : (__clzll(x) - 1) / 3;
}

struct TopoTile {
    int8_t d[kTopoTile + 2 * kTopoHalo];  // d at particle t0 - H + m (-1 outside [1, n))
    int8_t w[kTopoTile + kTopoHalo];      // w at window start t0 - H + m
    uint32_t cnt[kLevels][kTopoThreads];  // nodes per level and thread
};

// Per-particle level range [lo, hi] and, when the node at hi holds at most B
// particles, its count c (else 0): lo | (hi + 1) << 8 | c << 16.
__device__ inline uint32_t topo_pack(int lo, int hi, int c)
{
    return static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi + 1) << 8) |
           (static_cast<uint32_t>(c) << 16);
}
__device__ inline int topo_lo(uint32_t v) { return static_cast<int>(v & 0xffu); }
__device__ inline int topo_hi(uint32_t v) { return static_cast<int>((v >> 8) & 0xffu) - 1; }
__device__ inline int topo_c(uint32_t v) { return static_cast<int>((v >> 16) & 0xffu); }

// A thread's node count per level, in registers: 5-bit fields (at most 16
// particles per thread), levels 0-10 in c[0] and 11-21 in c[1].
struct LevelCounts {
    unsigned long long c[2] = {0ull, 0ull};

    // fields a..b (0 <= a <= b <= 10) of one word set to 1
    __device__ static unsigned long long fields(int a, int b)
    {
        constexpr unsigned long long kOnes = 0x4210842108421ull;  // bit 5k, k = 0..10
        return kOnes & ((1ull << (5 * (b + 1))) - 1ull) & ~((1ull << (5 * a)) - 1ull);
    }
    __device__ void add(int lo, int hi)
    {
        if (lo > hi)
            return;
        if (lo <= 10)
            c[0] += fields(lo, min(hi, 10));
        if (hi >= 11)
            c[1] += fields(max(lo, 11) - 11, hi - 11);
    }
    __device__ uint32_t get(int L) const
    {
        return static_cast<uint32_t>((L <= 10 ? c[0] >> (5 * L) : c[1] >> (5 * (L - 11))) & 31u);
    }
};

__device__ inline uint32_t wave_total(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Per tile: the level range of every particle (pinfo) and the nodes per
// level (tilecnt[L * ntiles + tile]). BT > 0: the bucket as a compile-time
// constant (the window loops unroll and their LDS reads overlap; one read at
// a time they set the kernel's time); BT = 0: the bucket Brt at run time.
template <int BT>
__global__ __launch_bounds__(kTopoThreads) void topo_count_kernel(
    const unsigned long long* __restrict__ keys, uint32_t n, int Brt, int D,
    uint32_t* __restrict__ pinfo, uint32_t* __restrict__ tilecnt, uint32_t ntiles)
{
    const int B = BT > 0 ? BT : Brt;
    __shared__ TopoTile S;
    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const uint32_t t0 = tile * kTopoTile;
    // every key load of the thread in flight at once, then the digits
    constexpr int kIters = (kTopoTile + 2 * kTopoHalo + kTopoThreads - 1) / kTopoThreads;
    unsigned long long ka[kIters], kb[kIters];
#pragma unroll
    for (int k = 0; k < kIters; ++k) {
        const int m = tid + k * kTopoThreads;
        const long long p = static_cast<long long>(t0) - kTopoHalo + m;
        const bool in = m < kTopoTile + 2 * kTopoHalo && p >= 1 && p < static_cast<long long>(n);
        ka[k] = in ? keys[p - 1] : 0ull;
        kb[k] = in ? keys[p] : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kIters; ++k) {
        const int m = tid + k * kTopoThreads;
        const long long p = static_cast<long long>(t0) - kTopoHalo + m;
        if (m < kTopoTile + 2 * kTopoHalo)
            S.d[m] = static_cast<int8_t>(p >= 1 && p < static_cast<long long>(n)
                                             ? shared_digits(ka[k], kb[k])
                                             : -1);
    }
    __syncthreads();
    for (int m = tid; m < kTopoTile + kTopoHalo; m += kTopoThreads) {
        int v = kMaxLevel;
#pragma unroll
        for (int b = 1; b <= (BT > 0 ? BT : kTopoHalo); ++b)  // m + B < kTopoTile + 2 * kTopoHalo
            if (BT > 0 || b <= B)
                v = min(v, static_cast<int>(S.d[m + b]));
        S.w[m] = static_cast<int8_t>(v);
    }
    __syncthreads();
    // particles interleaved over the threads (neighbouring lanes read
    // neighbouring bytes of d and w, and store neighbouring words)
    LevelCounts lc;
#pragma unroll 4
    for (int q = 0; q < kTopoPer; ++q) {
        const int r = q * kTopoThreads + tid;
        const uint32_t i = t0 + static_cast<uint32_t>(r);
        int lo = kLevels, hi = -1, c = 0;
        if (i < n) {
            int W = -1;
            if (BT > 0) {
#pragma unroll
                for (int b = 0; b <= BT; ++b)  // r + H - b >= 0
                    W = max(W, static_cast<int>(S.w[r + kTopoHalo - b]));
            } else {
                for (int b = 0; b <= B; ++b)
                    W = max(W, static_cast<int>(S.w[r + kTopoHalo - b]));
            }
            lo = i == 0 ? 0 : static_cast<int>(S.d[r + kTopoHalo]) + 1;
            hi = min(kMaxLevel, max(D, W + 1));
            // the leaf at hi ends at the next run start within B, if any
            // (d is -1 at and beyond n): the smallest such b
            if (lo <= hi) {
                if (BT > 0) {
#pragma unroll
                    for (int b = BT; b >= 1; --b)
                        if (S.d[r + kTopoHalo + b] < hi)
                            c = b;
                } else {
                    for (int b = 1; b <= B; ++b) {
                        if (S.d[r + kTopoHalo + b] < hi) {
                            c = b;
                            break;
                        }
                    }
                }
            }
        }
        if (i < n)
            pinfo[i] = topo_pack(lo, hi, c);
        lc.add(lo, hi);
    }
#pragma unroll
    for (int L = 0; L < kLevels; ++L)
        S.cnt[L][tid] = lc.get(L);
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    for (int L = wv; L < kLevels; L += kTopoThreads / 64) {
        const uint32_t v = wave_total(S.cnt[L][lane] + S.cnt[L][lane + 64] +
                                      S.cnt[L][lane + 128] + S.cnt[L][lane + 192]);
        if (lane == 0)
            tilecnt[static_cast<size_t>(L) * ntiles + tile] = v;
    }
}

// Exclusive scan of each level's tile counts (one block per level, in place);
// the level totals go to tilecnt[kLevels * ntiles + L].
__global__ __launch_bounds__(kScanBlock) void topo_scan_kernel(uint32_t* __restrict__ tilecnt,
                                                               uint32_t ntiles)
{
    __shared__ uint32_t tmp[kScanWaves];
    const int L = blockIdx.x;
    uint32_t* col = tilecnt + static_cast<size_t>(L) * ntiles;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < ntiles; b += kScanBlock) {
        const uint32_t t = b + threadIdx.x;
        const uint32_t v = t < ntiles ? col[t] : 0u;
        uint32_t total = 0;
        const uint32_t ex = block_exclusive256(v, tmp, &total);
        if (t < ntiles)
            col[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0)
        tilecnt[static_cast<size_t>(kLevels) * ntiles + L] = carry;
}

// The nodes that start in one tile, in breadth-first place, level by level
// (a thread's nodes of one level take consecutive slots, so neighbouring
// lanes write neighbouring slots). The first level whose nodes would not fit
// in node_cap is dropped with every deeper one (the nodes above it that
// would split stay leaves; flagged). Block 0 also writes the level table.
__global__ __launch_bounds__(kTopoThreads) void topo_emit_kernel(
    const unsigned long long* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ pinfo,
    const uint32_t* __restrict__ tilecnt, uint32_t ntiles, int node_cap,
    uint64_t* __restrict__ nprefix, uint8_t* __restrict__ nlevel, uint32_t* __restrict__ nstart,
    uint32_t* __restrict__ ncount, uint32_t* __restrict__ nchild, uint32_t* __restrict__ nfirst,
    uint32_t* __restrict__ nparent, int* __restrict__ info, int* __restrict__ err)
{
    __shared__ uint32_t cnt[kLevels][kTopoThreads];
    __shared__ uint32_t sbase[kLevels];  // level base + this tile's offset
    __shared__ int slmax;
    const int tid = threadIdx.x;
    const uint32_t tile = blockIdx.x;
    const uint32_t t0 = tile * kTopoTile;
    const uint32_t i0 = t0 + static_cast<uint32_t>(tid * kTopoPer);
    const uint32_t* tot = tilecnt + static_cast<size_t>(kLevels) * ntiles;
    // the thread's level ranges and keys, loaded together
    uint32_t pk[kTopoPer];
    unsigned long long kq[kTopoPer];
    if (i0 + kTopoPer <= n) {
        const uint4* in4 = reinterpret_cast<const uint4*>(pinfo + i0);
        const ulonglong2* k2 = reinterpret_cast<const ulonglong2*>(keys + i0);
#pragma unroll
        for (int k = 0; k < kTopoPer / 4; ++k) {
            const uint4 v = in4[k];
            pk[4 * k] = v.x;
            pk[4 * k + 1] = v.y;
            pk[4 * k + 2] = v.z;
            pk[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int k = 0; k < kTopoPer / 2; ++k) {
            const ulonglong2 v = k2[k];
            kq[2 * k] = v.x & kKeyMask;
            kq[2 * k + 1] = v.y & kKeyMask;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kTopoPer; ++q) {
            const bool in = i0 + q < n;
            pk[q] = in ? pinfo[i0 + q] : topo_pack(kLevels, -1, 0);
            kq[q] = in ? keys[i0 + q] & kKeyMask : 0ull;
        }
    }
    if (tid == 0)
        slmax = kMaxLevel;
    __syncthreads();
    uint64_t below = 0;  // nodes of the levels above level tid
    if (tid < kLevels) {
        for (int l = 0; l < tid; ++l)
            below += tot[l];
        if (below + tot[tid] > static_cast<uint64_t>(node_cap))
            atomicMin(&slmax, tid - 1);  // level 0 (the root) always fits
    }
    __syncthreads();
    const int lmax = slmax;
    if (tid < kLevels) {
        const uint32_t base = static_cast<uint32_t>(below);
        sbase[tid] = base + (tid <= lmax ? tilecnt[static_cast<size_t>(tid) * ntiles + tile] : 0u);
        if (tile == 0) {
            const bool kept = tid <= lmax;
            uint64_t kept_total = 0;
            for (int l = 0; l <= lmax; ++l)
                kept_total += tot[l];
            info[tid] = kept ? static_cast<int>(base) : static_cast<int>(kept_total);
            info[kLevels + tid] = kept ? static_cast<int>(tot[tid]) : 0;
            if (tid == 0) {
                int deepest = 0;
                for (int l = 0; l <= lmax; ++l)
                    if (tot[l] > 0u)
                        deepest = l;
                info[kInfoTotal] = static_cast<int>(kept_total);
                info[kInfoMaxLevel] = deepest;
                info[kInfoBig] = 0;
                if (lmax < kMaxLevel && tot[lmax + 1] > 0u)
                    atomicOr(err, kErrNodeOverflow);
            }
        }
    }
    // levels past lmax dropped: the node there becomes a leaf of unknown count
    int lmin_t = kLevels, lmax_t = -1;
#pragma unroll
    for (int q = 0; q < kTopoPer; ++q) {
        const int lo = topo_lo(pk[q]), hi = topo_hi(pk[q]);
        if (hi > lmax)
            pk[q] = topo_pack(lo, lmax, 0);
        if (lo <= min(hi, lmax)) {
            lmin_t = min(lmin_t, lo);
            lmax_t = max(lmax_t, min(hi, lmax));
        }
    }
    LevelCounts lc;
#pragma unroll
    for (int q = 0; q < kTopoPer; ++q)
        lc.add(topo_lo(pk[q]), topo_hi(pk[q]));
#pragma unroll
    for (int L = 0; L < kLevels; ++L)
        cnt[L][tid] = lc.get(L);
    __syncthreads();
    // per level: exclusive scan of the thread counts, in place
    const int lane = tid & 63, wv = tid >> 6;
    for (int L = wv; L < kLevels; L += kTopoThreads / 64) {
        const uint32_t a0 = cnt[L][4 * lane], a1 = cnt[L][4 * lane + 1];
        const uint32_t a2 = cnt[L][4 * lane + 2], a3 = cnt[L][4 * lane + 3];
        const uint32_t sum = a0 + a1 + a2 + a3;
        const uint32_t ex = wave_inclusive(sum) - sum;
        cnt[L][4 * lane] = ex;
        cnt[L][4 * lane + 1] = ex + a0;
        cnt[L][4 * lane + 2] = ex + a0 + a1;
        cnt[L][4 * lane + 3] = ex + a0 + a1 + a2;
    }
    __syncthreads();
    uint32_t prev[kTopoPer];  // particle q's node at the level before
#pragma unroll
    for (int q = 0; q < kTopoPer; ++q)
        prev[q] = 0xffffffffu;
    for (int L = lmin_t; L <= lmax_t; ++L) {
        uint32_t idx = sbase[L] + cnt[L][tid];
        const int sh = 3 * (kMaxLevel - L);
#pragma unroll
        for (int q = 0; q < kTopoPer; ++q) {
            const int lo = topo_lo(pk[q]), hi = topo_hi(pk[q]);
            if (L < lo || L > hi)
                continue;
            const uint32_t i = i0 + static_cast<uint32_t>(q);
            nprefix[idx] = L == 0 ? 0ull : kq[q] >> sh;
            nlevel[idx] = static_cast<uint8_t>(L);
            nstart[idx] = i;
            if (L < hi) {  // nfirst: written by the first child, next level
                nchild[idx] = kPending;
                ncount[idx] = kPending;
            } else {
                nfirst[idx] = i;
                nchild[idx] = 0u;
                const int c = topo_c(pk[q]);
                ncount[idx] = c > 0 ? static_cast<uint32_t>(c) : kPending;
            }
            if (L > lo) {  // the first child of the node above
                nparent[idx] = prev[q];
                nfirst[prev[q]] = idx;
            } else {
                nparent[idx] = L == 0 ? 0xffffffffu : kPending;
            }
            prev[q] = idx;
            ++idx;
        }
    }
}

// First particle past s whose level-L prefix differs from s's (or n):
// galloping, then bisection.
__device__ inline uint32_t run_end(const unsigned long long* __restrict__ keys, uint32_t n,
                                   uint32_t s, int L)
{
    const int sh = 3 * (kMaxLevel - L);
    const unsigned long long k0 = (keys[s] & kKeyMask) >> sh;
    uint32_t in = s, out = n;
    for (uint64_t step = 1;; step <<= 1) {
        const uint64_t p = static_cast<uint64_t>(in) + step;
        if (p >= n)
            break;
        if (((keys[p] & kKeyMask) >> sh) != k0) {
            out = static_cast<uint32_t>(p);
            break;
        }
        in = static_cast<uint32_t>(p);
    }
    while (out - in > 1u) {
        const uint32_t mid = in + (out - in) / 2u;
        if (((keys[mid] & kKeyMask) >> sh) != k0)
            out = mid;
        else
            in = mid;
    }
    return out;
}

// Number of children of an internal node (level L, prefix p) whose first
// child is f: the run of level-(L + 1) nodes from f with parent prefix p.
__device__ inline uint32_t topo_children(const uint64_t* __restrict__ nprefix,
                                         const int* __restrict__ info, int L,
                                         unsigned long long p, uint32_t f)
{
    const uint32_t end =
        static_cast<uint32_t>(info[L + 1]) + static_cast<uint32_t>(info[kLevels + L + 1]);
    // the next 7 slots' prefixes, loaded together (in bounds: clamped to f)
    unsigned long long q[7];
#pragma unroll
    for (int k = 0; k < 7; ++k)
        q[k] = nprefix[f + k + 1u < end ? f + k + 1u : f];
    uint32_t m = 1;
    bool run = true;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        run = run && f + k + 1u < end && (q[k] >> 3) == p;
        m += run ? 1u : 0u;
    }
    return m;
}

// What the tiles left pending, one thread per node. Only pending fields are
// written; every field this pass reads is final or is read only to tell a
// leaf (nchild 0) from an internal node (pending or final, never 0).
__global__ __launch_bounds__(256) void topo_link_kernel(
    const unsigned long long* __restrict__ keys, uint32_t n, const uint64_t* __restrict__ nprefix,
    const uint8_t* __restrict__ nlevel, const uint32_t* __restrict__ nstart,
    uint32_t* __restrict__ ncount, uint32_t* __restrict__ nchild,
    const uint32_t* __restrict__ nfirst, uint32_t* __restrict__ nparent,
    uint32_t* __restrict__ sleaf, const int* __restrict__ info, int* __restrict__ err)
{
    const int total = info[kInfoTotal];
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < total; x += gridDim.x * blockDim.x) {
        const int L = nlevel[x];
        const unsigned long long px = nprefix[x];
        if (nparent[x] == kPending) {
            // a later sibling: its parent is that of the first sibling (the
            // parent's first child, linked by the emit pass), at most 7 back
            const int lb = info[L];
            // the 7 slots before, loaded together (in bounds: clamped to x)
            unsigned long long q[7];
#pragma unroll
            for (int k = 0; k < 7; ++k)
                q[k] = nprefix[x - k - 1 >= lb ? x - k - 1 : x];
            int y = x;
            bool run = true;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                run = run && x - k - 1 >= lb && (q[k] >> 3) == (px >> 3);
                y -= run ? 1 : 0;
            }
            const uint32_t par = nparent[y];
            if (y == x || par == kPending)
                atomicOr(err, kErrTopology);
            else
                nparent[x] = par;
        }
        if (nchild[x] == kPending) {
            const uint32_t f = nfirst[x];
            if (L >= kMaxLevel || f >= static_cast<uint32_t>(total)) {
                atomicOr(err, kErrTopology);
                continue;
            }
            nchild[x] = topo_children(nprefix, info, L, px, f);
        }
        if (ncount[x] == kPending) {
            // the run ends where its last leaf ends: descend by last children
            uint32_t cur = static_cast<uint32_t>(x);
            int lc = L;
            unsigned long long pc = px;
            bool ok = true;
            for (int g = 0; g < kLevels && nchild[cur] != 0u; ++g) {
                const uint32_t f = nfirst[cur];
                if (lc >= kMaxLevel || f >= static_cast<uint32_t>(total)) {
                    ok = false;
                    break;
                }
                cur = f + topo_children(nprefix, info, lc, pc, f) - 1u;
                ++lc;
                pc = nprefix[cur];
            }
            if (!ok || nchild[cur] != 0u) {
                atomicOr(err, kErrTopology);
                continue;
            }
            const uint32_t cs = nstart[cur];
            const uint32_t cc = ncount[cur];
            const uint32_t e = cc != kPending ? cs + cc : run_end(keys, n, cs, lc);
            ncount[x] = e - nstart[x];
        }
        if (nchild[x] == 0u) {  // a leaf: its particles point to it
            const uint32_t s0 = nstart[x], e = min(s0 + ncount[x], n);
            for (uint32_t q = s0; q < e; ++q)
                sleaf[q] = static_cast<uint32_t>(x);
        }
    }
}

// Third moments about the COM, T_ijk = sum m y_i y_j y_k (10 components:
// xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz), accumulated in registers.
struct Oct {
    float t[10];
};

__device__ inline void oct_zero(Oct& o)
{
#pragma unroll
    for (int k = 0; k < 10; ++k)
        o.t[k] = 0.0f;
}

// o += m y y y (one particle at offset y from the COM)
__device__ inline void oct_add_point(Oct& o, float m, float x, float y, float z)
{
    const float mx = m * x, my = m * y, mz = m * z;
    o.t[0] += mx * x * x;
    o.t[1] += mx * x * y;
    o.t[2] += mx * x * z;
    o.t[3] += mx * y * y;
    o.t[4] += mx * y * z;
    o.t[5] += mx * z * z;
    o.t[6] += my * y * y;
    o.t[7] += my * y * z;
    o.t[8] += my * z * z;
    o.t[9] += mz * z * z;
}

// o += the third moment of a child about the parent COM (parallel-axis
// shift): T_c + (D_i S_jk + D_j S_ik + D_k S_ij) + M_c D_i D_j D_k, with D
// the child COM minus the parent COM and S = (Q + trs I) / 3 the child's
// second moment about its own COM (Q traceless).
__device__ inline void oct_add_child(Oct& o, const float* tc, const WalkNode& c, float dx,
                                     float dy, float dz)
{
    const float third = 1.0f / 3.0f;
    const float sxx = (c.qxx + c.trs) * third, syy = (c.qyy + c.trs) * third;
    const float szz = (-c.qxx - c.qyy + c.trs) * third;
    const float sxy = c.qxy * third, sxz = c.qxz * third, syz = c.qyz * third;
    const float D[3] = {dx, dy, dz};
    const float S[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}};
    const int I[10] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 2};
    const int J[10] = {0, 0, 0, 1, 1, 2, 1, 1, 2, 2};
    const int K[10] = {0, 1, 2, 1, 2, 2, 1, 2, 2, 2};
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const int i = I[k], j = J[k], l = K[k];
        o.t[k] += tc[k] + D[i] * S[j][l] + D[j] * S[i][l] + D[l] * S[i][j] +
                  c.mass * D[i] * D[j] * D[l];
    }
}

// |d|^2 with explicit roundings: the moments come out the same bits in every
// kernel that computes them, whatever the compiler fuses around them.
__device__ inline float norm2_rn(float x, float y, float z)
{
    return __fmaf_rn(z, z, __fmaf_rn(y, y, __fmul_rn(x, x)));
}

// FP64 lattice centre of a cell: the origin of its particles' FP32 offsets.
__device__ inline void cell_origin(const int* c, const RootGeom* __restrict__ geom, double* o)
{
    o[0] = __fma_rn(static_cast<double>(c[0]), geom->unit_d[0], geom->lo_d[0]);
    o[1] = __fma_rn(static_cast<double>(c[1]), geom->unit_d[1], geom->lo_d[1]);
    o[2] = __fma_rn(static_cast<double>(c[2]), geom->unit_d[2], geom->lo_d[2]);
}

// Moments of one node. A leaf reads its particles' offsets from its lattice
// centre through part(s, origin) (particle s, as stored in spart); an internal
// node combines its children, which must be complete. OCT: also the third
// moments (noct, 10 per node; read by the learned far-field correction).
struct NodeMeta {
    unsigned long long prefix;
    uint32_t nchild, first, count;
};

// The caller loads the node's fields (all in one round trip).
__device__ inline NodeMeta node_meta(int node, const uint64_t* __restrict__ nprefix,
                                     const uint32_t* __restrict__ ncount,
                                     const uint32_t* __restrict__ nchild,
                                     const uint32_t* __restrict__ nfirst)
{
    return NodeMeta{nprefix[node], nchild[node], nfirst[node], ncount[node]};
}

template <bool OCT, bool LEAF, class Part>
__device__ inline void node_moments(int node, int level, const NodeMeta& meta,
                                    const RootGeom* __restrict__ geom, WalkNode* __restrict__ wn,
                                    float* __restrict__ ext, float* __restrict__ noct, int refit,
                                    int* err, Part part)
{
    const float ux = geom->unit_f[0], uy = geom->unit_f[1], uz = geom->unit_f[2];
    {
        int c[3];
        lattice_centre(meta.prefix, level, c);
        const uint32_t nc = LEAF ? 0u : meta.nchild;
        const uint32_t first = meta.first;
        float M = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        float qxx = 0.0f, qxy = 0.0f, qxz = 0.0f, qyy = 0.0f, qyz = 0.0f, bmax = 0.0f, trs = 0.0f;
        // e (refit only): half-width, in lattice units along the widest
        // axis, of the cube about the lattice centre that holds the subtree's
        // particles (the cell's own half-width 2^(21 - level) unless they
        // left the cell)
        float e = 0.0f;
        uint32_t count = 0;
        Oct oc;
        if (OCT)
            oct_zero(oc);
        if (LEAF) {
            count = meta.count;
            double o[3];
            cell_origin(c, geom, o);
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (uint32_t s = first; s < first + count; ++s) {
                const float4 v = part(s, o);
                const float x = v.x, y = v.y, z = v.z, m = v.w;
                if (refit)  // the extent matters only to a refit
                    e = fmaxf(e, fmaxf(fabsf(x) / ux, fmaxf(fabsf(y) / uy, fabsf(z) / uz)));
                M = __fadd_rn(M, m);
                sx = __fmaf_rn(m, x, sx);
                sy = __fmaf_rn(m, y, sy);
                sz = __fmaf_rn(m, z, sz);
            }
            if (M > 0.0f) {
                cx = sx / M;
                cy = sy / M;
                cz = sz / M;
            }
            for (uint32_t s = first; s < first + count; ++s) {
                const float4 p = part(s, o);
                const float dx = __fsub_rn(p.x, cx), dy = __fsub_rn(p.y, cy),
                            dz = __fsub_rn(p.z, cz);
                const float r2 = norm2_rn(dx, dy, dz);
                const float tx = __fmul_rn(3.0f, dx), ty = __fmul_rn(3.0f, dy);
                qxx = __fmaf_rn(p.w, __fmaf_rn(tx, dx, -r2), qxx);
                qxy = __fmaf_rn(p.w, __fmul_rn(tx, dy), qxy);
                qxz = __fmaf_rn(p.w, __fmul_rn(tx, dz), qxz);
                qyy = __fmaf_rn(p.w, __fmaf_rn(ty, dy, -r2), qyy);
                qyz = __fmaf_rn(p.w, __fmul_rn(ty, dz), qyz);
                trs = __fmaf_rn(p.w, r2, trs);
                bmax = fmaxf(bmax, r2);
                if (OCT)
                    oct_add_point(oc, p.w, dx, dy, dz);
            }
            bmax = sqrtf(bmax);
        } else {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            // The (at most 8) children unrolled, each load in bounds (index
            // clamped to the last child, whose repeats weigh 0), so that the
            // loads issue together instead of one round trip per child.
#pragma unroll
            for (uint32_t jj = 0; jj < 8; ++jj) {
                const WalkNode ch = wn[first + (jj < nc ? jj : nc - 1u)];
                const float cm = jj < nc ? ch.mass : 0.0f;
                const float rx = __fmaf_rn(static_cast<float>(ch.cx - c[0]), ux, ch.comx);
                const float ry = __fmaf_rn(static_cast<float>(ch.cy - c[1]), uy, ch.comy);
                const float rz = __fmaf_rn(static_cast<float>(ch.cz - c[2]), uz, ch.comz);
                M = __fadd_rn(M, cm);
                sx = __fmaf_rn(cm, rx, sx);
                sy = __fmaf_rn(cm, ry, sy);
                sz = __fmaf_rn(cm, rz, sz);
            }
            if (M > 0.0f) {
                cx = sx / M;
                cy = sy / M;
                cz = sz / M;
            }
#pragma unroll
            for (uint32_t jj = 0; jj < 8; ++jj) {
                const bool live = jj < nc;
                const WalkNode ch = wn[first + (live ? jj : nc - 1u)];
                const float dx =
                    __fsub_rn(__fmaf_rn(static_cast<float>(ch.cx - c[0]), ux, ch.comx), cx);
                const float dy =
                    __fsub_rn(__fmaf_rn(static_cast<float>(ch.cy - c[1]), uy, ch.comy), cy);
                const float dz =
                    __fsub_rn(__fmaf_rn(static_cast<float>(ch.cz - c[2]), uz, ch.comz), cz);
                const float r2 = norm2_rn(dx, dy, dz);
                const float tx = __fmul_rn(3.0f, dx), ty = __fmul_rn(3.0f, dy);
                const float m = ch.mass;
                qxx = __fadd_rn(qxx, live ? __fmaf_rn(m, __fmaf_rn(tx, dx, -r2), ch.qxx) : 0.0f);
                qxy = __fadd_rn(qxy, live ? __fmaf_rn(m, __fmul_rn(tx, dy), ch.qxy) : 0.0f);
                qxz = __fadd_rn(qxz, live ? __fmaf_rn(m, __fmul_rn(tx, dz), ch.qxz) : 0.0f);
                qyy = __fadd_rn(qyy, live ? __fmaf_rn(m, __fmaf_rn(ty, dy, -r2), ch.qyy) : 0.0f);
                qyz = __fadd_rn(qyz, live ? __fmaf_rn(m, __fmul_rn(ty, dz), ch.qyz) : 0.0f);
                trs = __fadd_rn(trs, live ? __fmaf_rn(m, r2, ch.trs) : 0.0f);
                if (live && m > 0.0f)
                    bmax = fmaxf(bmax, __fadd_rn(sqrtf(r2), ch.bmax));
                if (live) {
                    if (refit) {
                        const int off =
                            max(abs(ch.cx - c[0]), max(abs(ch.cy - c[1]), abs(ch.cz - c[2])));
                        e = fmaxf(e, ext[first + jj] + static_cast<float>(off));
                    }
                    if (OCT)
                        oct_add_child(oc, noct + static_cast<size_t>(first + jj) * 10, ch, dx, dy,
                                      dz);
                }
            }
        }
        if (refit)
            ext[node] = e;
        if (OCT) {
            float* o = noct + static_cast<size_t>(node) * 10;
#pragma unroll
            for (int k = 0; k < 10; ++k)
                o[k] = oc.t[k];
        }
        // Refit (moments of the current positions on an earlier topology):
        // particles may have left their cells. The moments and bmax are
        // exact for the current positions, and a target inside a node's
        // bounding sphere opens it whatever its size. The level the walk
        // derives the cell size from is lowered only once the particles
        // reach beyond 1.5 times the cell's half-width. (The bound e of a
        // parent exceeds its half-width by half its child's excess: lowering
        // at any excess doubled the size of every ancestor of one escaped
        // particle, up to the root, and the 8M block-step walk took 1.3-4x
        // longer than after a rebuild.) A fresh build keeps the true level.
        int wlevel = level;
        if (refit) {
            while (wlevel > 0 && 1.5f * ldexpf(1.0f, kMaxLevel - wlevel) < e)
                --wlevel;
        }
        WalkNode w;
        w.comx = cx;
        w.comy = cy;
        w.comz = cz;
        w.mass = M;
        w.qxx = qxx;
        w.qxy = qxy;
        w.qxz = qxz;
        w.qyy = qyy;
        w.qyz = qyz;
        w.bmax = bmax;
        w.trs = trs;
        w.first = static_cast<int>(first);
        w.cx = c[0];
        w.cy = c[1];
        w.cz = c[2];
        if (nc != 0) {
            // internal node: the subtree's particle count, saturated (the walk
            // reads it only to pick the near-field treatment)
            count = min(meta.count, kMaxLeafCount);
        } else if (count > kMaxLeafCount) {
            // The walk then sees a truncated leaf: flag it; every particle was
            // still packed above, so no access goes out of bounds.
            atomicOr(err, kErrLeafTooLarge);
            count = kMaxLeafCount;
        }
        w.meta = make_meta(nc, static_cast<uint32_t>(wlevel), count);
        wn[node] = w;
    }
}

// Leaves of every level. Particles in chunks of kLeafChunk (kLeafPer per
// thread, coalesced): each stores its offset from its leaf's lattice centre
// (spart) and keeps it in LDS. The chunk's leaf starts are then listed and
// summed by consecutive threads (full wavefronts, not one lane in six),
// reading the chunk's offsets from LDS and recomputing those past the chunk
// (the same arithmetic, so the same bits). sleaf must be set.
constexpr int kLeafThreads = 256;
constexpr int kLeafPer = 4;
constexpr int kLeafChunk = kLeafThreads * kLeafPer;

template <bool OCT>
__global__ __launch_bounds__(kLeafThreads) void leaf_moments_kernel(
    const uint64_t* __restrict__ nprefix, const uint8_t* __restrict__ nlevel,
    const uint32_t* __restrict__ nstart, const uint32_t* __restrict__ ncount,
    const uint32_t* __restrict__ nchild, const uint32_t* __restrict__ nfirst,
    const uint32_t* __restrict__ sleaf, uint32_t n, const int* __restrict__ info,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const float* __restrict__ pm, const RootGeom* __restrict__ geom, WalkNode* __restrict__ wn,
    float4* __restrict__ spart, float* __restrict__ ext, float* __restrict__ noct, int refit,
    int* err)
{
    __shared__ float4 sp[kLeafChunk];
    __shared__ uint32_t starts[kLeafChunk];  // chunk-relative first particle of each leaf
    __shared__ uint32_t nstarts;
    const uint32_t b0 = blockIdx.x * kLeafChunk;
    const uint32_t total = static_cast<uint32_t>(info[kInfoTotal]);
    if (threadIdx.x == 0)
        nstarts = 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kLeafPer; ++k) {
        const uint32_t r = static_cast<uint32_t>(k * kLeafThreads) + threadIdx.x;
        const uint32_t p = b0 + r;
        bool first = false;
        if (p < n) {
            const uint32_t leaf = sleaf[p];
            if (leaf < total) {
                int c[3];
                lattice_centre(nprefix[leaf], nlevel[leaf], c);
                double o[3];
                cell_origin(c, geom, o);
                const float4 v = make_float4(
                    static_cast<float>(px[p] - o[0]), static_cast<float>(py[p] - o[1]),
                    static_cast<float>(pz[p] - o[2]), pm[p]);
                spart[p] = v;
                sp[r] = v;
                first = nstart[leaf] == p;
            } else {
                atomicOr(err, kErrTopology);
            }
        }
        // list the leaf starts (order within the list is free)
        const unsigned long long mask = __ballot(first);
        const int lane = threadIdx.x & 63;
        uint32_t base = 0;
        if (lane == 0 && mask)
            base = atomicAdd(&nstarts, static_cast<uint32_t>(__popcll(mask)));
        base = __shfl(base, 0, 64);
        if (first)
            starts[base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)))] = r;
    }
    __syncthreads();
    const uint32_t ns = nstarts;
    const uint32_t lim = b0 + kLeafChunk;
    for (uint32_t t = threadIdx.x; t < ns; t += kLeafThreads) {
        const uint32_t p = b0 + starts[t];
        const uint32_t leaf = sleaf[p];
        const NodeMeta meta = node_meta(static_cast<int>(leaf), nprefix, ncount, nchild, nfirst);
        node_moments<OCT, true>(
            static_cast<int>(leaf), nlevel[leaf], meta, geom, wn, ext, noct, refit, err,
            [&](uint32_t s, const double* o) {
                if (s < lim)
                    return sp[s - b0];
                return make_float4(static_cast<float>(px[s] - o[0]),
                                   static_cast<float>(py[s] - o[1]),
                                   static_cast<float>(pz[s] - o[2]), pm[s]);
            });
    }
}

// sleaf[s] = the leaf holding particle s (topology built level by level).
__global__ __launch_bounds__(256) void leaf_fill_kernel(const uint32_t* __restrict__ nstart,
                                                        const uint32_t* __restrict__ ncount,
                                                        const uint32_t* __restrict__ nchild,
                                                        const int* __restrict__ info, uint32_t n,
                                                        uint32_t* __restrict__ sleaf)
{
    const int total = info[kInfoTotal];
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < total; x += gridDim.x * blockDim.x) {
        if (nchild[x] != 0u)
            continue;
        const uint32_t s0 = nstart[x], e = min(s0 + ncount[x], n);
        for (uint32_t q = s0; q < e; ++q)
            sleaf[q] = static_cast<uint32_t>(x);
    }
}

// Internal nodes of one level (deepest level first, after the leaves), one
// thread per node of the level. (Listing the internal ones first so that
// whole wavefronts sum them was slower: the level is latency-bound, and the
// mostly idle wavefronts of this form keep more loads in flight.)
template <bool OCT>
__global__ __launch_bounds__(256) void tree_moments_kernel(
    const uint64_t* __restrict__ nprefix, const uint32_t* __restrict__ ncount,
    const uint32_t* __restrict__ nchild, const uint32_t* __restrict__ nfirst,
    const int* __restrict__ info, int level, const RootGeom* __restrict__ geom,
    WalkNode* __restrict__ wn, float* __restrict__ ext, float* __restrict__ noct, int refit,
    int* err)
{
    const int base = info[level];
    const int cnt = info[kLevels + level];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int node = base + k;
        const NodeMeta meta = node_meta(node, nprefix, ncount, nchild, nfirst);
        if (meta.nchild == 0)
            continue;
        node_moments<OCT, false>(node, level, meta, geom, wn, ext, noct, refit, err,
                                 [](uint32_t, const double*) { return float4{}; });
    }
}


// Walk groups. Node i owns groups when
//   * i is the root and holds <= G particles: one group;
//   * i holds > G particles and is a leaf: ceil(count / G) chunks;
//   * i holds > G particles and has children: its children holding <= G
//     particles, packed greedily in octant order into runs of consecutive
//     children with <= G particles in total (a child holding > G ends a run).
//     A run also ends before a child that would stretch the run's box (the
//     smallest box of octants covering it) over a sibling holding > G
//     particles: octants 3 and 4, for instance, span the whole cell, and a
//     group whose targets surround a dense sibling would open all of it.
// Every particle lands in exactly one group; a group's particles are
// contiguous in sorted order and lie inside node i's cell, whose lattice
// centre is the group's reference point. G = gmax (at most kGroupMax, one
// target per lane of the walk's wavefront).
__device__ inline bool octant_box_holds(uint32_t ax, uint32_t ay, uint32_t az, uint32_t o)
{
    // ax/ay/az: bit v set when the box covers value v of that axis' octant bit
    return ((ax >> ((o >> 2) & 1u)) & (ay >> ((o >> 1) & 1u)) & (az >> (o & 1u)) & 1u) != 0u;
}

template <bool EMIT>
__device__ inline uint32_t node_groups(int i, const uint64_t* __restrict__ nprefix,
                                       const uint32_t* __restrict__ nstart,
                                       const uint32_t* __restrict__ ncount,
                                       const uint32_t* __restrict__ nchild,
                                       const uint32_t* __restrict__ nfirst,
                                       const uint32_t* __restrict__ nparent, uint32_t base,
                                       uint4* __restrict__ groups, uint32_t gmax)
{
    const uint32_t c = ncount[i];
    if (c <= gmax) {
        if (nparent[i] != 0xffffffffu)
            return 0;
        if (EMIT)
            groups[base] = make_uint4(nstart[i], c, static_cast<uint32_t>(i), 0u);
        return 1;
    }
    const uint32_t nc = nchild[i];
    if (nc == 0) {
        const uint32_t k = (c + gmax - 1) / gmax;
        if (EMIT)
            for (uint32_t j = 0; j < k; ++j) {
                const uint32_t off = j * gmax;
                const uint32_t cnt = c - off < gmax ? c - off : gmax;
                groups[base + j] = make_uint4(nstart[i] + off, cnt, static_cast<uint32_t>(i), 0u);
            }
        return k;
    }
    const uint32_t f = nfirst[i];
    uint32_t dense = 0;  // octants of the children holding > G particles
    for (uint32_t j = f; j < f + nc; ++j)
        if (ncount[j] > gmax)
            dense |= 1u << static_cast<uint32_t>(nprefix[j] & 7u);
    uint32_t ng = 0, run_start = 0, run_count = 0;
    uint32_t bx = 0, by = 0, bz = 0;  // the run's octant box
    for (uint32_t j = f; j < f + nc; ++j) {
        const uint32_t cc = ncount[j];
        const bool small = cc <= gmax;
        const uint32_t o = static_cast<uint32_t>(nprefix[j] & 7u);
        const uint32_t nx = bx | (1u << ((o >> 2) & 1u)), ny = by | (1u << ((o >> 1) & 1u)),
                       nz = bz | (1u << (o & 1u));
        bool covers_dense = false;
        for (uint32_t q = 0; q < 8; ++q)
            covers_dense = covers_dense || (((dense >> q) & 1u) && octant_box_holds(nx, ny, nz, q));
        if (run_count > 0 && (!small || run_count + cc > gmax || covers_dense)) {
            if (EMIT)
                groups[base + ng] = make_uint4(run_start, run_count, static_cast<uint32_t>(i), 0u);
            ++ng;
            run_count = 0;
            bx = by = bz = 0;
        }
        if (small) {
            if (run_count == 0)
                run_start = nstart[j];
            run_count += cc;
            bx |= 1u << ((o >> 2) & 1u);
            by |= 1u << ((o >> 1) & 1u);
            bz |= 1u << (o & 1u);
        }
    }
    if (run_count > 0) {
        if (EMIT)
            groups[base + ng] = make_uint4(run_start, run_count, static_cast<uint32_t>(i), 0u);
        ++ng;
    }
    return ng;
}

__global__ __launch_bounds__(256) void group_count_kernel(
    const uint64_t* __restrict__ nprefix, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, const uint32_t* __restrict__ nchild,
    const uint32_t* __restrict__ nfirst, const uint32_t* __restrict__ nparent,
    const int* __restrict__ info, uint32_t* __restrict__ gcnt, uint32_t gmax)
{
    const int total = info[kInfoTotal];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
        gcnt[i] = node_groups<false>(i, nprefix, nstart, ncount, nchild, nfirst, nparent, 0, nullptr,
                                     gmax);
}

__global__ __launch_bounds__(256) void group_emit_kernel(
    const uint64_t* __restrict__ nprefix, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, const uint32_t* __restrict__ nchild,
    const uint32_t* __restrict__ nfirst, const uint32_t* __restrict__ nparent,
    const uint32_t* __restrict__ goff,
    const int* __restrict__ info, uint4* __restrict__ groups, uint32_t gmax)
{
    const int total = info[kInfoTotal];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
        node_groups<true>(i, nprefix, nstart, ncount, nchild, nfirst, nparent, goff[i], groups,
                          gmax);
}

inline unsigned level_grid(size_t bound)
{
    const size_t b = (bound + 255) / 256;
    return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(b, 4096)));
}

inline size_t level_bound(int level, size_t n, size_t cap)
{
    size_t b = 1;
    for (int l = 0; l < level && b < n; ++l)
        b *= 8;
    return std::min(std::min(b, n), cap);
}

}  // namespace

void launch_tree_build(Device& d, int leaf_depth, int bucket, hipStream_t s)
{
    const uint32_t n = static_cast<uint32_t>(d.n);
    const unsigned long long* sorted = reinterpret_cast<const unsigned long long*>(d.keys);
    if (n > 0 && bucket <= kTopoHalo && d.topo_cnt) {
        const uint32_t ntiles = topo_tiles(d.n);
        uint32_t* pinfo = d.bnd_tmp;  // n words (the level-by-level build's child bounds)
        if (bucket == 16)  // the default
            topo_count_kernel<16><<<ntiles, kTopoThreads, 0, s>>>(
                sorted, n, bucket, leaf_depth, pinfo, d.topo_cnt, ntiles);
        else
            topo_count_kernel<0><<<ntiles, kTopoThreads, 0, s>>>(
                sorted, n, bucket, leaf_depth, pinfo, d.topo_cnt, ntiles);
        topo_scan_kernel<<<kLevels, kScanBlock, 0, s>>>(d.topo_cnt, ntiles);
        topo_emit_kernel<<<ntiles, kTopoThreads, 0, s>>>(
            sorted, n, pinfo, d.topo_cnt, ntiles, static_cast<int>(d.node_cap), d.nprefix,
            d.nlevel, d.nstart, d.ncount, d.nchild, d.nfirst, d.nparent, d.level_info, d.err);
        topo_link_kernel<<<level_grid(std::min<size_t>(d.node_cap, 2 * d.n + 64)), 256, 0,
                           s>>>(sorted, n, d.nprefix, d.nlevel, d.nstart, d.ncount,
                                       d.nchild, d.nfirst, d.nparent, d.sleaf, d.level_info,
                                       d.err);
        return;
    }
    // Buckets wider than the tiles' halo: level by level, top down.
    tree_root_kernel<<<1, 64, 0, s>>>(d.nprefix, d.nlevel, d.nstart, d.ncount, d.nchild,
                                             d.nfirst, d.nparent, d.level_info, n);
    const unsigned long long* keys = reinterpret_cast<const unsigned long long*>(d.keys);
    const int cap = static_cast<int>(d.node_cap);
    for (int L = 0; L < kMaxLevel; ++L) {
        const size_t bound = level_bound(L, d.n, d.node_cap);
        const unsigned g = level_grid(bound * 64);  // a wavefront per node on small levels
        tree_count_kernel<<<g, 256, 0, s>>>(
            keys, d.nstart, d.ncount, d.nchild, d.nfirst, d.scan_tmp, d.bnd_tmp, d.big_list,
            d.level_info, L, leaf_depth, static_cast<uint32_t>(bucket));
        if (bound > kWavePerNodeMax)  // levels that can run thread-per-node
            tree_count_big_kernel<<<64, 256, 0, s>>>(keys, d.nstart, d.ncount, d.nchild,
                                                            d.nfirst, d.scan_tmp, d.bnd_tmp,
                                                            d.big_list, d.level_info, L);
        const unsigned nscan =
            static_cast<unsigned>(std::max<size_t>(1, (bound + kScanChunk - 1) / kScanChunk));
        tree_scan_fix_kernel<<<nscan, kScanBlock, 0, s>>>(d.scan_tmp, d.scan_blocks,
                                                                 d.level_info, L, cap, d.err);
        tree_emit_kernel<<<level_grid(bound), 256, 0, s>>>(
            d.nprefix, d.nlevel, d.nstart, d.ncount, d.nchild, d.nfirst, d.nparent, d.scan_tmp,
            d.bnd_tmp, d.level_info, L);
    }
    leaf_fill_kernel<<<level_grid(std::min<size_t>(d.node_cap, 2 * d.n + 64)), 256, 0,
                       s>>>(d.nstart, d.ncount, d.nchild, d.level_info, n, d.sleaf);
}

void launch_groups(Device& d, uint32_t gmax, hipStream_t s)
{
    const unsigned g = level_grid(d.node_cap);
    group_count_kernel<<<g, 256, 0, s>>>(d.nprefix, d.nstart, d.ncount, d.nchild, d.nfirst,
                                                d.nparent, d.level_info, d.gscan, gmax);
    launch_scan_exclusive(d.gscan, d.node_cap, d.level_info + kInfoTotal, 0, d.gscan_blocks,
                          nullptr, d.level_info + kInfoGroups, d.err, s);
    group_emit_kernel<<<g, 256, 0, s>>>(d.nprefix, d.nstart, d.ncount, d.nchild, d.nfirst,
                                               d.nparent, d.gscan, d.level_info, d.groups, gmax);
}

void launch_tree_moments(Device& d, bool refit)
{
    const int rf = refit ? 1 : 0;
    const uint32_t n = static_cast<uint32_t>(d.n);
    const unsigned gl = static_cast<unsigned>((d.n + kLeafChunk - 1) / kLeafChunk);
    if (gl > 0) {
        if (d.noct)
            leaf_moments_kernel<true><<<gl, kLeafThreads, 0, d.stream>>>(
                d.nprefix, d.nlevel, d.nstart, d.ncount, d.nchild, d.nfirst, d.sleaf, n,
                d.level_info, d.pos[0], d.pos[1], d.pos[2], d.mass, d.geom, d.wnode, d.spart,
                d.next, d.noct, rf, d.err);
        else
            leaf_moments_kernel<false><<<gl, kLeafThreads, 0, d.stream>>>(
                d.nprefix, d.nlevel, d.nstart, d.ncount, d.nchild, d.nfirst, d.sleaf, n,
                d.level_info, d.pos[0], d.pos[1], d.pos[2], d.mass, d.geom, d.wnode, d.spart,
                d.next, nullptr, rf, d.err);
    }
    // level kMaxLevel holds leaves only
    for (int L = kMaxLevel - 1; L >= 0; --L) {
        const size_t bound = level_bound(L, d.n, d.node_cap);
        const unsigned g = level_grid(bound);
        if (d.noct)
            tree_moments_kernel<true><<<g, 256, 0, d.stream>>>(
                d.nprefix, d.ncount, d.nchild, d.nfirst, d.level_info, L, d.geom, d.wnode, d.next,
                d.noct, rf, d.err);
        else
            tree_moments_kernel<false><<<g, 256, 0, d.stream>>>(
                d.nprefix, d.ncount, d.nchild, d.nfirst, d.level_info, L, d.geom, d.wnode, d.next,
                nullptr, rf, d.err);
    }
}

}  // namespace bh
