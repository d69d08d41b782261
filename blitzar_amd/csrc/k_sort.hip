// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, device-wide
// exclusive scan, equal-key tie ordering and the particle-store permutation.
//
// Contract followed: thrust::sort_by_key(keys, identity indices) at
// engine/physics/octree/cuda/linear/OctBuild.inl:96 -- ascending keys, equal
// keys keep ascending original particle index. The store is kept in the last
// sorted order, so the sort's own stability orders ties by previous slot;
// tie_fix_kernel then reorders every run of equal keys by particle id, which
// restores the reference contract exactly.
//
// One pass (8-bit digit) = histogram kernel (reads 8 B/key), one single-pass
// scan of the digit-major tile counts, and a scatter kernel that ranks keys inside a tile
// of 4096 with wave ballots, reorders the tile in LDS and writes digit runs
// contiguously (reads 12 B/key, writes 12 B/key).
#include "bh_internal.h"
#include "scan_lookback.h"

namespace bh {
namespace {

constexpr int kBlock = 256;
#ifndef BH_SORT_ITEMS
#define BH_SORT_ITEMS 16
#endif
// Keys per thread. A tile of 4096 keys (16 per thread) writes digit runs of 16
// keys on average, twice those of a 2048-key tile: fewer partial cache lines
// (70 us less per 8M-key sort than 8 per thread, at 2 blocks per CU).
constexpr int kItems = BH_SORT_ITEMS;
constexpr int kTile = kBlock * kItems;  // keys per tile
constexpr int kWaves = kBlock / 64;

__device__ inline unsigned long long lanemask_lt()
{
    const unsigned lane = threadIdx.x & 63u;
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

__global__ __launch_bounds__(kBlock) void hist_kernel(const unsigned long long* __restrict__ keys,
                                                      size_t n, int shift, uint32_t dmask,
                                                      uint32_t* __restrict__ hist,
                                                      uint32_t ntiles)
{
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile;
    for (int k = 0; k < kItems; ++k) {
        const size_t i = base + static_cast<size_t>(k) * kBlock + t;
        if (i < n)
            atomicAdd(&h[(keys[i] >> shift) & dmask], 1u);
    }
    __syncthreads();
    hist[static_cast<size_t>(t) * ntiles + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(
    const unsigned long long* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    unsigned long long* __restrict__ keys_out, uint32_t* __restrict__ vals_out, size_t n,
    int shift, uint32_t dmask, const uint32_t* __restrict__ hist, uint32_t ntiles)
{
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t tile_off[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t red[kWaves];
    __shared__ unsigned long long sk[kTile];
    __shared__ uint32_t sv[kTile];

    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    const size_t tile = blockIdx.x;
    const size_t tbase = tile * kTile;
    const uint32_t tile_n =
        static_cast<uint32_t>(n - tbase < static_cast<size_t>(kTile) ? n - tbase : kTile);

    for (int k = 0; k < kWaves; ++k)
        wcnt[k][t] = 0;
    gbase[t] = hist[static_cast<size_t>(t) * ntiles + tile];
    __syncthreads();

    unsigned long long key[kItems];
    uint32_t val[kItems];
    uint32_t rnk[kItems];
    const size_t wbase = tbase + static_cast<size_t>(w) * 64 * kItems;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const size_t i = wbase + static_cast<size_t>(it) * 64 + lane;
        const bool valid = i < n;
        key[it] = valid ? keys_in[i] : 0ull;
        val[it] = valid ? vals_in[i] : 0u;
    }
    const unsigned long long lt = lanemask_lt();
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const size_t i = wbase + static_cast<size_t>(it) * 64 + lane;
        const bool valid = i < n;
        const uint32_t d = static_cast<uint32_t>(key[it] >> shift) & dmask;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t before = static_cast<uint32_t>(__popcll(peers & lt));
        const uint32_t base = valid ? wcnt[w][d] : 0u;
        // The lowest lane of each digit group advances that wave's counter;
        // LDS operations of one wave complete in program order, so every
        // lane has read the old value above.
        if (valid && before == 0)
            wcnt[w][d] = base + static_cast<uint32_t>(__popcll(peers));
        rnk[it] = base + before;
    }
    __syncthreads();
    // Per digit: exclusive offsets across waves, then across digits in the tile.
    {
        uint32_t run = 0;
        for (int k = 0; k < kWaves; ++k) {
            const uint32_t c = wcnt[k][t];
            wcnt[k][t] = run;
            run += c;
        }
        uint32_t total;
        const uint32_t ex = block_exclusive256(run, red, &total);
        tile_off[t] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const size_t i = wbase + static_cast<size_t>(it) * 64 + lane;
        if (i < n) {
            const uint32_t d = static_cast<uint32_t>(key[it] >> shift) & dmask;
            const uint32_t pos = tile_off[d] + wcnt[w][d] + rnk[it];
            sk[pos] = key[it];
            sv[pos] = val[it];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kItems; ++m) {
        const uint32_t j = static_cast<uint32_t>(m) * kBlock + t;
        if (j < tile_n) {
            const unsigned long long k = sk[j];
            const uint32_t d = static_cast<uint32_t>(k >> shift) & dmask;
            const size_t out = static_cast<size_t>(gbase[d]) + (j - tile_off[d]);
            keys_out[out] = k;
            vals_out[out] = sv[j];
        }
    }
}

// ---- device-wide exclusive scan (element count from host or device) -------
__global__ __launch_bounds__(kScanBlock) void scan_lookback_kernel(
    uint32_t* __restrict__ data, const int* n_dev, size_t n_host, uint32_t* __restrict__ ws,
    uint32_t* total_out, int* total_out_i, int* err)
{
    size_t n = n_dev ? static_cast<size_t>(*n_dev > 0 ? *n_dev : 0) : n_host;
    n = n < static_cast<size_t>(gridDim.x) * kScanChunk ? n : static_cast<size_t>(gridDim.x) * kScanChunk;
    const ScanBlockResult r = scan_lookback_block(data, n, ws, err, kErrScan);
    if (r.last && threadIdx.x == 0) {
        if (total_out)
            *total_out = r.total;
        if (total_out_i)
            *total_out_i = static_cast<int>(r.total);
    }
}

// Runs of equal keys: order the run by particle id (insertion sort; runs are
// short unless particles coincide).
__global__ __launch_bounds__(256) void tie_fix_kernel(const unsigned long long* __restrict__ keys,
                                                      uint32_t* __restrict__ vals,
                                                      const uint32_t* __restrict__ id, size_t n)
{
    const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s + 1 >= n)
        return;
    const unsigned long long k = keys[s];
    if (keys[s + 1] != k || (s > 0 && keys[s - 1] == k))
        return;
    size_t e = s + 2;
    while (e < n && keys[e] == k)
        ++e;
    for (size_t a = s + 1; a < e; ++a) {
        const uint32_t v = vals[a];
        const uint32_t key_id = id[v];
        size_t b = a;
        while (b > s && id[vals[b - 1]] > key_id) {
            vals[b] = vals[b - 1];
            --b;
        }
        vals[b] = v;
    }
}

// The store follows the sort in two parts. Positions and masses, which the
// moments need, go first on the build's main stream; velocities, ids and the
// relative criterion's |a|, needed only by the walk and the integrator, go
// on the second stream beside the moments.
__global__ __launch_bounds__(256) void permute_store_kernel(
    const uint32_t* __restrict__ src, size_t n, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const float* __restrict__ m,
    double* __restrict__ opx, double* __restrict__ opy, double* __restrict__ opz,
    float* __restrict__ om)
{
    const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const uint32_t j = src[s];
    opx[s] = px[j];
    opy[s] = py[j];
    opz[s] = pz[j];
    om[s] = m[j];
}

__global__ __launch_bounds__(256) void permute_rest_kernel(
    const uint32_t* __restrict__ src, size_t n, const double* __restrict__ vx,
    const double* __restrict__ vy, const double* __restrict__ vz,
    const uint32_t* __restrict__ id, const float* __restrict__ am, double* __restrict__ ovx,
    double* __restrict__ ovy, double* __restrict__ ovz, uint32_t* __restrict__ oid,
    float* __restrict__ oam)
{
    const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const uint32_t j = src[s];
    ovx[s] = vx[j];
    ovy[s] = vy[j];
    ovz[s] = vz[j];
    oid[s] = id[j];
    if (am)
        oam[s] = am[j];
}

// The block time-step arrays follow the store through the same permutation.
__global__ __launch_bounds__(256) void permute_adaptive_kernel(
    const uint32_t* __restrict__ src, size_t n, const float* __restrict__ ax,
    const float* __restrict__ ay, const float* __restrict__ az, const uint32_t* __restrict__ lv,
    const uint32_t* __restrict__ lt, float* __restrict__ oax, float* __restrict__ oay,
    float* __restrict__ oaz, uint32_t* __restrict__ olv, uint32_t* __restrict__ olt)
{
    const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n)
        return;
    const uint32_t j = src[s];
    oax[s] = ax[j];
    oay[s] = ay[j];
    oaz[s] = az[j];
    olv[s] = lv[j];
    olt[s] = lt[j];
}

// rank[id[s]] = s: caller index -> store slot (rebuilt only when an API call
// needs it, not every step).
__global__ __launch_bounds__(256) void rank_kernel(const uint32_t* __restrict__ id, size_t n,
                                                   uint32_t* __restrict__ rank)
{
    const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < n)
        rank[id[s]] = static_cast<uint32_t>(s);
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

}  // namespace

size_t scan_block_sums_words(size_t max_n) { return scan_workspace_words(max_n); }

void launch_scan_exclusive(uint32_t* data, size_t max_n, const int* n_dev, size_t n_host,
                           uint32_t* block_sums, uint32_t* total_out, int* total_out_i,
                           int* err, hipStream_t s)
{
    const size_t cap = n_dev ? max_n : n_host;
    const unsigned nb = static_cast<unsigned>(std::max<size_t>(1, (cap + kScanChunk - 1) / kScanChunk));
    scan_lookback_kernel<<<nb, kScanBlock, 0, s>>>(data, n_dev, n_host, block_sums, total_out,
                                                   total_out_i, err);
}

size_t radix_hist_words(size_t n)
{
    const size_t tiles = std::max<size_t>(1, (n + kTile - 1) / kTile);
    return 256 * tiles + scan_block_sums_words(256 * tiles);
}

int radix_sort_passes(int key_bits) { return (key_bits + 7) / 8; }

bool radix_sort_from(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_a,
                     uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b, uint32_t* hist_ws,
                     size_t n, int key_bits, int* err, hipStream_t stream)
{
    const int passes = radix_sort_passes(key_bits);
    if (n <= 1)
        return false;
    const uint32_t ntiles = static_cast<uint32_t>((n + kTile - 1) / kTile);
    const size_t hn = static_cast<size_t>(ntiles) * 256;
    uint32_t* hist = hist_ws;
    uint32_t* sums = hist_ws + hn;
    const unsigned long long* kin = reinterpret_cast<const unsigned long long*>(keys_in);
    const uint32_t* vin = vals_in;
    unsigned long long* kout = reinterpret_cast<unsigned long long*>(keys_a);
    uint32_t* vout = vals_a;
    unsigned long long* knext = reinterpret_cast<unsigned long long*>(keys_b);
    uint32_t* vnext = vals_b;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        // The last digit covers only the remaining key bits.
        const int bits = key_bits - shift < 8 ? key_bits - shift : 8;
        const uint32_t dmask = (1u << bits) - 1u;
        hist_kernel<<<ntiles, kBlock, 0, stream>>>(kin, n, shift, dmask, hist, ntiles);
        launch_scan_exclusive(hist, hn, nullptr, hn, sums, nullptr, nullptr, err, stream);
        scatter_kernel<<<ntiles, kBlock, 0, stream>>>(kin, vin, kout, vout, n, shift, dmask,
                                                      hist, ntiles);
        // this pass's output is the next pass's input; the next output is the
        // other buffer pair
        kin = kout;
        vin = vout;
        std::swap(kout, knext);
        std::swap(vout, vnext);
    }
    return (passes & 1) != 0;
}

bool radix_sort_buffers(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                        uint32_t* hist_ws, size_t n, int key_bits, int* err, hipStream_t stream)
{
    return radix_sort_from(keys, vals, keys_alt, vals_alt, keys, vals, hist_ws, n, key_bits, err,
                           stream);
}

void launch_radix_sort(Device& d, size_t n, int key_bits)
{
    if (n <= 1)
        return;
    if (radix_sort_buffers(d.keys, d.keys_b, d.vals, d.vals_b, d.sort_hist, n, key_bits, d.err,
                           d.stream)) {
        // Result is in the alternate buffers: swap the roles so d.keys/d.vals hold it.
        std::swap(d.keys, d.keys_b);
        std::swap(d.vals, d.vals_b);
    }
}

void launch_tie_fix(Device& d)
{
    if (d.n < 2)
        return;
    const unsigned blocks = static_cast<unsigned>((d.n + 255) / 256);
    tie_fix_kernel<<<blocks, 256, 0, d.stream>>>(reinterpret_cast<const unsigned long long*>(d.keys),
                                                d.vals, d.id, d.n);
}

void launch_permute_store(Device& d, hipStream_t st)
{
    if (d.n == 0)
        return;
    const unsigned blocks = static_cast<unsigned>((d.n + 255) / 256);
    permute_store_kernel<<<blocks, 256, 0, st>>>(d.vals, d.n, d.pos[0], d.pos[1], d.pos[2],
                                                 d.mass, d.pos_b[0], d.pos_b[1], d.pos_b[2],
                                                 d.mass_b);
    for (int a = 0; a < 3; ++a)
        std::swap(d.pos[a], d.pos_b[a]);
    std::swap(d.mass, d.mass_b);
}

void launch_permute_rest(Device& d, hipStream_t st)
{
    if (d.n == 0)
        return;
    const unsigned blocks = static_cast<unsigned>((d.n + 255) / 256);
    permute_rest_kernel<<<blocks, 256, 0, st>>>(
        d.vals, d.n, d.vel[0], d.vel[1], d.vel[2], d.id, d.amag_live ? d.amag : nullptr,
        d.vel_b[0], d.vel_b[1], d.vel_b[2], d.id_b, d.amag_b);
    for (int a = 0; a < 3; ++a)
        std::swap(d.vel[a], d.vel_b[a]);
    std::swap(d.id, d.id_b);
    if (d.amag_live)
        std::swap(d.amag, d.amag_b);
    if (d.alevel) {
        permute_adaptive_kernel<<<blocks, 256, 0, st>>>(
            d.vals, d.n, d.aacc[0], d.aacc[1], d.aacc[2], d.alevel, d.atick, d.aacc_b[0],
            d.aacc_b[1], d.aacc_b[2], d.alevel_b, d.atick_b);
        for (int a = 0; a < 3; ++a)
            std::swap(d.aacc[a], d.aacc_b[a]);
        std::swap(d.alevel, d.alevel_b);
        std::swap(d.atick, d.atick_b);
    }
}

void launch_rank(Device& d)
{
    if (d.n == 0)
        return;
    rank_kernel<<<static_cast<unsigned>((d.n + 255) / 256), 256, 0, d.stream>>>(d.id, d.n, d.rank);
}

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0)
        return;
    fill_u32_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(p, v, n);
}

}  // namespace bh
