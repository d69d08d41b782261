// Single-pass device-wide exclusive scan of uint32 words (decoupled look-back).
//
// One launch of ceil(capacity / kScanChunk) blocks of kScanBlock threads scans
// the first n words in place (n may live on the device); blocks past the end
// of the data return at once. Block b scans words [b * kScanChunk, ...) and
// publishes its own sum at once; wave 0 then reads the status words of up to
// 64 predecessors per step until it meets one that carries an inclusive
// prefix. A block only waits for blocks of lower index, which the dispatcher
// starts before it, and never waits before publishing its own sum, so the
// lowest unfinished block can always finish. The number of polls is bounded
// all the same: a block that exceeds it raises an error bit and stops.
// (Block order comes from blockIdx rather than from a ticket counter: a
// counter is one address that every block of every scan would hit.)
//
// Workspace (scan_workspace_words(capacity) uint32 words, zeroed once at
// allocation, never cleared again):
//   ws[0]       epoch of the next launch (advanced by the block that ends the scan)
//   ws[64 ...]  one 64-bit status word per block:
//               epoch (30 bits) | flag (2 bits: 1 block sum, 2 inclusive prefix) | value (32 bits)
// A status word written by an earlier launch carries an older epoch and reads
// as empty, so the workspace needs no clearing between launches (and a launch
// replayed from a graph behaves like any other). Every block of a launch
// reads the epoch before it publishes, and the last block advances it only
// after its look-back has seen every earlier block's status.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace bh {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanBlock * kScanItems;  // 2048
constexpr int kScanWaves = kScanBlock / 64;
constexpr int kScanHeaderWords = 64;
// Upper bound on look-back polls of one block; beyond it the block flags an
// error and stops waiting (results of that scan are then wrong, not hung).
constexpr uint32_t kScanMaxPolls = 1u << 22;

inline size_t scan_workspace_words(size_t capacity)
{
    const size_t nb = (capacity + kScanChunk - 1) / kScanChunk + 1;
    return kScanHeaderWords + 2 * nb;
}

// Inclusive scan over one wave.
__device__ inline uint32_t wave_inclusive(uint32_t v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o)
            v += u;
    }
    return v;
}

// Exclusive scan over a 256-thread block; returns the exclusive prefix and
// the block total in *total. `tmp` must hold kScanWaves words.
__device__ inline uint32_t block_exclusive256(uint32_t v, uint32_t* tmp, uint32_t* total)
{
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive(v);
    if (lane == 63)
        tmp[w] = inc;
    __syncthreads();
    uint32_t off = 0, all = 0;
    for (int k = 0; k < kScanWaves; ++k) {
        const uint32_t s = tmp[k];
        if (k < w)
            off += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return off + inc - v;
}

__device__ inline unsigned long long scan_status(uint32_t epoch, uint32_t flag, uint32_t value)
{
    return (static_cast<unsigned long long>(epoch & 0x3fffffffu) << 34) |
           (static_cast<unsigned long long>(flag) << 32) | value;
}

struct ScanBlockResult {
    bool last;       // this block holds the scan's last word (or, when n == 0, block 0)
    uint32_t total;  // sum of all n words (valid in the last block)
};

// The whole scan as run by one block; every block of the launch must call it
// exactly once, with the same n.
__device__ inline ScanBlockResult scan_lookback_block(uint32_t* __restrict__ data, size_t n,
                                                      uint32_t* __restrict__ ws, int* err,
                                                      int err_bit)
{
    __shared__ uint32_t s_epoch, s_excl;
    __shared__ uint32_t red[kScanWaves];
    unsigned long long* status = reinterpret_cast<unsigned long long*>(ws + kScanHeaderWords);
    const size_t nblocks = (n + kScanChunk - 1) / kScanChunk;
    const uint32_t b = blockIdx.x;
    ScanBlockResult r{false, 0u};
    if (n == 0) {
        r.last = b == 0;
        return r;
    }
    if (b >= nblocks)
        return r;
    if (threadIdx.x == 0)
        s_epoch = __hip_atomic_load(&ws[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    const size_t base = static_cast<size_t>(b) * kScanChunk;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t i = base + static_cast<size_t>(threadIdx.x) * kScanItems + k;
        v[k] = i < n ? data[i] : 0u;
        s += v[k];
    }
    uint32_t agg;
    uint32_t off = block_exclusive256(s, red, &agg);  // its barriers also publish s_epoch
    const uint32_t epoch = s_epoch;

    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t excl = 0;
        if (b == 0) {
            if (lane == 0)
                __hip_atomic_store(&status[0], scan_status(epoch, 2u, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&status[b], scan_status(epoch, 1u, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            long long p = static_cast<long long>(b) - 1;  // nearest predecessor in this step
            uint32_t polls = 0;
            for (;;) {
                const long long idx = p - lane;
                uint32_t flag = 2u, val = 0u;  // before block 0: an inclusive zero
                if (idx >= 0) {
                    const unsigned long long st = __hip_atomic_load(
                        &status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const bool cur = static_cast<uint32_t>(st >> 34) == (epoch & 0x3fffffffu);
                    flag = cur ? static_cast<uint32_t>(st >> 32) & 3u : 0u;
                    val = static_cast<uint32_t>(st);
                }
                const unsigned long long incl = __ballot(flag == 2u);
                const unsigned long long empty = __ballot(flag == 0u);
                const int first = incl ? __builtin_ctzll(incl) : 64;
                const unsigned long long need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
                if (empty & need) {
                    if (++polls > kScanMaxPolls) {
                        if (lane == 0)
                            atomicOr(err, err_bit);
                        break;
                    }
                    continue;
                }
                uint32_t part = lane <= first ? val : 0u;
                for (int o = 32; o > 0; o >>= 1)
                    part += __shfl_xor(part, o, 64);
                excl += part;
                if (first < 64)
                    break;
                p -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&status[b], scan_status(epoch, 2u, excl + agg),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_excl = excl;
            if (b == nblocks - 1)
                __hip_atomic_store(&ws[0], epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    off += s_excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t i = base + static_cast<size_t>(threadIdx.x) * kScanItems + k;
        if (i < n)
            data[i] = off;
        off += v[k];
    }
    r.last = b == nblocks - 1;
    r.total = s_excl + agg;
    return r;
}

}  // namespace bh
