// Allocation-free variant of bh_hip_sort_pairs (include/bh_hip.h).
//
// bh_hip_sort_pairs allocates a second buffer pair, the per-tile digit
// histograms, the scan's status words and an error word for every call. Some
// callers cannot call hipMalloc at all (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation). For them,
// bh_hip_sort_pairs_noalloc runs the same kernels on device-resident inputs
// and outputs and keeps all of that scratch in one file-scope device array,
// `scratch`, of fixed size: it sorts at most kNoAllocMaxKeys pairs per call.
//
// Each GPU that the library runs on holds its own copy of the array, at its
// own address, until the process ends. The address is looked up once per
// device (hipGetSymbolAddress, with that device current) on the first call
// made on it; host code only ever passes that looked-up address to kernels.
// Every call on one GPU shares the array: calls that overlap on one GPU (two
// streams, two threads) race and give wrong results. The error word is
// cleared by a fill kernel at the start of each call; the scan status words
// carry an epoch (scan_lookback.h), so they need clearing only once, at
// set-up; the histograms are fully rewritten by every pass.
#include "../../include/bh_hip.h"
#include "bh_internal.h"
#include "scan_lookback.h"

#include <cstdio>
#include <mutex>
#include <string>

namespace {

constexpr size_t kNoAllocMaxKeys = size_t(1) << 21;  // 2,097,152 pairs
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
// Digit histograms (256 words per tile of 4,096 keys) plus the scan workspace
// over them; bh::radix_hist_words(kNoAllocMaxKeys) is checked against this at
// set-up.
constexpr size_t kHistWordsMax = 256 * (kNoAllocMaxKeys / 4096) + 1024;
constexpr size_t kOffErr = 0;
constexpr size_t kOffKeys = align_up(kOffErr + sizeof(int));
constexpr size_t kOffVals = align_up(kOffKeys + kNoAllocMaxKeys * sizeof(uint64_t));
constexpr size_t kOffHist = align_up(kOffVals + kNoAllocMaxKeys * sizeof(uint32_t));
constexpr size_t kScratchBytes = align_up(kOffHist + kHistWordsMax * sizeof(uint32_t));

// ~24.6 MiB per GPU. In host code this name is a host-side stand-in, not a
// device address: only the per-device pointer from hipGetSymbolAddress below
// is passed to kernels.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

void set_err(char* err, size_t cap, const std::string& msg)
{
    if (err && cap > 0)
        std::snprintf(err, cap, "%s", msg.c_str());
}

std::string hip_msg(const char* what, hipError_t e)
{
    return std::string("[hip] ") + what + " failed: " + hipGetErrorString(e);
}

// The current device's copy of `scratch`, set up on the first call made on
// that device: address lookup, then the histogram / scan-status region and the
// error word zeroed (a fill kernel, on the caller's stream, waited for).
unsigned char* device_scratch(hipStream_t stream, std::string* msg)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *msg = hip_msg("hipGetDevice", e);
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        *msg = "bh_hip_sort_pairs_noalloc: device index out of range";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch_ptr[dev])
        return g_scratch_ptr[dev];
    if (bh::radix_hist_words(kNoAllocMaxKeys) > kHistWordsMax) {
        *msg = "bh_hip_sort_pairs_noalloc: histogram region too small";
        return nullptr;
    }
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        (void)hipGetLastError();
        *msg = hip_msg("hipGetSymbolAddress(scratch)", e);
        return nullptr;
    }
    unsigned char* base = static_cast<unsigned char*>(ptr);
    bh::launch_fill_u32(reinterpret_cast<uint32_t*>(base + kOffErr), 0u, 1, stream);
    bh::launch_fill_u32(reinterpret_cast<uint32_t*>(base + kOffHist), 0u, kHistWordsMax, stream);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        *msg = hip_msg("scratch set-up", e);
        return nullptr;
    }
    g_scratch_ptr[dev] = base;
    return base;
}

}  // namespace

extern "C" {

bh_hip_result_t bh_hip_sort_pairs_noalloc(const uint64_t* keys, const uint32_t* vals, uint64_t n,
                                          int32_t key_bits, uint64_t* keys_out,
                                          uint32_t* vals_out, void* stream, char* err,
                                          size_t err_cap)
{
    if ((n && (!keys || !vals || !keys_out || !vals_out)) || key_bits < 1 || key_bits > 64) {
        set_err(err, err_cap, "bh_hip_sort_pairs_noalloc: invalid argument");
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (n > kNoAllocMaxKeys) {
        set_err(err, err_cap, "bh_hip_sort_pairs_noalloc: n = " + std::to_string(n) +
                                  " exceeds the maximum of " + std::to_string(kNoAllocMaxKeys) +
                                  " pairs per call (fixed scratch size)");
        return BH_HIP_INVALID_ARGUMENT;
    }
    if (n == 0)
        return BH_HIP_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::string msg;
    unsigned char* base = device_scratch(s, &msg);
    if (!base) {
        set_err(err, err_cap, msg);
        return BH_HIP_INTERNAL_ERROR;
    }
    int* flags = reinterpret_cast<int*>(base + kOffErr);
    uint64_t* skeys = reinterpret_cast<uint64_t*>(base + kOffKeys);
    uint32_t* svals = reinterpret_cast<uint32_t*>(base + kOffVals);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + kOffHist);

    hipError_t e = hipSuccess;
    if (n == 1) {
        if (keys_out != keys)
            e = hipMemcpyAsync(keys_out, keys, sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && vals_out != vals)
            e = hipMemcpyAsync(vals_out, vals, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_err(err, err_cap, hip_msg("bh_hip_sort_pairs_noalloc copy", e));
            return BH_HIP_INTERNAL_ERROR;
        }
        return BH_HIP_OK;
    }
    bh::launch_fill_u32(reinterpret_cast<uint32_t*>(flags), 0u, 1, s);
    // The passes alternate between two buffer pairs and must end in the
    // output. Disjoint input and output: the output is the pair the last pass
    // writes, the scratch pair the other. Sorting in place (either output
    // array is its input): pass 0 must not write the input, so it writes the
    // scratch pair; an odd pass count then ends there and is copied out.
    const bool in_place = keys_out == keys || vals_out == vals;
    const bool odd = (bh::radix_sort_passes(key_bits) & 1) != 0;
    bool copy_out = false;
    if (!in_place) {
        if (odd)
            bh::radix_sort_from(keys, vals, keys_out, vals_out, skeys, svals, hist, n, key_bits,
                                flags, s);
        else
            bh::radix_sort_from(keys, vals, skeys, svals, keys_out, vals_out, hist, n, key_bits,
                                flags, s);
    } else {
        bh::radix_sort_from(keys, vals, skeys, svals, keys_out, vals_out, hist, n, key_bits,
                            flags, s);
        copy_out = odd;
    }
    e = hipGetLastError();
    if (e == hipSuccess && copy_out) {
        e = hipMemcpyAsync(keys_out, skeys, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(vals_out, svals, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    }
    int host_flags = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&host_flags, flags, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_err(err, err_cap, hip_msg("bh_hip_sort_pairs_noalloc", e));
        return BH_HIP_INTERNAL_ERROR;
    }
    if (host_flags != 0) {
        set_err(err, err_cap, "[octree_hip] sort: scan did not complete");
        return BH_HIP_INTERNAL_ERROR;
    }
    return BH_HIP_OK;
}

}  // extern "C"
