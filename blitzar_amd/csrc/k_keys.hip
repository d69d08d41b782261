// AABB reduction, root geometry and Morton key generation for octree_hip.
//
// Reference behaviour followed:
//   AABB over positions         engine/physics/octree/cuda/linear/OctBuild.inl:58-72
//   key quantisation            engine/physics/octree/cuda/linear/OctKeyKernels.inl:22-46
//   bit interleave              engine/physics/octree/morton/OctMortonCodes.inl:7-22
// The quantisation is the reference's per-axis formula
//   t = clamp((p - lo) * inv, 0, 1),  q = min(2^21-1, (uint)fma(t, 2^21-1, 0.5))
// (nvcc's default --fmad=true contracts the reference's `t * 2097151.0f + 0.5f`
// into that FMA). Key mode 0 interleaves all 21 bits per axis (63-bit key);
// key mode 1 reproduces the reference's interleave, which runs on 32-bit lanes
// with the 10-bit masks and therefore keeps only the low 10 bits per axis.
//
// HBM traffic: AABB reads 24 B/particle (FP64 SoA); keys read 24 B and write
// 12 B per particle. The AABB combines per-block partial extrema in a second
// one-block kernel (uses Device::diag, 6 doubles per block, 1024 blocks).
#include "bh_internal.h"

#include <cfloat>

namespace bh {
namespace {

constexpr unsigned long long kSign = 0x8000000000000000ull;

__device__ inline unsigned long long order_bits(double x)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b & kSign) ? ~b : (b | kSign);
}

__device__ inline double from_order_bits(unsigned long long e)
{
    const unsigned long long b = (e & kSign) ? (e & ~kSign) : ~e;
    return __longlong_as_double(static_cast<long long>(b));
}

__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Grid-stride min/max of the three FP64 coordinate arrays; each block writes
// its six partial extrema (no atomics; the final kernel combines them, so the
// result is deterministic).
__global__ __launch_bounds__(256) void aabb_kernel(const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz, size_t n,
                                                   double* __restrict__ partial)
{
    __shared__ double red[6][4];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const double x = px[i], y = py[i], z = pz[i];
        lo[0] = fmin(lo[0], x);
        lo[1] = fmin(lo[1], y);
        lo[2] = fmin(lo[2], z);
        hi[0] = fmax(hi[0], x);
        hi[1] = fmax(hi[1], y);
        hi[2] = fmax(hi[2], z);
    }
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        const double l = wave_min(lo[a]);
        const double h = wave_max(hi[a]);
        if ((threadIdx.x & 63) == 0) {
            red[a][w] = l;
            red[3 + a][w] = h;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        double v = red[a][0];
        for (int k = 1; k < 4; ++k)
            v = a < 3 ? fmin(v, red[a][k]) : fmax(v, red[a][k]);
        partial[6 * blockIdx.x + a] = v;
    }
}

// Combine the per-block extrema (one block).
__global__ __launch_bounds__(256) void aabb_final_kernel(const double* __restrict__ partial,
                                                         int nb, unsigned long long* bits)
{
    __shared__ double red[6][4];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], partial[6 * b + a]);
            hi[a] = fmax(hi[a], partial[6 * b + 3 + a]);
        }
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        const double l = wave_min(lo[a]);
        const double h = wave_max(hi[a]);
        if ((threadIdx.x & 63) == 0) {
            red[a][w] = l;
            red[3 + a][w] = h;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        double v = red[a][0];
        for (int k = 1; k < 4; ++k)
            v = a < 3 ? fmin(v, red[a][k]) : fmax(v, red[a][k]);
        bits[a] = order_bits(v);
    }
}

// Root geometry from the FP64 AABB (one thread).
__global__ void geometry_kernel(const unsigned long long* bits, RootGeom* g)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    for (int a = 0; a < 3; ++a) {
        const float lo = static_cast<float>(from_order_bits(bits[a]));
        const float hi = static_cast<float>(from_order_bits(bits[3 + a]));
        const float inv = __fdiv_rn(1.0f, fmaxf(hi - lo, 1.0e-12f));
        g->lo[a] = lo;
        g->inv[a] = inv;
        // The lattice unit is half of one quantisation step of the FP32 frame:
        // step = 1 / (inv * (2^21 - 1)).
        const double step = 1.0 / (static_cast<double>(inv) * static_cast<double>(kQMax));
        g->lo_d[a] = static_cast<double>(lo);
        g->unit_d[a] = 0.5 * step;
        g->unit_f[a] = static_cast<float>(0.5 * step);
        g->width_f[a] = static_cast<float>(step);
    }
}

__device__ inline unsigned long long spread21(unsigned int v)
{
    unsigned long long x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// The reference interleave as it executes: 32-bit lanes, 10-bit masks.
__device__ inline unsigned long long spread_ref32(unsigned int v)
{
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return static_cast<unsigned long long>(v);
}

__device__ inline unsigned int quantise(float p, float lo, float inv)
{
    const float t = fminf(fmaxf((p - lo) * inv, 0.0f), 1.0f);
    const unsigned int q = static_cast<unsigned int>(__fmaf_rn(t, 2097151.0f, 0.5f));
    return min(kQMax, q);
}

template <int MODE>
__global__ __launch_bounds__(256) void keys_kernel(const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz, size_t n,
                                                   const RootGeom* __restrict__ g,
                                                   unsigned long long* __restrict__ keys,
                                                   uint32_t* __restrict__ vals)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const unsigned int qx = quantise(static_cast<float>(px[i]), g->lo[0], g->inv[0]);
    const unsigned int qy = quantise(static_cast<float>(py[i]), g->lo[1], g->inv[1]);
    const unsigned int qz = quantise(static_cast<float>(pz[i]), g->lo[2], g->inv[2]);
    unsigned long long k;
    if (MODE == 1)
        k = (spread_ref32(qx) << 2) | (spread_ref32(qy) << 1) | spread_ref32(qz);
    else
        k = (spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz);
    keys[i] = k;
    vals[i] = static_cast<uint32_t>(i);
}

}  // namespace

void launch_aabb(Device& d)
{
    int nb = d.aabb_parts;  // partials the drift kernel left for these positions
    d.aabb_parts = 0;
    if (nb <= 0) {
        const size_t blocks = std::min<size_t>((d.n + 255) / 256, kAabbBlocks);
        nb = static_cast<int>(std::max<size_t>(blocks, 1));
        aabb_kernel<<<nb, 256, 0, d.stream>>>(d.pos[0], d.pos[1], d.pos[2], d.n, d.diag);
    }
    aabb_final_kernel<<<1, 256, 0, d.stream>>>(d.diag, nb, d.aabb_bits);
    geometry_kernel<<<1, 64, 0, d.stream>>>(d.aabb_bits, d.geom);
}

void launch_keys(Device& d, int key_mode)
{
    const unsigned blocks = static_cast<unsigned>((d.n + 255) / 256);
    if (blocks == 0)
        return;
    if (key_mode == 1)
        keys_kernel<1><<<blocks, 256, 0, d.stream>>>(
            d.pos[0], d.pos[1], d.pos[2], d.n, d.geom,
            reinterpret_cast<unsigned long long*>(d.keys), d.vals);
    else
        keys_kernel<0><<<blocks, 256, 0, d.stream>>>(
            d.pos[0], d.pos[1], d.pos[2], d.n, d.geom,
            reinterpret_cast<unsigned long long*>(d.keys), d.vals);
}

}  // namespace bh
