// FP64 kick/drift, host<->store reordering, FP64 direct-sum and energy
// diagnostics for octree_hip.
//
// Kick/drift follows the reference KDK kernels
// (engine/physics/core/cuda/prelude/CudParticleKernels.inl:94-220) in FP64:
//   v += a dt/2 ; x += v dt ; <rebuild + walk> ; v += a dt/2
// with the first force of a step reused from the end of the previous one. The
// Euler mode is the reference CPU update
// (engine/physics/octree/model/OctParticleSystemHost.cpp:61-69):
//   v = clamp(v + a dt, max_speed) ; x += v dt.
// Kick/drift traffic per particle and step: read x 24, v 24, a 12; write x 24,
// v 24 (108 B).
//
// Direct sum: engine/physics/treepm/model/TpmCpu.cpp:78-116, here with FP64
// positions (the truth the force error is measured against).
// Energies / momenta: engine/physics/fmm/metrics/FmmMetrics.cpp:47-83.
#include "bh_internal.h"

#include <cfloat>

namespace bh {
namespace {

// Position wrapped into [0, box) (box <= 0: unchanged).
__device__ inline double wrap_box(double x, double box)
{
    if (box <= 0.0)
        return x;
    x -= box * floor(x / box);
    return x < box ? x : 0.0;  // x / box rounding just below 1
}

__device__ inline void clamp3(double& x, double& y, double& z, double limit)
{
    const double s = sqrt(x * x + y * y + z * z);
    if (limit > 0.0 && s > limit && s > 1e-6) {
        const double k = limit / s;
        x *= k;
        y *= k;
        z *= k;
    }
}

__global__ __launch_bounds__(256) void kick_drift_kernel(
    double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const float* __restrict__ ax, const float* __restrict__ ay, const float* __restrict__ az,
    size_t n, double dtk, double dtd, double vmax, double box)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double x = vx[i] + static_cast<double>(ax[i]) * dtk;
    double y = vy[i] + static_cast<double>(ay[i]) * dtk;
    double z = vz[i] + static_cast<double>(az[i]) * dtk;
    clamp3(x, y, z, vmax);
    vx[i] = x;
    vy[i] = y;
    vz[i] = z;
    if (dtd != 0.0) {
        px[i] = wrap_box(px[i] + x * dtd, box);
        py[i] = wrap_box(py[i] + y * dtd, box);
        pz[i] = wrap_box(pz[i] + z * dtd, box);
    }
}

__device__ inline double wave_min_d(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double wave_max_d(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Kick and drift (grid-stride) that also keeps the extrema of the new
// positions: each block writes its six partial extrema in the layout of the
// build's AABB pass, which then only combines them (same values, so the same
// root geometry bit for bit).
__global__ __launch_bounds__(256) void kick_drift_aabb_kernel(
    double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const float* __restrict__ ax, const float* __restrict__ ay, const float* __restrict__ az,
    size_t n, double dtk, double dtd, double vmax, double box, double* __restrict__ partial)
{
    __shared__ double red[6][4];
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        double x = vx[i] + static_cast<double>(ax[i]) * dtk;
        double y = vy[i] + static_cast<double>(ay[i]) * dtk;
        double z = vz[i] + static_cast<double>(az[i]) * dtk;
        clamp3(x, y, z, vmax);
        vx[i] = x;
        vy[i] = y;
        vz[i] = z;
        const double nx = wrap_box(px[i] + x * dtd, box);
        const double ny = wrap_box(py[i] + y * dtd, box);
        const double nz = wrap_box(pz[i] + z * dtd, box);
        px[i] = nx;
        py[i] = ny;
        pz[i] = nz;
        lo[0] = fmin(lo[0], nx);
        lo[1] = fmin(lo[1], ny);
        lo[2] = fmin(lo[2], nz);
        hi[0] = fmax(hi[0], nx);
        hi[1] = fmax(hi[1], ny);
        hi[2] = fmax(hi[2], nz);
    }
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        const double l = wave_min_d(lo[a]);
        const double h = wave_max_d(hi[a]);
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

__global__ __launch_bounds__(256) void wrap_kernel(double* __restrict__ px,
                                                   double* __restrict__ py,
                                                   double* __restrict__ pz, size_t n, double box)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    px[i] = wrap_box(px[i], box);
    py[i] = wrap_box(py[i], box);
    pz[i] = wrap_box(pz[i], box);
}

__global__ __launch_bounds__(256) void init_ids_kernel(uint32_t* __restrict__ id,
                                                       uint32_t* __restrict__ rank, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        id[i] = static_cast<uint32_t>(i);
        rank[i] = static_cast<uint32_t>(i);
    }
}

// stage holds caller-ordered xyz triples; write the store slot rank[i].
__global__ __launch_bounds__(256) void scatter3_kernel(const double* __restrict__ stage,
                                                       const uint32_t* __restrict__ rank,
                                                       double* __restrict__ x,
                                                       double* __restrict__ y,
                                                       double* __restrict__ z, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t s = rank[i];
    x[s] = stage[3 * i];
    y[s] = stage[3 * i + 1];
    z[s] = stage[3 * i + 2];
}

__global__ __launch_bounds__(256) void gather3_kernel(double* __restrict__ stage,
                                                      const uint32_t* __restrict__ rank,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t s = rank[i];
    stage[3 * i] = x[s];
    stage[3 * i + 1] = y[s];
    stage[3 * i + 2] = z[s];
}

__global__ __launch_bounds__(256) void gather3f_kernel(double* __restrict__ stage,
                                                       const uint32_t* __restrict__ rank,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ y,
                                                       const float* __restrict__ z, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t s = rank[i];
    stage[3 * i] = static_cast<double>(x[s]);
    stage[3 * i + 1] = static_cast<double>(y[s]);
    stage[3 * i + 2] = static_cast<double>(z[s]);
}

__device__ inline double block_sum256(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < 4; ++k)
        t += red[k];
    __syncthreads();
    return t;
}

// One 256-thread block per target: FP64 sum over every source j != target.
__global__ __launch_bounds__(256) void direct_fp64_kernel(
    const uint32_t* __restrict__ targets, const uint32_t* __restrict__ rank,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const float* __restrict__ pm, size_t n, double eps2, double min_d2, double* __restrict__ out)
{
    __shared__ double red[4];
    const uint32_t self = rank[targets[blockIdx.x]];
    const double xi = px[self], yi = py[self], zi = pz[self];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (size_t j = threadIdx.x; j < n; j += blockDim.x) {
        if (j == self)
            continue;
        const double dx = px[j] - xi, dy = py[j] - yi, dz = pz[j] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        if (r2 <= min_d2)
            continue;
        const double inv = 1.0 / sqrt(r2);
        const double s = static_cast<double>(pm[j]) * inv * inv * inv;
        ax += dx * s;
        ay += dy * s;
        az += dz * s;
    }
    ax = block_sum256(ax, red);
    ay = block_sum256(ay, red);
    az = block_sum256(az, red);
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = ax;
        out[3 * blockIdx.x + 1] = ay;
        out[3 * blockIdx.x + 2] = az;
    }
}

// Per-block partial sums of KE, P(3), L(3), mass.
__global__ __launch_bounds__(256) void energy_partial_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const float* __restrict__ pm, size_t n, double* __restrict__ partial)
{
    __shared__ double red[4];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const double m = pm[i];
        const double x = px[i], y = py[i], z = pz[i];
        const double u = vx[i], v = vy[i], w = vz[i];
        acc[0] += 0.5 * m * (u * u + v * v + w * w);
        acc[1] += m * u;
        acc[2] += m * v;
        acc[3] += m * w;
        acc[4] += m * (y * w - z * v);
        acc[5] += m * (z * u - x * w);
        acc[6] += m * (x * v - y * u);
        acc[7] += m;
    }
    for (int k = 0; k < 8; ++k) {
        const double t = block_sum256(acc[k], red);
        if (threadIdx.x == 0)
            partial[8 * blockIdx.x + k] = t;
    }
}

__global__ void energy_final_kernel(const double* __restrict__ partial, int nb,
                                    double* __restrict__ out)
{
    const int k = threadIdx.x;
    if (k >= 8)
        return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b)
        s += partial[8 * b + k];
    out[k] = s;
}

// Softened potential of sampled targets: phi_i = -sum_{j != i} m_j / r_ij.
__global__ __launch_bounds__(256) void potential_kernel(
    const uint32_t* __restrict__ slots, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const float* __restrict__ pm,
    size_t n, double eps2, double min_d2, double* __restrict__ out)
{
    __shared__ double red[4];
    const uint32_t self = slots[blockIdx.x];
    const double xi = px[self], yi = py[self], zi = pz[self];
    double phi = 0.0;
    for (size_t j = threadIdx.x; j < n; j += blockDim.x) {
        if (j == self)
            continue;
        const double dx = px[j] - xi, dy = py[j] - yi, dz = pz[j] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        if (r2 > min_d2)
            phi -= static_cast<double>(pm[j]) / sqrt(r2);
    }
    phi = block_sum256(phi, red);
    if (threadIdx.x == 0)
        out[blockIdx.x] = phi * static_cast<double>(pm[self]);
}

inline unsigned blocks_for(size_t n) { return static_cast<unsigned>((n + 255) / 256); }
// grids of the kernels that reduce into one device counter: at most 2,048
// blocks (8 per CU), each adding once
inline unsigned reduce_blocks(size_t n) { return std::min(blocks_for(n), 2048u); }

}  // namespace

void launch_kick(Device& d, double dt_kick, double max_speed)
{
    if (d.n == 0)
        return;
    kick_drift_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(
        d.pos[0], d.pos[1], d.pos[2], d.vel[0], d.vel[1], d.vel[2], d.acc[0], d.acc[1], d.acc[2],
        d.n, dt_kick, 0.0, max_speed, 0.0);
}

void launch_kick_drift(Device& d, double dt_kick, double dt_drift, double max_speed, double box)
{
    if (d.n == 0)
        return;
    const unsigned nb = std::min(blocks_for(d.n), static_cast<unsigned>(kAabbBlocks));
    kick_drift_aabb_kernel<<<nb, 256, 0, d.stream>>>(
        d.pos[0], d.pos[1], d.pos[2], d.vel[0], d.vel[1], d.vel[2], d.acc[0], d.acc[1], d.acc[2],
        d.n, dt_kick, dt_drift, max_speed, box, d.diag);
    d.aabb_parts = static_cast<int>(nb);
}

void launch_wrap(Device& d, double box)
{
    d.aabb_parts = 0;
    if (d.n == 0 || box <= 0.0)
        return;
    wrap_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(d.pos[0], d.pos[1], d.pos[2], d.n, box);
}

void launch_init_ids(Device& d)
{
    if (d.n == 0)
        return;
    init_ids_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(d.id, d.rank, d.n);
}

void launch_upload_scatter(Device& d, int which)
{
    d.aabb_parts = 0;
    if (d.n == 0)
        return;
    double** dst = which == 0 ? d.pos : d.vel;
    scatter3_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(d.stage, d.rank, dst[0], dst[1],
                                                           dst[2], d.n);
}

void launch_download_gather(Device& d, int which)
{
    if (d.n == 0)
        return;
    if (which == 2) {
        gather3f_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(d.stage, d.rank, d.acc[0],
                                                               d.acc[1], d.acc[2], d.n);
        return;
    }
    double** src = which == 0 ? d.pos : d.vel;
    gather3_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(d.stage, d.rank, src[0], src[1], src[2],
                                                          d.n);
}

void launch_direct_fp64(Device& d, const uint32_t* targets_dev, uint32_t nt, double eps,
                        double min_d2, double* out_dev)
{
    if (nt == 0 || d.n == 0)
        return;
    direct_fp64_kernel<<<nt, 256, 0, d.stream>>>(targets_dev, d.rank, d.pos[0], d.pos[1],
                                                 d.pos[2], d.mass, d.n, eps * eps, min_d2,
                                                 out_dev);
}

void launch_energy(Device& d, double* out_dev)
{
    d.aabb_parts = 0;  // diag is overwritten
    const int nb = 256;
    energy_partial_kernel<<<nb, 256, 0, d.stream>>>(d.pos[0], d.pos[1], d.pos[2], d.vel[0],
                                                    d.vel[1], d.vel[2], d.mass, d.n, d.diag);
    energy_final_kernel<<<1, 64, 0, d.stream>>>(d.diag, nb, out_dev);
}

void launch_potential(Device& d, const uint32_t* slots_dev, uint32_t ns, double eps,
                      double min_d2, double* out_dev)
{
    if (ns == 0 || d.n == 0)
        return;
    potential_kernel<<<ns, 256, 0, d.stream>>>(slots_dev, d.pos[0], d.pos[1], d.pos[2], d.mass,
                                               d.n, eps * eps, min_d2, out_dev);
}

// ---- dyadic block time steps -------------------------------------------------
// Semantics of engine/physics/core/cuda/integration/CudAdaptive.inl:31-57
// (level rule), 147-172 (drift), 181-228 (activity, correction); FP64 state,
// FP32 accelerations.
namespace {

__device__ inline uint32_t adaptive_level(float ax, float ay, float az, double vx, double vy,
                                          double vz, const AdaptiveParams& a)
{
    const double am = sqrt(static_cast<double>(ax) * ax + static_cast<double>(ay) * ay +
                           static_cast<double>(az) * az);
    const double vm = sqrt(vx * vx + vy * vy + vz * vz);
    const double adt = am > 1e-6 ? a.eta * sqrt(a.soft / am) : a.dt;
    const double vdt = a.velocity && vm > 1e-6 ? a.eta * a.soft / vm : a.dt;
    const double stable = fmin(a.dt, fmin(adt, vdt));
    for (int l = 0; l <= a.levels; ++l)
        if (a.dt / static_cast<double>(1u << l) <= stable)
            return static_cast<uint32_t>(l);
    return static_cast<uint32_t>(a.levels);
}

__global__ __launch_bounds__(256) void adaptive_init_kernel(
    const float* __restrict__ ax, const float* __restrict__ ay, const float* __restrict__ az,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz,
    uint32_t* __restrict__ level, uint32_t* __restrict__ last, size_t n, AdaptiveParams a,
    uint32_t tick)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    bx[i] = ax[i];
    by[i] = ay[i];
    bz[i] = az[i];
    level[i] = adaptive_level(ax[i], ay[i], az[i], vx[i], vy[i], vz[i], a);
    last[i] = tick;
}

// Predictor for every particle: x += v q + a q^2 / 2, v = clamp(v + a q).
__global__ __launch_bounds__(256) void adaptive_drift_kernel(
    double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const float* __restrict__ ax, const float* __restrict__ ay, const float* __restrict__ az,
    size_t n, double q, double vmax, double box)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double hq = 0.5 * q * q;
    const double gx = ax[i], gy = ay[i], gz = az[i];
    double x = vx[i], y = vy[i], z = vz[i];
    px[i] = wrap_box(px[i] + x * q + gx * hq, box);
    py[i] = wrap_box(py[i] + y * q + gy * hq, box);
    pz[i] = wrap_box(pz[i] + z * q + gz * hq, box);
    x += gx * q;
    y += gy * q;
    z += gz * q;
    clamp3(x, y, z, vmax);
    vx[i] = x;
    vy[i] = y;
    vz[i] = z;
}

// Corrector of one particle active at `tick` (its new force is in acc).
__device__ inline void correct_one(
    double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const float* __restrict__ nx, const float* __restrict__ ny, const float* __restrict__ nz,
    float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz,
    uint32_t* __restrict__ level, uint32_t* __restrict__ last, size_t i, const AdaptiveParams& a,
    uint32_t tick)
{
    const double h = static_cast<double>(tick - last[i]) * a.quantum;
    const double cx = static_cast<double>(nx[i]) - bx[i];
    const double cy = static_cast<double>(ny[i]) - by[i];
    const double cz = static_cast<double>(nz[i]) - bz[i];
    const double hh = 0.5 * h * h;
    px[i] = wrap_box(px[i] + cx * hh, a.box);
    py[i] = wrap_box(py[i] + cy * hh, a.box);
    pz[i] = wrap_box(pz[i] + cz * hh, a.box);
    double x = vx[i] + cx * h, y = vy[i] + cy * h, z = vz[i] + cz * h;
    clamp3(x, y, z, a.max_speed);
    vx[i] = x;
    vy[i] = y;
    vz[i] = z;
    bx[i] = nx[i];
    by[i] = ny[i];
    bz[i] = nz[i];
    last[i] = tick;
    level[i] = adaptive_level(nx[i], ny[i], nz[i], x, y, z, a);
}

// Corrector for the particles active at `tick` (their new force is in acc).
__global__ __launch_bounds__(256) void adaptive_correct_kernel(
    double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const float* __restrict__ nx, const float* __restrict__ ny, const float* __restrict__ nz,
    float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz,
    uint32_t* __restrict__ level, uint32_t* __restrict__ last, size_t n, AdaptiveParams a,
    uint32_t tick, unsigned long long* __restrict__ count)
{
    // Grid-stride over the particles; the count of corrected particles is
    // summed per block and added with one atomic per block (a single counter
    // every wave updates would serialise them).
    __shared__ unsigned long long block_count[4];
    unsigned long long ndue = 0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const uint32_t lv = min(level[i], static_cast<uint32_t>(a.levels));
        const uint32_t cadence = 1u << (static_cast<uint32_t>(a.levels) - lv);
        if ((tick & (cadence - 1u)) != 0u)
            continue;
        ++ndue;
        correct_one(px, py, pz, vx, vy, vz, nx, ny, nz, bx, by, bz, level, last, i, a, tick);
    }
    if (count) {
        for (int o = 32; o > 0; o >>= 1)
            ndue += __shfl_xor(ndue, o, 64);
        if ((threadIdx.x & 63) == 0)
            block_count[threadIdx.x >> 6] = ndue;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long t =
                block_count[0] + block_count[1] + block_count[2] + block_count[3];
            if (t)
                atomicAdd(count, t);
        }
    }
}

// Smallest tick after t at which some particle is due: min over particles of
// (t / cadence + 1) * cadence (ticks count from a multiple of 2^L).
__global__ __launch_bounds__(256) void adaptive_next_due_kernel(const uint32_t* __restrict__ level,
                                                                size_t n, int L, uint32_t t,
                                                                uint32_t* __restrict__ out)
{
    // grid-stride; one atomicMin per block
    __shared__ uint32_t block_due[4];
    uint32_t due = 0xffffffffu;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const uint32_t lv = min(level[i], static_cast<uint32_t>(L));
        const uint32_t c = 1u << (static_cast<uint32_t>(L) - lv);
        due = min(due, (t / c + 1u) * c);
    }
    for (int o = 32; o > 0; o >>= 1)
        due = min(due, static_cast<uint32_t>(__shfl_xor(static_cast<int>(due), o, 64)));
    if ((threadIdx.x & 63) == 0)
        block_due[threadIdx.x >> 6] = due;
    __syncthreads();
    if (threadIdx.x == 0) {
        due = min(min(block_due[0], block_due[1]), min(block_due[2], block_due[3]));
        if (due != 0xffffffffu)
            atomicMin(out, due);
    }
}

__global__ __launch_bounds__(256) void gather_u32_kernel(const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ rank,
                                                         size_t n, uint32_t* __restrict__ out)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = src[rank[i]];
}

}  // namespace

void launch_adaptive_init(Device& d, const AdaptiveParams& a, uint32_t tick)
{
    d.aabb_parts = 0;
    if (d.n == 0)
        return;
    adaptive_init_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(
        d.acc[0], d.acc[1], d.acc[2], d.vel[0], d.vel[1], d.vel[2], d.aacc[0], d.aacc[1],
        d.aacc[2], d.alevel, d.atick, d.n, a, tick);
}

void launch_adaptive_drift(Device& d, const AdaptiveParams& a, double span)
{
    d.aabb_parts = 0;
    if (d.n == 0)
        return;
    adaptive_drift_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(
        d.pos[0], d.pos[1], d.pos[2], d.vel[0], d.vel[1], d.vel[2], d.aacc[0], d.aacc[1],
        d.aacc[2], d.n, span, a.max_speed, a.box);
}

void launch_adaptive_next_due(Device& d, const AdaptiveParams& a, uint32_t t, uint32_t* out)
{
    d.aabb_parts = 0;
    launch_fill_u32(out, 0xffffffffu, 1, d.stream);
    if (d.n == 0)
        return;
    adaptive_next_due_kernel<<<reduce_blocks(d.n), 256, 0, d.stream>>>(d.alevel, d.n, a.levels,
                                                                        t, out);
}

void launch_adaptive_correct(Device& d, const AdaptiveParams& a, uint32_t tick,
                             unsigned long long* count)
{
    d.aabb_parts = 0;
    if (d.n == 0)
        return;
    adaptive_correct_kernel<<<reduce_blocks(d.n), 256, 0, d.stream>>>(
        d.pos[0], d.pos[1], d.pos[2], d.vel[0], d.vel[1], d.vel[2], d.acc[0], d.acc[1], d.acc[2],
        d.aacc[0], d.aacc[1], d.aacc[2], d.alevel, d.atick, d.n, a, tick, count);
}

void launch_gather_u32(Device& d, const uint32_t* src_by_slot, uint32_t* out_by_caller)
{
    if (d.n == 0)
        return;
    gather_u32_kernel<<<blocks_for(d.n), 256, 0, d.stream>>>(src_by_slot, d.rank, d.n,
                                                             out_by_caller);
}

}  // namespace bh
