// Learned far-field correction (LFF) for octree_hip.
//
// A small MLP predicts, for one target and one accepted tree node, the
// residual a_exact - a_multipole of the monopole + quadrupole approximation,
// in units of the monopole magnitude M / r^2. It is applied to the nodes a
// coarse walk (theta ~ 0.9) accepts close to the target
// (size / distance >= band), where the multipole error concentrates.
//
// Network: 16 inputs -> 64 -> 64 -> 3, SiLU.
// Features of (target x, node of mass M, COM c, cell size s, traceless
// quadrupole Q, lattice centre g, bounding radius bmax, particle count n),
// with softening eps:
//   f0 = log2(n)/16, f1..f6 = Q/(M s^2) (xx xy xz yy yz zz),
//   f7..f9 = (x - c)/s, f10 = |x - c|/s, f11..f13 = (c - g)/s, f14 = bmax/s,
//   f15 = eps^2 / (|x - c|^2 + eps^2).
// The residual is taken against the walk's second-order softened expansion
// (monopole, traceless quadrupole and the eps^2 trs term). Only nodes of at
// least min_count (>= 2) particles are corrected: below a few tens of
// particles the residual is the shot noise of the node's own particles, which
// no node-level feature predicts; the walk of the corrected mode opens such
// near nodes instead.
// Parameters, FP32, packed: W1[16][64], b1[64], W2[64][64], b2[64], W3[64][3],
// b3[3] (5443 floats); the output scale is folded into W3/b3.
//
// Kernels:
//   lff_tuple_kernel   : per recorded (group, node) and target: features and
//                        the exact residual (direct sum over the node's
//                        particles, FP32) -> training tuples
//   lff_grad_kernel    : FP32 forward + backward of the MSE loss over a
//                        minibatch; per-workgroup partial gradients
//   lff_adam_kernel    : reduces partial gradients, Adam update
//   lff_pack_kernel    : weights -> BF16 MFMA A-operand fragments
//   lff_apply_kernel   : BF16 MFMA inference (FP32 accumulate) over each
//                        group's recorded nodes; corrections added to acc
//   lff_forward_kernel : the same MFMA inference on explicit features
//
// MFMA mapping (v_mfma_f32_16x16x16_bf16). Layers are computed transposed,
// H_next^T = W^T H^T: the weights are the A operand (rows = output units),
// activations the B operand (columns = 16 target-node pairs). In the
// 16x16x16 layouts lane l holds A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and
// C[4(l/16)+j][l%16] (j = 0..3), so a layer's C fragments are exactly the B
// fragments of the next layer's k-tiles and no data moves between lanes.
#include "bh_internal.h"

namespace bh {
namespace {

constexpr int NF = kLffInputs;   // 16
constexpr int NH = kLffHidden;   // 64
constexpr int NO = kLffOutputs;  // 3
constexpr int NP = kLffParams;   // 5443
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NF * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NO;
static_assert(OFF_B3 + NO == NP, "parameter layout");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16v4 __attribute__((ext_vector_type(4)));

// Training arithmetic (FP32, correctly rounded division).
__device__ inline float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ inline float siluf(float z) { return z * sigmoidf(z); }
__device__ inline float dsiluf(float z)
{
    const float s = sigmoidf(z);
    return s * (1.0f + z * (1.0f - s));
}
// Inference activation: one v_exp_f32 and one v_rcp_f32.
__device__ inline float silu_fast(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ inline bf16x4 pack4(float a, float b, float c, float d)
{
    bf16v4 v;
    v[0] = static_cast<__bf16>(a);
    v[1] = static_cast<__bf16>(b);
    v[2] = static_cast<__bf16>(c);
    v[3] = static_cast<__bf16>(d);
    return __builtin_bit_cast(bf16x4, v);
}

__device__ inline uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Node part of the features, shared by every target of the node (s = the
// cell size, cnt = the particle count).
struct NodeFeat {
    float f0, q[6], c[3], b, inv_s;
};

__device__ inline NodeFeat node_feat(const WalkNode& w, float s, uint32_t cnt)
{
    NodeFeat n;
    n.inv_s = 1.0f / s;
    n.f0 = log2f(static_cast<float>(cnt > 0 ? cnt : 1)) * (1.0f / 16.0f);
    const float qn = w.mass > 0.0f ? 1.0f / (w.mass * s * s) : 0.0f;
    n.q[0] = w.qxx * qn;
    n.q[1] = w.qxy * qn;
    n.q[2] = w.qxz * qn;
    n.q[3] = w.qyy * qn;
    n.q[4] = w.qyz * qn;
    n.q[5] = (-w.qxx - w.qyy) * qn;
    n.c[0] = w.comx * n.inv_s;
    n.c[1] = w.comy * n.inv_s;
    n.c[2] = w.comz * n.inv_s;
    n.b = w.bmax * n.inv_s;
    return n;
}

// All 16 features of (node, target at r = x - c), d2 = |r|^2.
__device__ inline void pair_features(const NodeFeat& n, float rx, float ry, float rz, float d2,
                                     float eps2, float F[NF])
{
    F[0] = n.f0;
    for (int k = 0; k < 6; ++k)
        F[1 + k] = n.q[k];
    F[7] = rx * n.inv_s;
    F[8] = ry * n.inv_s;
    F[9] = rz * n.inv_s;
    F[10] = sqrtf(d2) * n.inv_s;
    F[11] = n.c[0];
    F[12] = n.c[1];
    F[13] = n.c[2];
    F[14] = n.b;
    F[15] = d2 + eps2 > 0.0f ? eps2 / (d2 + eps2) : 0.0f;
}

// Second-order softened expansion of node w at offset d = c - x (the walk's
// formula, FP32).
__device__ inline void quad_accel(const WalkNode& w, float dx, float dy, float dz, float eps2,
                                  float& ax, float& ay, float& az)
{
    const float r2 = dx * dx + dy * dy + dz * dz + eps2;
    const float ri = __frsqrt_rn(r2);
    const float ri2 = ri * ri;
    const float ri3 = ri * ri2;
    const float qzz = -w.qxx - w.qyy;
    const float qx = w.qxx * dx + w.qxy * dy + w.qxz * dz;
    const float qy = w.qxy * dx + w.qyy * dy + w.qyz * dz;
    const float qz = w.qxz * dx + w.qyz * dy + qzz * dz;
    const float dqd = dx * qx + dy * qy + dz * qz;
    const float ri5 = ri3 * ri2;
    const float s = w.mass * ri3 + 2.5f * (dqd - eps2 * w.trs) * ri5 * ri2;
    ax = s * dx - ri5 * qx;
    ay = s * dy - ri5 * qy;
    az = s * dz - ri5 * qz;
}

__device__ inline void group_targets(const uint4& gr, const WalkNode* __restrict__ nodes,
                                     const float4* __restrict__ spart,
                                     const uint32_t* __restrict__ sleaf, float ux, float uy,
                                     float uz, int lane, int& rcx, int& rcy, int& rcz, float& tx,
                                     float& ty, float& tz)
{
    const int gnode = static_cast<int>(gr.z);
    rcx = nodes[gnode].cx;
    rcy = nodes[gnode].cy;
    rcz = nodes[gnode].cz;
    tx = ty = tz = 0.0f;
    if (lane < static_cast<int>(gr.y)) {
        const int s = static_cast<int>(gr.x) + lane;
        const float4 p = spart[s];
        const int lf = static_cast<int>(sleaf[s]);
        tx = rel_pos(nodes[lf].cx - rcx, ux, p.x);
        ty = rel_pos(nodes[lf].cy - rcy, uy, p.y);
        tz = rel_pos(nodes[lf].cz - rcz, uz, p.z);
    }
}

// ---- training tuples -------------------------------------------------------
// Every group_stride-th group; each (target, recorded node) pair in the band
// is kept when its hash falls at or below keep (a uniform subsample). With
// X == nullptr the kernel only counts the pairs (no direct sums).
__global__ __launch_bounds__(64) void lff_tuple_kernel(
    const WalkNode* __restrict__ nodes, const uint32_t* __restrict__ nstart,
    const uint32_t* __restrict__ ncount, const float4* __restrict__ spart,
    const uint32_t* __restrict__ sleaf, const RootGeom* __restrict__ geom,
    const uint4* __restrict__ groups, const int* __restrict__ info,
    const uint32_t* __restrict__ lists, const uint32_t* __restrict__ list_n, int list_cap,
    int list_groups, int group_stride, float band, float eps2, uint32_t min_count, uint32_t keep,
    uint32_t seed, float* __restrict__ X, float* __restrict__ Y,
    unsigned int* __restrict__ counter, unsigned int max_tuples)
{
    const int ngroups = min(info[kInfoGroups], list_groups);
    const float ux = geom->unit_f[0], uy = geom->unit_f[1], uz = geom->unit_f[2];
    const float wfmax = fmaxf(fmaxf(geom->width_f[0], geom->width_f[1]), geom->width_f[2]);
    const int lane = threadIdx.x;
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int g = blockIdx.x * group_stride; g < ngroups; g += gridDim.x * group_stride) {
        const uint4 gr = groups[g];
        int rcx, rcy, rcz;
        float tx, ty, tz;
        group_targets(gr, nodes, spart, sleaf, ux, uy, uz, lane, rcx, rcy, rcz, tx, ty, tz);
        const bool active = lane < static_cast<int>(gr.y);
        const uint32_t nl = min(list_n[g], static_cast<uint32_t>(list_cap));
        for (uint32_t k = 0; k < nl; ++k) {
            const int node = static_cast<int>(lists[static_cast<size_t>(g) * list_cap + k]);
            const WalkNode w = nodes[node];
            const uint32_t a = nstart[node], cnt = ncount[node];
            if (cnt < min_count)
                continue;  // uniform across the wave
            const float size = cell_size(wfmax, static_cast<int>(meta_level(w.meta)));
            const float cx = rel_pos(w.cx - rcx, ux, w.comx);
            const float cy = rel_pos(w.cy - rcy, uy, w.comy);
            const float cz = rel_pos(w.cz - rcz, uz, w.comz);
            const float rx = tx - cx, ry = ty - cy, rz = tz - cz;
            const float d2 = rx * rx + ry * ry + rz * rz;
            const uint32_t h = mix32(static_cast<uint32_t>(g) * 0x9e3779b1u ^
                                     mix32(static_cast<uint32_t>(node) + seed) ^
                                     (static_cast<uint32_t>(lane) << 24));
            const bool take = active && size * size >= band * band * d2 && h <= keep;
            const unsigned long long tm = __ballot(take);
            if (tm == 0ull)
                continue;
            unsigned int base = 0;
            if (lane == 0)
                base = atomicAdd(counter, static_cast<unsigned int>(__popcll(tm)));
            base = __shfl(base, 0, 64);
            if (!X)
                continue;
            float ex = 0.0f, ey = 0.0f, ez = 0.0f;
            for (uint32_t j = a; j < a + cnt; ++j) {
                const float4 q = spart[j];
                const int lf = static_cast<int>(sleaf[j]);
                const float px = rel_pos(nodes[lf].cx - rcx, ux, q.x);
                const float py = rel_pos(nodes[lf].cy - rcy, uy, q.y);
                const float pz = rel_pos(nodes[lf].cz - rcz, uz, q.z);
                const float dx = px - tx, dy = py - ty, dz = pz - tz;
                const float r2 = dx * dx + dy * dy + dz * dz + eps2;
                const float ri = __frsqrt_rn(r2);
                const float sc = r2 > 0.0f ? q.w * ri * ri * ri : 0.0f;
                ex += sc * dx;
                ey += sc * dy;
                ez += sc * dz;
            }
            const unsigned int idx = base + static_cast<unsigned int>(__popcll(tm & lt));
            if (take && idx < max_tuples) {
                float qx, qy, qz;
                quad_accel(w, cx - tx, cy - ty, cz - tz, eps2, qx, qy, qz);
                const float mono = w.mass / (d2 + eps2);
                float F[NF];
                pair_features(node_feat(w, size, cnt), rx, ry, rz, d2, eps2, F);
                for (int f = 0; f < NF; ++f)
                    X[static_cast<size_t>(idx) * NF + f] = F[f];
                Y[static_cast<size_t>(idx) * NO + 0] = (ex - qx) / mono;
                Y[static_cast<size_t>(idx) * NO + 1] = (ey - qy) / mono;
                Y[static_cast<size_t>(idx) * NO + 2] = (ez - qz) / mono;
            }
        }
    }
}

// ---- training: FP32 forward/backward -----------------------------------------
// One workgroup = one wave; lane i handles sample i of a 64-sample chunk.
// Gradients of the mean squared error over `nbatch` samples; each workgroup
// writes its partial gradient (and partial loss) for the samples it handled.
__global__ __launch_bounds__(64) void lff_grad_kernel(
    const float* __restrict__ X, const float* __restrict__ Y, const uint32_t* __restrict__ order,
    size_t first, uint32_t nbatch, float yscale, const float* __restrict__ P,
    float* __restrict__ partial, float* __restrict__ loss_partial)
{
    __shared__ float sP[NP];
    __shared__ float sX[64][NF + 1];
    __shared__ float sZ1[64][NH + 1];
    __shared__ float sH1[64][NH + 1];
    __shared__ float sZ2[64][NH + 1];
    __shared__ float sD1[64][NH + 1];
    __shared__ float sD2[64][NH + 1];
    __shared__ float sDY[64][NO];
    const int t = threadIdx.x;
    for (int k = t; k < NP; k += 64)
        sP[k] = P[k];
    float gW1[NF], gW2[NH], gW3[NO], gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, loss = 0.0f;
    for (int k = 0; k < NF; ++k)
        gW1[k] = 0.0f;
    for (int k = 0; k < NH; ++k)
        gW2[k] = 0.0f;
    for (int k = 0; k < NO; ++k)
        gW3[k] = 0.0f;
    const float inv_b = 1.0f / static_cast<float>(nbatch);
    __syncthreads();
    const uint32_t nchunks = (nbatch + 63) / 64;
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t bi = ch * 64 + t;
        const bool valid = bi < nbatch;
        const size_t src = valid ? static_cast<size_t>(order[first + bi]) : 0;
        float x[NF], y[NO];
        for (int k = 0; k < NF; ++k) {
            x[k] = valid ? X[src * NF + k] : 0.0f;
            sX[t][k] = x[k];
        }
        for (int k = 0; k < NO; ++k)
            y[k] = valid ? Y[src * NO + k] * yscale : 0.0f;
        // forward (activations of this lane's sample live in its LDS rows)
        for (int a = 0; a < NH; ++a) {
            float z = sP[OFF_B1 + a];
            for (int k = 0; k < NF; ++k)
                z += x[k] * sP[OFF_W1 + k * NH + a];
            sZ1[t][a] = z;
            sH1[t][a] = siluf(z);
        }
        float out[NO];
        for (int k = 0; k < NO; ++k)
            out[k] = sP[OFF_B3 + k];
        for (int h = 0; h < NH; ++h) {
            float z = sP[OFF_B2 + h];
            for (int a = 0; a < NH; ++a)
                z += sH1[t][a] * sP[OFF_W2 + a * NH + h];
            sZ2[t][h] = z;
            const float h2 = siluf(z);
            for (int k = 0; k < NO; ++k)
                out[k] += h2 * sP[OFF_W3 + h * NO + k];
        }
        // backward
        float dy[NO];
        for (int k = 0; k < NO; ++k) {
            const float e = out[k] - y[k];
            dy[k] = valid ? 2.0f * e * inv_b : 0.0f;
            loss += valid ? e * e : 0.0f;
            sDY[t][k] = dy[k];
        }
        for (int h = 0; h < NH; ++h) {
            float g = 0.0f;
            for (int k = 0; k < NO; ++k)
                g += dy[k] * sP[OFF_W3 + h * NO + k];
            sD2[t][h] = g * dsiluf(sZ2[t][h]);
        }
        for (int a = 0; a < NH; ++a) {
            float g = 0.0f;
            for (int h = 0; h < NH; ++h)
                g += sD2[t][h] * sP[OFF_W2 + a * NH + h];
            sD1[t][a] = g * dsiluf(sZ1[t][a]);
        }
        __syncthreads();
        // weight gradients: lane t owns output column t of W1 and W2, row t of W3
        for (int i = 0; i < 64; ++i) {
            const float d1 = sD1[i][t];
            const float d2 = sD2[i][t];
            gb1 += d1;
            gb2 += d2;
            for (int k = 0; k < NF; ++k)
                gW1[k] += sX[i][k] * d1;
            for (int a = 0; a < NH; ++a)
                gW2[a] += sH1[i][a] * d2;
            const float h2 = siluf(sZ2[i][t]);
            for (int k = 0; k < NO; ++k)
                gW3[k] += h2 * sDY[i][k];
            if (t < NO)
                gb3 += sDY[i][t];
        }
        __syncthreads();
    }
    float* out = partial + static_cast<size_t>(blockIdx.x) * NP;
    for (int k = 0; k < NF; ++k)
        out[OFF_W1 + k * NH + t] = gW1[k];
    out[OFF_B1 + t] = gb1;
    for (int a = 0; a < NH; ++a)
        out[OFF_W2 + a * NH + t] = gW2[a];
    out[OFF_B2 + t] = gb2;
    for (int k = 0; k < NO; ++k)
        out[OFF_W3 + t * NO + k] = gW3[k];
    if (t < NO)
        out[OFF_B3 + t] = gb3;
    // partial loss (sum of squared errors of this workgroup)
    for (int o = 32; o > 0; o >>= 1)
        loss += __shfl_xor(loss, o, 64);
    if (t == 0)
        loss_partial[blockIdx.x] = loss;
}

// Sums the partial gradients into grad (when non-null) and applies Adam when
// lr > 0. step is 1-based.
__global__ __launch_bounds__(256) void lff_adam_kernel(const float* __restrict__ partial, int nb,
                                                       float* __restrict__ P,
                                                       float* __restrict__ m,
                                                       float* __restrict__ v,
                                                       float* __restrict__ grad, float lr,
                                                       float b1, float b2, int step)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= NP)
        return;
    float g = 0.0f;
    for (int b = 0; b < nb; ++b)
        g += partial[static_cast<size_t>(b) * NP + k];
    if (grad)
        grad[k] = g;
    if (lr > 0.0f) {
        const float mk = b1 * m[k] + (1.0f - b1) * g;
        const float vk = b2 * v[k] + (1.0f - b2) * g * g;
        m[k] = mk;
        v[k] = vk;
        const float mh = mk / (1.0f - powf(b1, static_cast<float>(step)));
        const float vh = vk / (1.0f - powf(b2, static_cast<float>(step)));
        P[k] -= lr * mh / (sqrtf(vh) + 1e-8f);
    }
}

__global__ void lff_loss_sum_kernel(const float* __restrict__ lp, int nb, float* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b)
        s += lp[b];
    *out += s;
}

// ---- inference: BF16 MFMA -----------------------------------------------------
// Fragment tables: frag[f * 64 + lane], f = 0..3 W1^T (m-tiles), 4..19 W2^T
// (4 * mt + kt), 20..23 W3^T (k-tiles; output rows >= 3 are zero).
// bias = b1[64], b2[64], b3 padded to 16.
__global__ __launch_bounds__(64) void lff_pack_kernel(const float* __restrict__ P,
                                                      bf16x4* __restrict__ frag,
                                                      float* __restrict__ bias)
{
    const int l = threadIdx.x;
    const int r = l & 15, q = l >> 4;
    for (int mt = 0; mt < 4; ++mt) {
        float w[4];
        for (int j = 0; j < 4; ++j)
            w[j] = P[OFF_W1 + (4 * q + j) * NH + 16 * mt + r];
        frag[mt * 64 + l] = pack4(w[0], w[1], w[2], w[3]);
    }
    for (int mt = 0; mt < 4; ++mt)
        for (int kt = 0; kt < 4; ++kt) {
            float w[4];
            for (int j = 0; j < 4; ++j)
                w[j] = P[OFF_W2 + (16 * kt + 4 * q + j) * NH + 16 * mt + r];
            frag[(4 + 4 * mt + kt) * 64 + l] = pack4(w[0], w[1], w[2], w[3]);
        }
    for (int kt = 0; kt < 4; ++kt) {
        float w[4];
        for (int j = 0; j < 4; ++j)
            w[j] = r < NO ? P[OFF_W3 + (16 * kt + 4 * q + j) * NO + r] : 0.0f;
        frag[(20 + kt) * 64 + l] = pack4(w[0], w[1], w[2], w[3]);
    }
    bias[l] = P[OFF_B1 + l];
    bias[64 + l] = P[OFF_B2 + l];
    if (l < 16)
        bias[128 + l] = l < NO ? P[OFF_B3 + l] : 0.0f;
}

struct MlpRegs {
    bf16x4 w1[4];
    bf16x4 w2[16];
    bf16x4 w3[4];
    f32x4 b1[4];
    f32x4 b2[4];
    f32x4 b3;
};

__device__ inline void load_mlp(MlpRegs& R, const bf16x4* __restrict__ frag,
                                const float* __restrict__ bias, int lane)
{
    const int q = lane >> 4;
    for (int f = 0; f < 4; ++f)
        R.w1[f] = frag[f * 64 + lane];
    for (int f = 0; f < 16; ++f)
        R.w2[f] = frag[(4 + f) * 64 + lane];
    for (int f = 0; f < 4; ++f)
        R.w3[f] = frag[(20 + f) * 64 + lane];
    for (int mt = 0; mt < 4; ++mt)
        for (int j = 0; j < 4; ++j) {
            R.b1[mt][j] = bias[16 * mt + 4 * q + j];
            R.b2[mt][j] = bias[64 + 16 * mt + 4 * q + j];
        }
    for (int j = 0; j < 4; ++j)
        R.b3[j] = bias[128 + 4 * q + j];
}

// x = this lane's 4 input features (4q..4q+3 of pair l%16). Returns the
// output fragment: lanes 0..15 hold outputs 0..3 (0..2 used) of pair l%16.
__device__ inline f32x4 mlp_mfma(const MlpRegs& R, const bf16x4 x)
{
    bf16x4 h[4];
    for (int mt = 0; mt < 4; ++mt) {
        const f32x4 c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w1[mt], x, R.b1[mt], 0, 0, 0);
        h[mt] = pack4(silu_fast(c[0]), silu_fast(c[1]), silu_fast(c[2]), silu_fast(c[3]));
    }
    bf16x4 h2[4];
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 c = R.b2[mt];
        for (int kt = 0; kt < 4; ++kt)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w2[4 * mt + kt], h[kt], c, 0, 0, 0);
        h2[mt] = pack4(silu_fast(c[0]), silu_fast(c[1]), silu_fast(c[2]), silu_fast(c[3]));
    }
    f32x4 c = R.b3;
    for (int kt = 0; kt < 4; ++kt)
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w3[kt], h2[kt], c, 0, 0, 0);
    return c;
}

__global__ __launch_bounds__(64) void lff_forward_kernel(const float* __restrict__ X, uint32_t n,
                                                         const bf16x4* __restrict__ frag,
                                                         const float* __restrict__ bias,
                                                         float* __restrict__ Y)
{
    const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
    MlpRegs R;
    load_mlp(R, frag, bias, lane);
    for (uint32_t base = blockIdx.x * 16; base < n; base += gridDim.x * 16) {
        const uint32_t i = base + static_cast<uint32_t>(r);
        float f[4];
        for (int j = 0; j < 4; ++j)
            f[j] = i < n ? X[static_cast<size_t>(i) * NF + 4 * q + j] : 0.0f;
        const f32x4 c = mlp_mfma(R, pack4(f[0], f[1], f[2], f[3]));
        if (q == 0 && i < n)
            for (int k = 0; k < NO; ++k)
                Y[static_cast<size_t>(i) * NO + k] = c[k];
    }
}

// Per group: for every recorded node and every target with
// size / |x - c| >= band, add MLP(features) * M / (|x - c|^2 + eps^2) to the
// target's acceleration; then apply the acceleration clamp.
__global__ __launch_bounds__(64) void lff_apply_kernel(
    const WalkNode* __restrict__ nodes, const uint32_t* __restrict__ ncount,
    const float4* __restrict__ spart,
    const uint32_t* __restrict__ sleaf, const RootGeom* __restrict__ geom,
    const uint4* __restrict__ groups, const int* __restrict__ info,
    const uint32_t* __restrict__ lists, const uint32_t* __restrict__ list_n, int list_cap,
    int list_groups, const bf16x4* __restrict__ frag, const float* __restrict__ bias, float band,
    float eps2, uint32_t min_count, float max_acc, float* __restrict__ accx, float* __restrict__ accy,
    float* __restrict__ accz)
{
    __shared__ float sT[64][3];
    const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
    const float ux = geom->unit_f[0], uy = geom->unit_f[1], uz = geom->unit_f[2];
    const float wfmax = fmaxf(fmaxf(geom->width_f[0], geom->width_f[1]), geom->width_f[2]);
    const float band2 = band * band;
    MlpRegs R;
    load_mlp(R, frag, bias, lane);
    const int ngroups = info[kInfoGroups];
    for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const uint4 gr = groups[g];
        int rcx, rcy, rcz;
        float tx, ty, tz;
        group_targets(gr, nodes, spart, sleaf, ux, uy, uz, lane, rcx, rcy, rcz, tx, ty, tz);
        __syncthreads();
        sT[lane][0] = tx;
        sT[lane][1] = ty;
        sT[lane][2] = tz;
        __syncthreads();
        const int nact = static_cast<int>(gr.y);
        float cx[4] = {0, 0, 0, 0}, cy[4] = {0, 0, 0, 0}, cz[4] = {0, 0, 0, 0};
        const uint32_t nl = g < list_groups ? min(list_n[g], static_cast<uint32_t>(list_cap)) : 0u;
        for (uint32_t k = 0; k < nl; ++k) {
            const int node = static_cast<int>(lists[static_cast<size_t>(g) * list_cap + k]);
            const WalkNode w = nodes[node];
            const uint32_t cnt = ncount[node];
            if (cnt < min_count)
                continue;  // uniform across the wave
            const float size = cell_size(wfmax, static_cast<int>(meta_level(w.meta)));
            const float ncx = rel_pos(w.cx - rcx, ux, w.comx);
            const float ncy = rel_pos(w.cy - rcy, uy, w.comy);
            const float ncz = rel_pos(w.cz - rcz, uz, w.comz);
            const NodeFeat nf = node_feat(w, size, cnt);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                if (16 * ct < nact) {  // uniform across the wave
                    const int p = 16 * ct + r;
                    const float rx = sT[p][0] - ncx, ry = sT[p][1] - ncy, rz = sT[p][2] - ncz;
                    const float d2 = rx * rx + ry * ry + rz * rz;
                    const bool in = p < nact && size * size >= band2 * d2;
                    if (__ballot(in) != 0ull) {  // uniform: skip tiles with no pair in the band
                        float F[NF];
                        pair_features(nf, rx, ry, rz, d2, eps2, F);
                        float v[4];
                        for (int j = 0; j < 4; ++j)  // lane quarter q holds features 4q..4q+3
                            v[j] = q == 0 ? F[j] : q == 1 ? F[4 + j] : q == 2 ? F[8 + j] : F[12 + j];
                        const f32x4 c = mlp_mfma(R, pack4(v[0], v[1], v[2], v[3]));
                        if (q == 0 && in) {
                            const float mono = w.mass / (d2 + eps2);
                            cx[ct] += c[0] * mono;
                            cy[ct] += c[1] * mono;
                            cz[ct] += c[2] * mono;
                        }
                    }
                }
            }
        }
        if (q == 0) {
            for (int ct = 0; ct < 4; ++ct) {
                const int p = 16 * ct + r;
                if (p < nact) {
                    const size_t s = static_cast<size_t>(gr.x) + static_cast<size_t>(p);
                    float ax = accx[s] + cx[ct], ay = accy[s] + cy[ct], az = accz[s] + cz[ct];
                    const float mag = sqrtf(ax * ax + ay * ay + az * az);
                    if (max_acc > 0.0f && mag > max_acc && mag > 1.0e-6f) {
                        const float kk = max_acc / mag;
                        ax *= kk;
                        ay *= kk;
                        az *= kk;
                    }
                    accx[s] = ax;
                    accy[s] = ay;
                    accz[s] = az;
                }
            }
        }
    }
}

}  // namespace

void launch_lff_tuples(Device& d, const LffState& L, int group_stride, float band, float eps2,
                       int min_count, uint32_t keep, uint32_t seed, bool count_only,
                       unsigned int* counter, unsigned int max_tuples)
{
    const unsigned grid = static_cast<unsigned>(std::max<size_t>(1, d.stack_waves));
    lff_tuple_kernel<<<grid, 64, 0, d.stream>>>(
        d.wnode, d.nstart, d.ncount, d.spart, d.sleaf, d.geom, d.groups, d.level_info, L.lists,
        L.list_n, L.list_cap, static_cast<int>(L.list_groups), group_stride, band, eps2,
        static_cast<uint32_t>(std::max(2, min_count)), keep, seed, count_only ? nullptr : L.X,
        L.Y, counter, max_tuples);
}

void launch_lff_grad(Device& d, const LffState& L, const float* X, const float* Y,
                     const uint32_t* order, size_t first, uint32_t nbatch, float yscale,
                     const float* P, int nblocks)
{
    lff_grad_kernel<<<nblocks, 64, 0, d.stream>>>(X, Y, order, first, nbatch, yscale, P,
                                                  L.partial, L.loss_partial);
}

void launch_lff_adam(Device& d, const LffState& L, int nblocks, float* P, float* grad, float lr,
                     int step)
{
    lff_adam_kernel<<<(NP + 255) / 256, 256, 0, d.stream>>>(L.partial, nblocks, P, L.adam_m,
                                                           L.adam_v, grad, lr, 0.9f, 0.999f, step);
    lff_loss_sum_kernel<<<1, 64, 0, d.stream>>>(L.loss_partial, nblocks, L.loss_acc);
}

void launch_lff_pack(Device& d, const LffState& L)
{
    lff_pack_kernel<<<1, 64, 0, d.stream>>>(L.params, reinterpret_cast<bf16x4*>(L.frag), L.bias);
}

void launch_lff_forward(Device& d, const LffState& L, const float* X, uint32_t n, float* Y)
{
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 15) / 16, 4096));
    lff_forward_kernel<<<std::max(1u, grid), 64, 0, d.stream>>>(
        X, n, reinterpret_cast<const bf16x4*>(L.frag), L.bias, Y);
}

void launch_lff_apply(Device& d, const LffState& L, float band, float eps2, int min_count,
                      float max_acc)
{
    const unsigned grid = static_cast<unsigned>(std::max<size_t>(1, d.stack_waves));
    lff_apply_kernel<<<grid, 64, 0, d.stream>>>(
        d.wnode, d.ncount, d.spart, d.sleaf, d.geom, d.groups, d.level_info, L.lists, L.list_n, L.list_cap,
        static_cast<int>(L.list_groups), reinterpret_cast<const bf16x4*>(L.frag), L.bias, band,
        eps2, static_cast<uint32_t>(std::max(2, min_count)), max_acc, d.acc[0], d.acc[1],
        d.acc[2]);
}

}  // namespace bh
