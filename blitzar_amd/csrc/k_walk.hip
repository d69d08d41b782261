// Wavefront-grouped Barnes-Hut walk for octree_hip.
//
// Reference semantics (per target, engine/physics/octree/traversal/
// OctreeTraversal.cpp:129-222 and engine/physics/octree/cuda/force/
// OctGpu.inl:100-228): skip empty nodes; a node is accepted as a single source
// at its centre of mass when the target is not inside it and
//   size / sqrt(|com - x|^2 + eps^2) < theta      (criterion "com"), or
//   size / dist(x, box)               < theta      (criterion "bounds");
// octree_hip adds criterion "extent": bmax / sqrt(|com - x|^2 + eps^2) < theta,
// bmax the radius about the COM that encloses the node's particles (the
// multipole truncation error scales with bmax / distance, not cell size);
// otherwise it is opened; particles of opened leaves are summed directly
// (pair force engine/physics/octree/force/OctreeForce.cpp:23-57); the sum is
// clamped to max_acceleration (engine/physics/octree/force/OctForces.cpp:128-133).
//
// MI355X mapping. One 64-lane workgroup (one wavefront) owns a group of at
// most 64 targets (one target per lane) -- the particles of one tree node, so
// the group is spatially compact -- and walks the tree once for the whole
// group. A node is accepted only when the criterion holds for every
// point of the group's bounding box (distance measured from the box, which is
// conservative), and only when no target lies within the node's bounding
// sphere (bmax) -- the group form of the reference's "contains self" test.
// Accepted nodes (monopole + traceless quadrupole) and the particles of opened
// leaves are staged in LDS lists and evaluated by all 64 lanes with broadcast
// LDS reads. Traversal keeps a LIFO of node indices in global memory, popped
// 64 at a time. Groups are handed out by an atomic ticket; every workgroup
// exits when the ticket passes the group count.
//
// Precision: positions in the walk are FP32 offsets from the exact lattice
// centre of the group's node: float(c_node - c_ref) * unit + offset,
// with integer lattice coordinates c, so the walk itself uses no FP64.
//
// Second order (multipole_order 2) is the expansion of the Plummer-softened
// kernel g(r) = (r^2 + eps^2)^-1/2 about the COM, d = com - x, R^2 = |d|^2 + eps^2:
//   a = M d / R^3 - Q d / R^5 + 5/2 (d.Q d - eps^2 trs) d / R^7,
// Q the traceless quadrupole, trs = sum m |y|^2 (the eps^2 trs term is what
// the traceless Q cannot carry once eps > 0). multipole_order 1 is the
// reference's softened monopole.
//
// Cost model for the roofline: 20 FLOP per particle-particle and 50 FLOP per
// particle-cell (monopole + quadrupole) interaction.
#include "bh_internal.h"

namespace bh {
namespace {

// LDS list capacities (cells flush at kCellCap - 64, particles at kPartCap);
// together with the leaf descriptors they set the LDS per wavefront and so
// the walk's residency: 96 and 96 take 6.5 KB (24 waves per CU by LDS, 20 by
// VGPRs) where 128 and 128 took 8.3 KB (18 waves per CU); the relative walk
// is 2 % faster at 8M.
#ifndef BH_WALK_CELL_CAP
#define BH_WALK_CELL_CAP 96
#endif
#ifndef BH_WALK_PART_CAP
#define BH_WALK_PART_CAP 96
#endif
constexpr int kCellCap = BH_WALK_CELL_CAP;
constexpr int kPartCap = BH_WALK_PART_CAP;

// Group tickets. On large walks (at least kQueueMinGroupsPerWave groups per
// resident wave) the walk order is cut into chunks of kTicketChunk
// consecutive groups, dealt round-robin to kQueues counters; a wave draws
// from the queue of its XCD, so the groups one XCD walks at a time are
// neighbours in space and share tree nodes in that XCD's L2. Smaller walks
// use one counter: with few groups per wave, the per-XCD split costs more in
// balance than it gains in locality.
#ifndef BH_WALK_TICKET_QUEUES
#define BH_WALK_TICKET_QUEUES 8
#endif
#ifndef BH_WALK_TICKET_CHUNK
#define BH_WALK_TICKET_CHUNK 16
#endif
constexpr int kQueues = BH_WALK_TICKET_QUEUES;
constexpr unsigned kTicketChunk = BH_WALK_TICKET_CHUNK;
constexpr int kQueueMinGroupsPerWave = 16;
static_assert(kQueues >= 1 && kQueues <= kTicketQueues, "ticket queues");

// position in the walk order of the k-th ticket of queue q (of nq)
__device__ inline int ticket_position(unsigned k, int q, int nq)
{
    return static_cast<int>(((k / kTicketChunk) * static_cast<unsigned>(nq) +
                             static_cast<unsigned>(q)) * kTicketChunk + k % kTicketChunk);
}

// Group-box reductions, once per group. (Written with DPP row rotations and
// row broadcasts, each step waits on the one before; with the shuffles below
// other waves fill those waits, and the 1M-particle COM walk was 3 % faster.)
__device__ inline float wmin(float v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float wmax(float v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
// Inclusive scan over the 64 lanes (all active) in six DPP adds: shifts of
// 1, 2, 4 and 8 within each 16-lane row, then lane 15 of each row into the
// next (rows 1 and 3) and lane 31 into rows 2 and 3. Lanes a shift leaves
// without a source, and rows the row mask excludes, add the old value 0.
// (Written with lane shuffles, each step is an LDS permute and a wait.)
__device__ inline uint32_t winclusive(uint32_t v)
{
    int x = static_cast<int>(v);
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return static_cast<uint32_t>(x);
}

// the value of lane 63 (a wave-uniform scalar)
__device__ inline uint32_t wlast(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// TreePM short-range weight of the Newtonian pair force:
// w(r) = erfc(r / 2rs) + r / (rs sqrt(pi)) exp(-r^2 / 4rs^2)
// (engine/physics/octree/force/OctreeForce.cpp:49-57). With x = r / rs it is
// E(x) Q(x), E = exp(-x^2 / 4) and Q(x) = erfcx(x / 2) + x / sqrt(pi) smooth
// (erfcx the scaled complementary error function). Q is this build's
// degree-13 least-squares polynomial on [0, 8] (Chebyshev points, weight E),
// held at x = 8 beyond: the FP32 evaluation errs by at most 1.6e-7 in w for
// every x (checked against erfc on the host). One exp2 and 13 FMAs against
// about 50 instructions for erfcf and expf.
__constant__ float kTpmQ[14] = {
    1.000000000e+00f,  2.537409785e-07f, 2.499964684e-01f,  -9.401225299e-02f,
    3.119471297e-02f,  -9.308321401e-03f, 2.498052781e-03f, -5.898143281e-04f,
    1.174437930e-04f,  -1.865963713e-05f, 2.218408781e-06f, -1.821118900e-07f,
    9.095778353e-09f,  -2.065529969e-10f};

// w(x) and E(x) = exp(-x^2 / 4), x = r / rs
__device__ inline float tpm_weight_e(float x, float& E)
{
    const float xc = fminf(x, 8.0f);
    float q = kTpmQ[13];
#pragma unroll
    for (int k = 12; k >= 0; --k)
        q = fmaf(q, xc, kTpmQ[k]);
    E = __builtin_amdgcn_exp2f(-(x * x) * 0.3606737554f);  // log2(e) / 4
    return q * E;
}

__device__ inline float tpm_weight(float r, float inv_rs)
{
    float E;
    return tpm_weight_e(r * inv_rs, E);
}

// List evaluation. A lane takes entries j0, j0 + dj, ...: dj = 1 for groups
// of more than 32 targets; smaller groups give each target dj lanes (see
// walk_kernel), which split the list and are summed at the end of the group.
template <int ORDER, bool TPM, int DJ>
__device__ inline void eval_cells_dj(const float4* __restrict__ ca, const float4* __restrict__ cb,
                                     const float2* __restrict__ cc, int ncell, int j0, float tx,
                                     float ty, float tz, const WalkParams& P, float& ax, float& ay,
                                     float& az)
{
    constexpr int dj = DJ;  // a compile-time stride keeps the unrolled LDS reads on one base
    const float inv_rs = TPM ? 1.0f / P.rs : 0.0f;
    const float box = TPM ? P.box : 0.0f;
    const float ib = box > 0.0f ? 1.0f / box : 0.0f;
#pragma unroll 2
    for (int j = j0; j < ncell; j += dj) {
        const float4 a = ca[j];
        float dx = a.x - tx, dy = a.y - ty, dz = a.z - tz;
        if (TPM && box > 0.0f) {  // nearest image of this pair
            dx -= box * rintf(dx * ib);
            dy -= box * rintf(dy * ib);
            dz -= box * rintf(dz * ib);
        }
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, P.eps2)));
        const float ri = __frsqrt_rn(r2);
        const float ri2 = ri * ri;
        const float ri3 = ri * ri2;
        float wgt = 1.0f;
        if (TPM && ORDER < 2)
            wgt = tpm_weight(r2 * ri, inv_rs);
        if (TPM && ORDER >= 2) {
            // Second-order expansion of the short-range kernel g(R) d,
            // g = w(R) / R^3, about the COM (S = (Q + trs I) / 3 the raw
            // second moment, Q traceless):
            //   a = [M g + G2 (dQd + trs |d|^2) / 6 + 5/6 G1 trs] d + G1 / 3 Q d
            // with G1 = g'(R) / R, G2 = G1'(R) / R. For the erfc weight,
            // w'(R) = -R^2 E / (2 sqrt(pi) rs^3), E = exp(-R^2 / 4 rs^2), so
            //   G1 = -c / R^2 - 3 w / R^5,
            //   G2 = c (1 / (2 rs^2 R^2) + 5 / R^4) + 15 w / R^7,
            // c = E / (2 sqrt(pi) rs^3). With w = 1, c = 0 it is the softened
            // Newtonian expansion of the open walk.
            const float4 b = cb[j];
            const float2 c = cc[j];  // (qyz, trs)
            const float qx = fmaf(b.z, dz, fmaf(b.y, dy, b.x * dx));
            const float qy = fmaf(c.x, dz, fmaf(b.w, dy, b.y * dx));
            const float qz = fmaf(-(b.x + b.w), dz, fmaf(c.x, dy, b.z * dx));
            const float dqd = fmaf(dz, qz, fmaf(dy, qy, dx * qx));
            const float R = r2 * ri;
            float E;
            const float w = tpm_weight_e(R * inv_rs, E);
            const float cE = E * (0.28209479177387814f * inv_rs * inv_rs * inv_rs);
            const float ri4 = ri2 * ri2;
            const float wr5 = w * ri4 * ri;
            const float g = w * ri3;
            const float G1 = fmaf(-cE, ri2, -3.0f * wr5);
            const float G2 = fmaf(cE, fmaf(0.5f * inv_rs * inv_rs, ri2, 5.0f * ri4),
                                  15.0f * wr5 * ri2);
            const float d2 = r2 - P.eps2;
            const float s = fmaf(G2 * (1.0f / 6.0f), fmaf(c.y, d2, dqd),
                                 fmaf(a.w, g, (5.0f / 6.0f) * G1 * c.y));
            const float t = G1 * (1.0f / 3.0f);
            ax = fmaf(t, qx, fmaf(s, dx, ax));
            ay = fmaf(t, qy, fmaf(s, dy, ay));
            az = fmaf(t, qz, fmaf(s, dz, az));
        } else if (ORDER >= 2) {
            const float4 b = cb[j];
            const float2 c = cc[j];
            // b = (qxx, qxy, qxz, qyy), c = (qyz, eps^2 trs); qzz = -qxx - qyy
            const float qx = fmaf(b.z, dz, fmaf(b.y, dy, b.x * dx));
            const float qy = fmaf(c.x, dz, fmaf(b.w, dy, b.y * dx));
            const float qz = fmaf(-(b.x + b.w), dz, fmaf(c.x, dy, b.z * dx));
            const float dqd = fmaf(dz, qz, fmaf(dy, qy, dx * qx));
            // M/R^3 + 5/2 (dQd - eps^2 trs)/R^7 = (M + 5/2 (dQd - eps^2 trs)/R^4) / R^3
            const float u = (dqd - c.y) * ri2;
            const float s = fmaf(2.5f * u, ri2, a.w) * ri3 * wgt;
            const float t = ri3 * ri2 * wgt;
            ax = fmaf(-t, qx, fmaf(s, dx, ax));
            ay = fmaf(-t, qy, fmaf(s, dy, ay));
            az = fmaf(-t, qz, fmaf(s, dz, az));
        } else {
            const float s = a.w * ri3 * wgt;
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
        }
    }
}

template <bool TPM, int DJ>
__device__ inline void eval_parts_dj(const float4* __restrict__ pl, int np, int j0, float tx,
                                     float ty, float tz, const WalkParams& P, float& ax, float& ay,
                                     float& az)
{
    constexpr int dj = DJ;
    if (!TPM && P.eps2 > 0.0f && P.min_d2 <= P.eps2) {
        // every pair has r2 >= eps2 >= min_d2: the skip test can only drop a
        // coincident pair, whose contribution is exactly zero anyway
        const float eps2 = P.eps2;
#pragma unroll 4
        for (int j = j0; j < np; j += dj) {
            const float4 a = pl[j];
            const float dx = a.x - tx, dy = a.y - ty, dz = a.z - tz;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            const float ri = __frsqrt_rn(r2);
            const float s = a.w * ri * (ri * ri);
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
        }
        return;
    }
    const float inv_rs = TPM ? 1.0f / P.rs : 0.0f;
    const float box = TPM ? P.box : 0.0f;
    const float ib = box > 0.0f ? 1.0f / box : 0.0f;
#pragma unroll 4
    for (int j = j0; j < np; j += dj) {
        const float4 a = pl[j];
        float dx = a.x - tx, dy = a.y - ty, dz = a.z - tz;
        if (TPM && box > 0.0f) {  // nearest image of this pair
            dx -= box * rintf(dx * ib);
            dy -= box * rintf(dy * ib);
            dz -= box * rintf(dz * ib);
        }
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float r2 = d2 + P.eps2;
        const float ri = __frsqrt_rn(r2);
        bool use = r2 > P.min_d2;
        float wgt = 1.0f;
        if (TPM) {
            wgt = tpm_weight(r2 * ri, inv_rs);
            use = use && (P.rcut2 <= 0.0f || d2 <= P.rcut2);
        }
        const float s = use ? a.w * ri * ri * ri * wgt : 0.0f;
        ax += s * dx;
        ay += s * dy;
        az += s * dz;
    }
}

template <int ORDER, bool TPM>
__device__ inline void eval_cells(const float4* __restrict__ ca, const float4* __restrict__ cb,
                                  const float2* __restrict__ cc, int ncell, int j0, int dj,
                                  float tx, float ty, float tz, const WalkParams& P, float& ax,
                                  float& ay, float& az)
{
    if (dj == 1)
        eval_cells_dj<ORDER, TPM, 1>(ca, cb, cc, ncell, j0, tx, ty, tz, P, ax, ay, az);
    else if (dj == 2)
        eval_cells_dj<ORDER, TPM, 2>(ca, cb, cc, ncell, j0, tx, ty, tz, P, ax, ay, az);
    else
        eval_cells_dj<ORDER, TPM, 4>(ca, cb, cc, ncell, j0, tx, ty, tz, P, ax, ay, az);
}

template <bool TPM>
__device__ inline void eval_parts(const float4* __restrict__ pl, int np, int j0, int dj, float tx,
                                  float ty, float tz, const WalkParams& P, float& ax, float& ay,
                                  float& az)
{
    if (dj == 1)
        eval_parts_dj<TPM, 1>(pl, np, j0, tx, ty, tz, P, ax, ay, az);
    else if (dj == 2)
        eval_parts_dj<TPM, 2>(pl, np, j0, tx, ty, tz, P, ax, ay, az);
    else
        eval_parts_dj<TPM, 4>(pl, np, j0, tx, ty, tz, P, ax, ay, az);
}

// Group far field: the local (Taylor) expansion about the group centre of an
// accepted node's softened second-order field,
//   a(g + delta) = A + J delta + 1/2 H delta delta,
// r = COM - g, R^2 = |r|^2 + eps^2, S = (Q + trs I)/3:
//   A_i  = M r_i R^-3 + 7.5 (rSr) r_i R^-7 - 3 (Sr)_i R^-5 - 1.5 trs r_i R^-5
//   J_il = M (3 r_i r_l R^-5 - d_il R^-3)
//          + (52.5 rSr R^-9 - 7.5 trs R^-7) r_i r_l - 15 R^-7 ((Sr)_i r_l + (Sr)_l r_i)
//          + d_il (1.5 trs R^-5 - 7.5 rSr R^-7) + 3 R^-5 S_il
//   H_ijk = M (15 r_i r_j r_k R^-7 - 3 (d_ij r_k + d_ik r_j + d_jk r_i) R^-5)
// (the monopole to second order in delta, the quadrupole to first).
// Storage: F[0..2] A, F[3..8] J (xx xy xz yy yz zz), F[9..18] H (xxx xxy xxz xyy
// xyz xzz yyy yyz yzz zzz).
constexpr int kFarTerms = 19;
#ifndef BH_FAR_GROUP_WEIGHT
#define BH_FAR_GROUP_WEIGHT 3.0f
#endif
constexpr float kFarGroupWeight = BH_FAR_GROUP_WEIGHT;

__device__ inline void far_accumulate(float rx, float ry, float rz, const WalkNode& w, float eps2,
                                      float* F)
{
    const float R2 = rx * rx + ry * ry + rz * rz + eps2;
    const float ri = __frsqrt_rn(R2);
    const float ri2 = ri * ri;
    const float ri3 = ri * ri2, ri5 = ri3 * ri2, ri7 = ri5 * ri2, ri9 = ri7 * ri2;
    const float M = w.mass, t = w.trs;
    const float third = 1.0f / 3.0f;
    const float sxx = (w.qxx + t) * third, syy = (w.qyy + t) * third;
    const float szz = (-w.qxx - w.qyy + t) * third;
    const float sxy = w.qxy * third, sxz = w.qxz * third, syz = w.qyz * third;
    const float qx = sxx * rx + sxy * ry + sxz * rz;
    const float qy = sxy * rx + syy * ry + syz * rz;
    const float qz = sxz * rx + syz * ry + szz * rz;
    const float rsr = rx * qx + ry * qy + rz * qz;
    // A
    const float am = M * ri3 + 7.5f * rsr * ri7 - 1.5f * t * ri5;
    F[0] += am * rx - 3.0f * ri5 * qx;
    F[1] += am * ry - 3.0f * ri5 * qy;
    F[2] += am * rz - 3.0f * ri5 * qz;
    // J
    const float jr = 3.0f * M * ri5 + 52.5f * rsr * ri9 - 7.5f * t * ri7;
    const float jd = -M * ri3 + 1.5f * t * ri5 - 7.5f * rsr * ri7;
    const float js = 3.0f * ri5, jq = -15.0f * ri7;
    F[3] += jr * rx * rx + jq * 2.0f * qx * rx + jd + js * sxx;
    F[4] += jr * rx * ry + jq * (qx * ry + qy * rx) + js * sxy;
    F[5] += jr * rx * rz + jq * (qx * rz + qz * rx) + js * sxz;
    F[6] += jr * ry * ry + jq * 2.0f * qy * ry + jd + js * syy;
    F[7] += jr * ry * rz + jq * (qy * rz + qz * ry) + js * syz;
    F[8] += jr * rz * rz + jq * 2.0f * qz * rz + jd + js * szz;
    // H
    const float c7 = 15.0f * M * ri7, c5 = 3.0f * M * ri5;
    F[9] += (c7 * rx * rx - 3.0f * c5) * rx;
    F[10] += (c7 * rx * rx - c5) * ry;
    F[11] += (c7 * rx * rx - c5) * rz;
    F[12] += (c7 * ry * ry - c5) * rx;
    F[13] += c7 * rx * ry * rz;
    F[14] += (c7 * rz * rz - c5) * rx;
    F[15] += (c7 * ry * ry - 3.0f * c5) * ry;
    F[16] += (c7 * ry * ry - c5) * rz;
    F[17] += (c7 * rz * rz - c5) * ry;
    F[18] += (c7 * rz * rz - 3.0f * c5) * rz;
}

__device__ inline void far_evaluate(const float* F, float dx, float dy, float dz, float& ax,
                                    float& ay, float& az)
{
    const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    const float xy = dx * dy, xz = dx * dz, yz = dy * dz;
    ax += F[0] + F[3] * dx + F[4] * dy + F[5] * dz +
          0.5f * (F[9] * xx + F[12] * yy + F[14] * zz) + F[10] * xy + F[11] * xz + F[13] * yz;
    ay += F[1] + F[4] * dx + F[6] * dy + F[7] * dz +
          0.5f * (F[10] * xx + F[15] * yy + F[17] * zz) + F[12] * xy + F[13] * xz + F[16] * yz;
    az += F[2] + F[5] * dx + F[7] * dy + F[8] * dz +
          0.5f * (F[11] * xx + F[16] * yy + F[18] * zz) + F[13] * xy + F[14] * xz + F[17] * yz;
}

// Far nodes of a group are queued in LDS (node indices) and folded into the
// group's expansion in batches of up to kFarCap, with every lane taking whole
// nodes; the lanes' partial sums are reduced once per batch into fsum. The
// 19 accumulators are then live only inside the batch, not across the list
// evaluation loops.
constexpr int kFarCap = 256;

__device__ inline void far_flush(const uint32_t* __restrict__ fl, int nfl,
                                 const WalkNode* __restrict__ nodes, int rcx, int rcy, int rcz,
                                 float ux, float uy, float uz, float gcx, float gcy, float gcz,
                                 float eps2, float* __restrict__ fsum)
{
    float F[kFarTerms];
#pragma unroll
    for (int k = 0; k < kFarTerms; ++k)
        F[k] = 0.0f;
#pragma unroll 1
    for (int i = static_cast<int>(threadIdx.x); i < nfl; i += 64) {
        const WalkNode w = nodes[fl[i]];
        const float rx = static_cast<float>(w.cx - rcx) * ux + w.comx - gcx;
        const float ry = static_cast<float>(w.cy - rcy) * uy + w.comy - gcy;
        const float rz = static_cast<float>(w.cz - rcz) * uz + w.comz - gcz;
        far_accumulate(rx, ry, rz, w, eps2, F);
    }
#pragma unroll
    for (int k = 0; k < kFarTerms; ++k) {
        float v = F[k];
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        F[k] = v;
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kFarTerms; ++k)
            fsum[k] += F[k];
    }
}

template <int ORDER, int CRIT, bool TPM, bool REC, bool FAR>
#ifndef BH_WALK_MIN_WAVES
#define BH_WALK_MIN_WAVES 5
#endif
// Walks without the far field are held to 96 VGPRs (5 waves per SIMD): the
// traversal is latency-bound and needs the resident waves.
__global__ __launch_bounds__(64, FAR ? 1 : BH_WALK_MIN_WAVES) void walk_kernel(
    WalkParams P, const WalkNode* __restrict__ nodes, const float4* __restrict__ spart,
    const uint32_t* __restrict__ sleaf, const RootGeom* __restrict__ geom,
    const uint4* __restrict__ groups, const int* __restrict__ info,
    int node_cap, unsigned int* __restrict__ work, int* __restrict__ stack_all, int stack_cap,
    float* __restrict__ accx, float* __restrict__ accy, float* __restrict__ accz,
    int* __restrict__ err, unsigned long long* __restrict__ stats, uint32_t* __restrict__ rec_lists,
    uint32_t* __restrict__ rec_n, int rec_cap, int rec_groups, float rec_band2,
    uint32_t rec_min_count)
{
    __shared__ float4 cA[kCellCap];
    __shared__ float4 cB[kCellCap];
    __shared__ float2 cC[kCellCap];
    __shared__ float4 pl[kPartCap];
    __shared__ int lfFirst[64];
    __shared__ uint32_t lfPre[64];
    __shared__ float lfSx[64], lfSy[64], lfSz[64];
    __shared__ uint32_t fl[FAR ? kFarCap : 1];
    __shared__ float fsum[FAR ? kFarTerms : 1];

    const int lane = threadIdx.x;
    int* stk = stack_all + static_cast<size_t>(blockIdx.x) * static_cast<size_t>(stack_cap);
    const float ux = geom->unit_f[0], uy = geom->unit_f[1], uz = geom->unit_f[2];
    const float wfx = geom->width_f[0], wfy = geom->width_f[1], wfz = geom->width_f[2];
    const float wfmax = fmaxf(fmaxf(wfx, wfy), wfz);
    const float rcut = TPM && P.rcut2 > 0.0f ? sqrtf(P.rcut2) : 0.0f;
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    unsigned long long npp = 0, npc = 0;
    // this wave's interaction totals, added to `stats` once when it exits
    // (one set of atomics per wave instead of per group: the counters are
    // single addresses every wave would otherwise contend for)
    unsigned long long tot_pp = 0, tot_pc = 0, tot_mono = 0, tot_far = 0, tot_eval = 0;
    const int ngroups = info[kInfoGroups];

    // A wave draws from its XCD's queue and, once that is drained, from the
    // others in turn; every queue position below ngroups is handed out exactly
    // once. (A ticket is drawn only when the wave is free: a wave that held
    // its next group while walking the current one would keep that group from
    // an idle wave at the end of the walk.)
    const int nq = static_cast<long long>(ngroups) >=
                           static_cast<long long>(kQueueMinGroupsPerWave) * gridDim.x
                       ? kQueues
                       : 1;
    int queue = static_cast<int>(blockIdx.x % static_cast<unsigned>(nq));
    int drained = 0;
    for (;;) {
        int t = ngroups;
        if (lane == 0) {
            for (;;) {
                t = ticket_position(atomicAdd(&work[queue * kTicketStride], 1u), queue, nq);
                if (t < ngroups || drained == nq - 1)
                    break;
                ++drained;
                queue = (queue + 1) % nq;
            }
            t = t < ngroups ? t : ngroups;
        }
        // the whole wave is active here, so the first lane is lane 0
        const int ticket = __builtin_amdgcn_readfirstlane(t);
        if (ticket >= ngroups)
            break;
        // Groups are stored in the breadth-first order of their nodes, shallow
        // (sparse, cheap) levels first; taking them deepest first puts the
        // expensive dense groups at the start instead of in the tail.
        const int g = ngroups - 1 - ticket;

        const uint4 gr = groups[g];
        int nact = static_cast<int>(gr.y);
        // Targets of this walk: the group's particles, or under block time
        // steps those due a force now. The box, the relative tolerance and
        // the stored forces cover these only (the corrector reads no other
        // particle's force, and each particle is due again before its next
        // kick), so a group with a few due particles walks with their box.
        // The due particles are compacted to the first lanes (their offsets
        // pass through lfFirst, unused until the traversal), so that a group
        // with few of them splits its lists like a small group.
        if (P.alevel) {
            bool due = false;
            if (lane < nact) {
                const uint32_t lv = min(P.alevel[gr.x + static_cast<uint32_t>(lane)],
                                        static_cast<uint32_t>(P.alevels));
                due = (P.atick & ((1u << (static_cast<uint32_t>(P.alevels) - lv)) - 1u)) == 0u;
            }
            const unsigned long long dm = __ballot(due);
            if (dm == 0ull)
                continue;
            if (due)
                lfFirst[__popcll(dm & lt)] = lane;
            nact = static_cast<int>(__popcll(dm));
            __syncthreads();
        }
        // groups of <= 32 (<= 16) targets: 2 (4) lanes per target, each
        // evaluating every 2nd (4th) list entry
        const int dj = nact <= 16 ? 4 : (nact <= 32 ? 2 : 1);
        const int tw = 64 / dj;  // lanes per list share
        const int tl = lane & (tw - 1);
        const int j0 = lane / tw;
        const bool active = tl < nact;
        const int s = static_cast<int>(gr.x) + (P.alevel ? (active ? lfFirst[tl] : 0) : tl);
        const bool target = active;
        const int gnode = static_cast<int>(gr.z);
        const int rcx = nodes[gnode].cx, rcy = nodes[gnode].cy, rcz = nodes[gnode].cz;

        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        if (active) {
            const float4 p = spart[s];
            const int lf = static_cast<int>(sleaf[s]);
            tx = rel_pos(nodes[lf].cx - rcx, ux, p.x);
            ty = rel_pos(nodes[lf].cy - rcy, uy, p.y);
            tz = rel_pos(nodes[lf].cz - rcz, uz, p.z);
        }
        const float gx0 = wmin(target ? tx : 3.0e38f), gx1 = wmax(target ? tx : -3.0e38f);
        const float gy0 = wmin(target ? ty : 3.0e38f), gy1 = wmax(target ? ty : -3.0e38f);
        const float gz0 = wmin(target ? tz : 3.0e38f), gz1 = wmax(target ? tz : -3.0e38f);
        const float gcx = 0.5f * (gx0 + gx1), gcy = 0.5f * (gy0 + gy1), gcz = 0.5f * (gz0 + gz1);
        const float ghx = 0.5f * (gx1 - gx0), ghy = 0.5f * (gy1 - gy0), ghz = 0.5f * (gz1 - gz0);
        // relative criterion: the group's smallest |a| of the last evaluation,
        // floored at 1 % of its largest (a target whose force nearly cancels
        // has no meaningful relative error and would open everything)
        float rtol = 0.0f, rtol_mono = 0.0f;
        if (CRIT == 3) {
            const float am = target ? P.amag[s] : 0.0f;
            const float amin = wmin(target ? am : 3.0e38f), amax = wmax(am);
            rtol = P.ralpha * fmaxf(amin, 0.01f * amax);
            rtol_mono = P.rmono * rtol;
        }
        unsigned long long nmono = 0;

        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        const float grad = FAR ? sqrtf(ghx * ghx + ghy * ghy + ghz * ghz) : 0.0f;
        unsigned long long nfar = 0;
        int nfl = 0;
        if (FAR && lane < kFarTerms)
            fsum[lane] = 0.0f;
        int ncell = 0, npart = 0;
        int nrec = 0;
        int sp = 1;
        if (lane == 0)
            stk[0] = 0;  // root
        __syncthreads();

        // Every iteration pops at least one node and a node is pushed at most
        // once per group in a well-formed tree, so node_cap iterations bound
        // the loop; the guard also stops a malformed tree from spinning.
        int iters = 0;
        while (sp > 0) {
            if (++iters > node_cap) {
                if (lane == 0)
                    atomicOr(err, kErrStackOverflow);
                break;
            }
            const int take = sp < 64 ? sp : 64;
            sp -= take;
            int node = lane < take ? stk[sp + lane] : -1;
            if (node >= node_cap) {
                atomicOr(err, kErrStackOverflow);
                node = -1;
            }
            __syncthreads();

            bool accept = false, direct = false, open = false, mono = false;
            WalkNode w = {};
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            float px = 0.0f, py = 0.0f, pz = 0.0f;
            float size = 0.0f;
            bool near = false;
            bool wide = false;
            if (node >= 0) {
                w = nodes[node];
                if (w.mass > 0.0f) {
                    sx = static_cast<float>(w.cx - rcx) * ux;
                    sy = static_cast<float>(w.cy - rcy) * uy;
                    sz = static_cast<float>(w.cz - rcz) * uz;
                    px = sx + w.comx;
                    py = sy + w.comy;
                    pz = sz + w.comz;
                    if (TPM && P.box > 0.0f) {
                        if (4.0f * w.bmax > P.box) {
                            wide = true;  // spans several images: always opened
                        } else {
                            // the node's image nearest to the group
                            const float ib = 1.0f / P.box;
                            const float hx = -P.box * rintf((px - gcx) * ib);
                            const float hy = -P.box * rintf((py - gcy) * ib);
                            const float hz = -P.box * rintf((pz - gcz) * ib);
                            sx += hx;
                            sy += hy;
                            sz += hz;
                            px += hx;
                            py += hy;
                            pz += hz;
                        }
                    }
                    const float dx = fmaxf(fabsf(px - gcx) - ghx, 0.0f);
                    const float dy = fmaxf(fabsf(py - gcy) - ghy, 0.0f);
                    const float dz = fmaxf(fabsf(pz - gcz) - ghz, 0.0f);
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    bool ok = d2 > w.bmax * w.bmax;
                    const int lvl = static_cast<int>(meta_level(w.meta));
                    size = cell_size(wfmax, lvl);
                    if (CRIT == 1) {
                        const float hs = ldexpf(0.5f, kMaxLevel - lvl);
                        const float bx = fmaxf(fabsf(sx - gcx) - ghx - hs * wfx, 0.0f);
                        const float by = fmaxf(fabsf(sy - gcy) - ghy - hs * wfy, 0.0f);
                        const float bz = fmaxf(fabsf(sz - gcz) - ghz - hs * wfz, 0.0f);
                        const float b2 = bx * bx + by * by + bz * bz;
                        ok = ok && b2 > 0.0f && size * size < P.theta2 * b2;
                    } else if (CRIT == 2) {
                        ok = ok && w.bmax * w.bmax < P.theta2 * (d2 + P.eps2);
                    } else if (CRIT == 3) {
                        const float r2 = d2 + P.eps2;
                        const float b2 = w.bmax * w.bmax;
                        const float err = ORDER >= 2 ? w.mass * b2 * w.bmax : w.mass * b2;
                        // (v_sqrt_f32: the acceptance needs no correctly rounded root)
                        const float lim = ORDER >= 2 ? rtol * r2 * r2 * __builtin_amdgcn_sqrtf(r2)
                                                     : rtol * r2 * r2;
                        ok = ok && size * size < P.theta2 * r2 && err <= lim;
                        // quadrupole term within tolerance too: a point mass suffices
                        mono = ORDER >= 2 && ok && w.mass * b2 <= rtol_mono * r2 * r2;
                    } else {
                        ok = ok && size * size < P.theta2 * (d2 + P.eps2);
                    }
                    if (TPM && P.rcut2 > 0.0f && !wide) {
                        // Prune nodes whose bounding sphere lies beyond the cutoff for the
                        // whole group.
                        const float lim = rcut + w.bmax;
                        if (d2 > lim * lim) {
                            ok = false;
                            w.mass = 0.0f;
                        }
                    }
                    if (wide)
                        ok = false;
                    if (REC && ok) {
                        // near the group: small nodes are opened, large ones
                        // accepted and recorded for the learned correction
                        // (a group beyond the lists' capacity opens them all)
                        near = size * size >= rec_band2 * d2;
                        if (near && (meta_count(w.meta) < rec_min_count || g >= rec_groups)) {
                            ok = false;
                            near = false;
                        }
                    }
                    if (w.mass > 0.0f) {
                        const uint32_t nch = meta_nchild(w.meta);
                        accept = ok;
                        direct = !ok && nch == 0;
                        open = !ok && nch != 0;
                    }
                }
            }

            if (REC) {
                near = near && accept;
                const unsigned long long rm = __ballot(near);
                if (near) {
                    const int slot = nrec + static_cast<int>(__popcll(rm & lt));
                    if (slot < rec_cap) {
                        rec_lists[static_cast<size_t>(g) * rec_cap + slot] =
                            static_cast<uint32_t>(node);
                    } else {
                        // the group's list is full: open the node instead
                        // (its error is then the plain walk's, never an
                        // uncorrected near node's)
                        const uint32_t nch = meta_nchild(w.meta);
                        accept = false;
                        direct = nch == 0;
                        open = nch != 0;
                    }
                }
                nrec += static_cast<int>(__popcll(rm));
            }

            if (FAR) {
                // far from the whole group: queued for the group's local
                // expansion instead of the per-target list. The group radius
                // weighs kFarGroupWeight times the node's: the expansion is
                // second order in the offset from the group centre, with a
                // larger error constant than the node's own second-order
                // multipole.
                bool far = false;
                if (accept) {
                    const float rx = px - gcx, ry = py - gcy, rz = pz - gcz;
                    const float lim = w.bmax + kFarGroupWeight * grad;
                    far = lim * lim < P.far2 * (rx * rx + ry * ry + rz * rz);
                    accept = accept && !far;
                }
                const unsigned long long fm = __ballot(far);
                if (far)
                    fl[nfl + static_cast<int>(__popcll(fm & lt))] = static_cast<uint32_t>(node);
                nfl += static_cast<int>(__popcll(fm));
                nfar += __popcll(fm);
            }

            // 1) accepted nodes -> LDS cell list (point masses -> particle list)
            if (CRIT == 3) {
                mono = mono && accept;
                accept = accept && !mono;
                const unsigned long long mm = __ballot(mono);
                const int nm = static_cast<int>(__popcll(mm));
                if (nm > 0) {
                    if (npart + nm >= kPartCap) {  // uniform; keeps npart < kPartCap
                        eval_parts<TPM>(pl, npart, j0, dj, tx, ty, tz, P, ax, ay, az);
                        npp += static_cast<unsigned long long>(npart);
                        npart = 0;
                        __syncthreads();
                    }
                    if (mono)
                        pl[npart + static_cast<int>(__popcll(mm & lt))] =
                            make_float4(px, py, pz, w.mass);
                    npart += nm;
                    nmono += static_cast<unsigned long long>(nm);
                }
            }
            const unsigned long long am = __ballot(accept);
            if (accept) {
                const int slot = ncell + static_cast<int>(__popcll(am & lt));
                cA[slot] = make_float4(px, py, pz, w.mass);
                cB[slot] = make_float4(w.qxx, w.qxy, w.qxz, w.qyy);
                // the open walk's expansion needs eps^2 trs, the TreePM one trs
                cC[slot] = make_float2(w.qyz, TPM ? w.trs : P.eps2 * w.trs);
            }
            ncell += static_cast<int>(__popcll(am));

            // 2) children of opened nodes -> traversal stack
            const uint32_t nc = open ? meta_nchild(w.meta) : 0u;
            const uint32_t cinc = winclusive(nc);
            const int ctot = static_cast<int>(wlast(cinc));
            if (sp + ctot > stack_cap) {
                if (lane == 0)
                    atomicOr(err, kErrStackOverflow);
            } else {
                if (open) {
                    const int base = sp + static_cast<int>(cinc - nc);
                    for (uint32_t j = 0; j < nc; ++j)
                        stk[base + static_cast<int>(j)] = w.first + static_cast<int>(j);
                }
                sp += ctot;
            }

            // 3) opened leaves -> descriptors, then their particles -> LDS
            const unsigned long long dm = __ballot(direct);
            const int nd = static_cast<int>(__popcll(dm));
            const uint32_t lc = direct ? meta_count(w.meta) : 0u;
            const uint32_t linc = winclusive(lc);
            const uint32_t ltot = wlast(linc);
            if (direct) {
                const int slot = static_cast<int>(__popcll(dm & lt));
                lfFirst[slot] = w.first;
                lfPre[slot] = linc - lc;
                lfSx[slot] = sx;
                lfSy[slot] = sy;
                lfSz[slot] = sz;
            }
            __syncthreads();

            if (FAR && nfl > kFarCap - 64) {
                far_flush(fl, nfl, nodes, rcx, rcy, rcz, ux, uy, uz, gcx, gcy, gcz, P.eps2, fsum);
                nfl = 0;
            }

            if (ncell > kCellCap - 64) {
                eval_cells<ORDER, TPM>(cA, cB, cC, ncell, j0, dj, tx, ty, tz, P, ax, ay, az);
                npc += static_cast<unsigned long long>(ncell);
                ncell = 0;
                __syncthreads();
            }

            uint32_t done = 0;
            while (done < ltot) {
                const uint32_t room = static_cast<uint32_t>(kPartCap - npart);
                const uint32_t chunk = ltot - done < room ? ltot - done : room;
                for (uint32_t f = static_cast<uint32_t>(lane); f < chunk; f += 64u) {
                    const uint32_t gi = done + f;
                    int lo = 0, hi = nd - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (lfPre[mid] <= gi)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    const float4 q = spart[lfFirst[lo] + static_cast<int>(gi - lfPre[lo])];
                    pl[npart + static_cast<int>(f)] =
                        make_float4(q.x + lfSx[lo], q.y + lfSy[lo], q.z + lfSz[lo], q.w);
                }
                __syncthreads();
                npart += static_cast<int>(chunk);
                done += chunk;
                if (npart == kPartCap) {
                    eval_parts<TPM>(pl, npart, j0, dj, tx, ty, tz, P, ax, ay, az);
                    npp += static_cast<unsigned long long>(npart);
                    npart = 0;
                    __syncthreads();
                }
            }
        }
        eval_cells<ORDER, TPM>(cA, cB, cC, ncell, j0, dj, tx, ty, tz, P, ax, ay, az);
        eval_parts<TPM>(pl, npart, j0, dj, tx, ty, tz, P, ax, ay, az);
        npc += static_cast<unsigned long long>(ncell);
        npp += static_cast<unsigned long long>(npart);
        for (int o = 32; o >= tw; o >>= 1) {  // the list shares of one target
            ax += __shfl_xor(ax, o, 64);
            ay += __shfl_xor(ay, o, 64);
            az += __shfl_xor(az, o, 64);
        }
        if (FAR) {
            // the rest of the queue, then the group's expansion at each target
            far_flush(fl, nfl, nodes, rcx, rcy, rcz, ux, uy, uz, gcx, gcy, gcz, P.eps2, fsum);
            nfl = 0;
            __syncthreads();
            float F[kFarTerms];
#pragma unroll
            for (int k = 0; k < kFarTerms; ++k)
                F[k] = fsum[k];
            far_evaluate(F, tx - gcx, ty - gcy, tz - gcz, ax, ay, az);
        }
        // per-target interaction counts for the roofline (only active lanes are real targets)
        npp = npp * static_cast<unsigned long long>(nact);
        npc = npc * static_cast<unsigned long long>(nact);
        tot_pp += npp;
        tot_pc += npc;
        if (CRIT == 3)
            tot_mono += nmono * static_cast<unsigned long long>(nact);
        if (FAR) {
            tot_far += nfar;                                      // node-group
            tot_eval += static_cast<unsigned long long>(nact);   // evaluations
        }
        npp = 0;
        npc = 0;
        if (REC && lane == 0 && g < rec_groups)
            rec_n[g] = static_cast<uint32_t>(nrec < rec_cap ? nrec : rec_cap);
        __syncthreads();

        if (target && j0 == 0) {
            // OctHostMath.cpp:50-57
            const float mag = sqrtf(ax * ax + ay * ay + az * az);
            if (P.max_acc > 0.0f && mag > P.max_acc && mag > 1.0e-6f) {
                const float k = P.max_acc / mag;
                ax *= k;
                ay *= k;
                az *= k;
            }
            accx[s] = ax;
            accy[s] = ay;
            accz[s] = az;
            if (P.amag)
                P.amag[s] = sqrtf(ax * ax + ay * ay + az * az);
        }
    }
    if (stats && lane == 0) {
        atomicAdd(&stats[0], tot_pp);
        atomicAdd(&stats[1], tot_pc);
        if (CRIT == 3 && tot_mono)
            atomicAdd(&stats[5], tot_mono);
        if (FAR) {
            atomicAdd(&stats[3], tot_far);
            atomicAdd(&stats[4], tot_eval);
        }
    }
}

template <int ORDER, int CRIT, bool TPM, bool REC = false, bool FAR = false>
void launch_walk_t(Device& d, const WalkParams& p, unsigned grid, unsigned long long* stats,
                   const LffState* L = nullptr, float band = 0.0f, int min_count = 0)
{
    walk_kernel<ORDER, CRIT, TPM, REC, FAR><<<grid, 64, 0, d.stream>>>(
        p, d.wnode, d.spart, d.sleaf, d.geom, d.groups, d.level_info,
        static_cast<int>(d.node_cap), d.work_counter,
        d.stack, static_cast<int>(d.stack_cap), d.acc[0], d.acc[1], d.acc[2], d.err, stats,
        L ? L->lists : nullptr, L ? L->list_n : nullptr, L ? L->list_cap : 0,
        L ? static_cast<int>(L->list_groups) : 0, band * band,
        static_cast<uint32_t>(min_count > 0 ? min_count : 0));
}

}  // namespace

void launch_walk_record(Device& d, const WalkParams& p, unsigned long long* stats,
                        const LffState& L, float band, int min_count)
{
    if (d.n == 0)
        return;
    launch_fill_u32(d.work_counter, 0u, kTicketWords, d.stream);
    const unsigned grid = static_cast<unsigned>(d.stack_waves);
    if (p.criterion == 1)
        launch_walk_t<2, 1, false, true>(d, p, grid, stats, &L, band, min_count);
    else if (p.criterion == 2)
        launch_walk_t<2, 2, false, true>(d, p, grid, stats, &L, band, min_count);
    else
        launch_walk_t<2, 0, false, true>(d, p, grid, stats, &L, band, min_count);
}

void launch_walk(Device& d, const WalkParams& p, unsigned long long* stats)
{
    if (d.n == 0)
        return;
    launch_fill_u32(d.work_counter, 0u, kTicketWords, d.stream);
    const unsigned grid = static_cast<unsigned>(d.stack_waves);
    const bool tpm = p.rs > 0.0f;
    if (p.far2 > 0.0f && !tpm && p.order >= 2) {
        // quadrupole walk with the group far-field expansion
        if (p.criterion == 1)
            launch_walk_t<2, 1, false, false, true>(d, p, grid, stats);
        else if (p.criterion == 2)
            launch_walk_t<2, 2, false, false, true>(d, p, grid, stats);
        else if (p.criterion == 3)
            launch_walk_t<2, 3, false, false, true>(d, p, grid, stats);
        else
            launch_walk_t<2, 0, false, false, true>(d, p, grid, stats);
        return;
    }
    if (p.criterion == 3) {
        if (p.order >= 2)
            tpm ? launch_walk_t<2, 3, true>(d, p, grid, stats)
                : launch_walk_t<2, 3, false>(d, p, grid, stats);
        else
            tpm ? launch_walk_t<1, 3, true>(d, p, grid, stats)
                : launch_walk_t<1, 3, false>(d, p, grid, stats);
    } else if (p.criterion == 2) {
        if (p.order >= 2)
            tpm ? launch_walk_t<2, 2, true>(d, p, grid, stats)
                : launch_walk_t<2, 2, false>(d, p, grid, stats);
        else
            tpm ? launch_walk_t<1, 2, true>(d, p, grid, stats)
                : launch_walk_t<1, 2, false>(d, p, grid, stats);
    } else if (p.order >= 2) {
        if (p.criterion == 1)
            tpm ? launch_walk_t<2, 1, true>(d, p, grid, stats)
                : launch_walk_t<2, 1, false>(d, p, grid, stats);
        else
            tpm ? launch_walk_t<2, 0, true>(d, p, grid, stats)
                : launch_walk_t<2, 0, false>(d, p, grid, stats);
    } else {
        if (p.criterion == 1)
            tpm ? launch_walk_t<1, 1, true>(d, p, grid, stats)
                : launch_walk_t<1, 1, false>(d, p, grid, stats);
        else
            tpm ? launch_walk_t<1, 0, true>(d, p, grid, stats)
                : launch_walk_t<1, 0, false>(d, p, grid, stats);
    }
}

}  // namespace bh
