// Particle-mesh long-range force of the periodic TreePM split (rocFFT).
//
// The short-range walk (k_walk.hip, template TPM) weights the Newtonian pair
// force by w(r) = erfc(r / 2 r_s) + r / (r_s sqrt(pi)) exp(-r^2 / 4 r_s^2);
// this mesh supplies the complement:
//   * cloud-in-cell deposit onto an N^3 mesh, cell h = L / N, cell centres at
//     (i + 1/2) h, periodic (FP32 atomics; the particle position is split
//     into cell index and fraction in FP64);
//   * rho_k = R2C FFT of the density;
//   * per axis a: A_k = -i k_a phi_k with
//       phi_k = -4 pi G rho_k exp(-k^2 r_s^2) / k^2 / W(k)^2,
//       W(k) = prod sinc^2(k h / 2)  (deposit and interpolation windows),
//     then a C2R FFT (unnormalised: the 1 / N^3 is applied on interpolation);
//   * cloud-in-cell interpolation of the three fields, added to acc.
// Mesh layout: row-major, z fastest: (ix N + iy) N + iz; the half spectrum
// is (ix N + iy) (N/2 + 1) + kz. Wavenumber index i maps to i for i < N/2
// and i - N otherwise (kz = 0..N/2).
//
// The reference's PM mesh (engine/physics/treepm/cuda/fft/TpmFft.inl:9-141,
// field TpmField.inl:7-90) runs on cuFFT with r_s = 0 in periodic mode
// (TpmCosmology.inl:52); this is the split form of the same mesh.
#include "bh_internal.h"

namespace bh {
namespace {

__device__ inline void cic_split(double x, double cells_per_len, int N, int& i0, float& f)
{
    const double u = x * cells_per_len - 0.5;
    const double fl = floor(u);
    f = static_cast<float>(u - fl);
    int i = static_cast<int>(fl) % N;
    i0 = i < 0 ? i + N : i;
}

__global__ __launch_bounds__(256) void pm_deposit_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const float* __restrict__ m, size_t n, int N, double cells_per_len, float inv_cell_vol,
    float* __restrict__ rho)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int ix, iy, iz;
    float fx, fy, fz;
    cic_split(px[i], cells_per_len, N, ix, fx);
    cic_split(py[i], cells_per_len, N, iy, fy);
    cic_split(pz[i], cells_per_len, N, iz, fz);
    const float w = m[i] * inv_cell_vol;
    const int jx = ix + 1 == N ? 0 : ix + 1, jy = iy + 1 == N ? 0 : iy + 1,
              jz = iz + 1 == N ? 0 : iz + 1;
    const size_t NN = static_cast<size_t>(N);
    const int xs[2] = {ix, jx}, ys[2] = {iy, jy}, zs[2] = {iz, jz};
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c)
                atomicAdd(&rho[(static_cast<size_t>(xs[a]) * NN + ys[b]) * NN + zs[c]],
                          w * wx[a] * wy[b] * wz[c]);
}

__device__ inline float sincf_half(float kh)  // sin(kh/2) / (kh/2)
{
    const float x = 0.5f * kh;
    return fabsf(x) < 1e-4f ? 1.0f - x * x * (1.0f / 6.0f) : sinf(x) / x;
}

// out_k = -i k_axis phi_k for one axis (float2 = interleaved complex).
__global__ __launch_bounds__(256) void pm_green_kernel(const float2* __restrict__ rho_k,
                                                       float2* __restrict__ out_k, int N,
                                                       float L, float rs, float G, int axis)
{
    const size_t Nh = static_cast<size_t>(N / 2 + 1);
    const size_t total = static_cast<size_t>(N) * N * Nh;
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total)
        return;
    const int kz = static_cast<int>(t % Nh);
    const int iy = static_cast<int>((t / Nh) % N);
    const int ix = static_cast<int>(t / (Nh * N));
    const float two_pi_L = 6.283185307179586f / L;
    const float kx = two_pi_L * static_cast<float>(ix < N / 2 ? ix : ix - N);
    const float ky = two_pi_L * static_cast<float>(iy < N / 2 ? iy : iy - N);
    const float kzf = two_pi_L * static_cast<float>(kz);
    const float k2 = kx * kx + ky * ky + kzf * kzf;
    if (k2 == 0.0f) {
        out_k[t] = make_float2(0.0f, 0.0f);
        return;
    }
    const float h = L / static_cast<float>(N);
    const float s = sincf_half(kx * h) * sincf_half(ky * h) * sincf_half(kzf * h);
    const float W = s * s;
    const float g = -12.566370614359172f * G * __expf(-k2 * rs * rs) / (k2 * W * W);
    const float2 r = rho_k[t];
    const float ka = axis == 0 ? kx : axis == 1 ? ky : kzf;
    // -i ka (g r): (re, im) -> (ka g r.im, -ka g r.re)
    out_k[t] = make_float2(ka * g * r.y, -ka * g * r.x);
}

__global__ __launch_bounds__(256) void pm_interp_kernel(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    size_t n, int N, double cells_per_len, const float* __restrict__ fx_mesh,
    const float* __restrict__ fy_mesh, const float* __restrict__ fz_mesh, float scale,
    float max_acc, float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int ix, iy, iz;
    float fx, fy, fz;
    cic_split(px[i], cells_per_len, N, ix, fx);
    cic_split(py[i], cells_per_len, N, iy, fy);
    cic_split(pz[i], cells_per_len, N, iz, fz);
    const int jx = ix + 1 == N ? 0 : ix + 1, jy = iy + 1 == N ? 0 : iy + 1,
              jz = iz + 1 == N ? 0 : iz + 1;
    const size_t NN = static_cast<size_t>(N);
    const int xs[2] = {ix, jx}, ys[2] = {iy, jy}, zs[2] = {iz, jz};
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                const size_t cidx = (static_cast<size_t>(xs[a]) * NN + ys[b]) * NN + zs[c];
                const float w = wx[a] * wy[b] * wz[c];
                gx += w * fx_mesh[cidx];
                gy += w * fy_mesh[cidx];
                gz += w * fz_mesh[cidx];
            }
    float x = ax[i] + scale * gx, y = ay[i] + scale * gy, z = az[i] + scale * gz;
    const float mag = sqrtf(x * x + y * y + z * z);
    if (max_acc > 0.0f && mag > max_acc && mag > 1.0e-6f) {  // OctHostMath.cpp:50-57
        const float k = max_acc / mag;
        x *= k;
        y *= k;
        z *= k;
    }
    ax[i] = x;
    ay[i] = y;
    az[i] = z;
}

unsigned grid_for(size_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

bool pm_setup(PmState& pm, int N, hipStream_t stream)
{
    if (pm.N == N && pm.fwd)
        return true;
    pm_release(pm);
    if (rocfft_setup() != rocfft_status_success)
        return false;
    const size_t lengths[3] = {static_cast<size_t>(N), static_cast<size_t>(N),
                               static_cast<size_t>(N)};
    if (rocfft_plan_create(&pm.fwd, rocfft_placement_notinplace,
                           rocfft_transform_type_real_forward, rocfft_precision_single, 3, lengths,
                           1, nullptr) != rocfft_status_success ||
        rocfft_plan_create(&pm.inv, rocfft_placement_notinplace,
                           rocfft_transform_type_real_inverse, rocfft_precision_single, 3, lengths,
                           1, nullptr) != rocfft_status_success)
        return false;
    size_t w1 = 0, w2 = 0;
    if (rocfft_plan_get_work_buffer_size(pm.fwd, &w1) != rocfft_status_success ||
        rocfft_plan_get_work_buffer_size(pm.inv, &w2) != rocfft_status_success)
        return false;
    pm.work_bytes = std::max(w1, w2);
    const size_t cells = static_cast<size_t>(N) * N * N;
    const size_t half = static_cast<size_t>(N) * N * (N / 2 + 1);
    bool ok = hipMalloc(&pm.rho, cells * sizeof(float)) == hipSuccess &&
              hipMalloc(&pm.rho_k, half * 2 * sizeof(float)) == hipSuccess &&
              hipMalloc(&pm.tmp_k, half * 2 * sizeof(float)) == hipSuccess;
    for (int a = 0; a < 3 && ok; ++a)
        ok = hipMalloc(&pm.field[a], cells * sizeof(float)) == hipSuccess;
    if (ok && pm.work_bytes)
        ok = hipMalloc(&pm.work, pm.work_bytes) == hipSuccess;
    if (!ok)
        return false;
    if (rocfft_execution_info_create(&pm.info) != rocfft_status_success ||
        rocfft_execution_info_set_stream(pm.info, stream) != rocfft_status_success ||
        (pm.work_bytes &&
         rocfft_execution_info_set_work_buffer(pm.info, pm.work, pm.work_bytes) !=
             rocfft_status_success))
        return false;
    pm.N = N;
    return true;
}

void pm_release(PmState& pm)
{
    if (pm.fwd)
        rocfft_plan_destroy(pm.fwd);
    if (pm.inv)
        rocfft_plan_destroy(pm.inv);
    if (pm.info)
        rocfft_execution_info_destroy(pm.info);
    for (void* p : {static_cast<void*>(pm.rho), static_cast<void*>(pm.rho_k),
                    static_cast<void*>(pm.tmp_k), static_cast<void*>(pm.field[0]),
                    static_cast<void*>(pm.field[1]), static_cast<void*>(pm.field[2]), pm.work})
        if (p)
            (void)hipFree(p);
    pm = PmState{};
}

bool launch_pm_accel(Device& d, PmState& pm, double L, float rs, float max_acc)
{
    if (d.n == 0)
        return true;
    const int N = pm.N;
    const size_t cells = static_cast<size_t>(N) * N * N;
    const size_t half = static_cast<size_t>(N) * N * (N / 2 + 1);
    const double cpl = static_cast<double>(N) / L;
    const double h = L / N;
    launch_fill_u32(reinterpret_cast<uint32_t*>(pm.rho), 0u, cells, d.stream);
    pm_deposit_kernel<<<grid_for(d.n), 256, 0, d.stream>>>(
        d.pos[0], d.pos[1], d.pos[2], d.mass, d.n, N, cpl, static_cast<float>(1.0 / (h * h * h)),
        pm.rho);
    void* in[1] = {pm.rho};
    void* out[1] = {pm.rho_k};
    if (rocfft_execute(pm.fwd, in, out, pm.info) != rocfft_status_success)
        return false;
    for (int a = 0; a < 3; ++a) {
        pm_green_kernel<<<grid_for(half), 256, 0, d.stream>>>(
            reinterpret_cast<const float2*>(pm.rho_k), reinterpret_cast<float2*>(pm.tmp_k), N,
            static_cast<float>(L), rs, 1.0f, a);
        void* ik[1] = {pm.tmp_k};
        void* of[1] = {pm.field[a]};
        if (rocfft_execute(pm.inv, ik, of, pm.info) != rocfft_status_success)
            return false;
    }
    pm_interp_kernel<<<grid_for(d.n), 256, 0, d.stream>>>(
        d.pos[0], d.pos[1], d.pos[2], d.n, N, cpl, pm.field[0], pm.field[1], pm.field[2],
        static_cast<float>(1.0 / static_cast<double>(cells)), max_acc, d.acc[0], d.acc[1],
        d.acc[2]);
    return true;
}

}  // namespace bh
