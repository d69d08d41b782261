// Deterministic initial conditions (include/bh_ic.h).
#include "ic.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

// SplitMix64 finaliser as a counter-based generator: the k-th draw of particle
// i in stream s is mix(seed, s, i, k).
inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t base;
    uint64_t k = 0;
    Stream(uint64_t seed, uint64_t stream, uint64_t index)
        : base(mix64(mix64(mix64(seed) ^ (stream * 0xd1b54a32d192ed03ull)) ^ index))
    {
    }
    // Uniform in the open interval (0, 1).
    double uniform()
    {
        const uint64_t r = mix64(base + 0x632be59bd9b4e019ull * (++k));
        return (static_cast<double>(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
};

inline void isotropic(Stream& s, double r, double* out)
{
    const double c = 2.0 * s.uniform() - 1.0;
    const double ph = 2.0 * kPi * s.uniform();
    const double sn = std::sqrt(std::max(0.0, 1.0 - c * c));
    out[0] = r * sn * std::cos(ph);
    out[1] = r * sn * std::sin(ph);
    out[2] = r * c;
}

// Subtract the mass-weighted mean position and velocity (serial sums: the
// result does not depend on threading).
void centre(uint64_t n, double* pos3, double* vel3, const float* mass)
{
    double m = 0.0, cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        m += mass[i];
        for (int a = 0; a < 3; ++a) {
            cx[a] += mass[i] * pos3[3 * i + a];
            cv[a] += mass[i] * vel3[3 * i + a];
        }
    }
    if (m <= 0.0)
        return;
    for (int a = 0; a < 3; ++a) {
        cx[a] /= m;
        cv[a] /= m;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
        for (int a = 0; a < 3; ++a) {
            pos3[3 * i + a] -= cx[a];
            vel3[3 * i + a] -= cv[a];
        }
}

// Hernquist isotropic distribution function (Hernquist 1990, eq. 17) in units
// G = M = a = 1, as a function of the binding energy eps = -E in (0, 1).
inline double hernquist_f(double eps)
{
    if (eps <= 0.0)
        return 0.0;
    if (eps >= 1.0)
        eps = 1.0 - 1e-15;
    const double q = std::sqrt(eps);
    const double q2 = eps;
    const double s = std::sqrt(1.0 - q2);
    const double num = 3.0 * std::asin(q) + q * s * (1.0 - 2.0 * q2) * (8.0 * q2 * q2 - 8.0 * q2 - 3.0);
    return num / std::pow(1.0 - q2, 2.5) / (8.0 * std::sqrt(2.0) * kPi * kPi * kPi);
}

}  // namespace

extern "C" {

void bh_ic_plummer(uint64_t n, uint64_t seed, double total_mass, double scale_radius,
                   double* pos3, double* vel3, float* mass)
{
    const double a = scale_radius;
    const double vscale = std::sqrt(total_mass / a);
    const float m = static_cast<float>(total_mass / static_cast<double>(n));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
        Stream s(seed, 0, static_cast<uint64_t>(i));
        // Radius from the cumulative mass X = r^3 / (1 + r^2)^(3/2).
        const double x = s.uniform();
        const double r = a / std::sqrt(std::pow(x, -2.0 / 3.0) - 1.0);
        isotropic(s, r, pos3 + 3 * i);
        // Speed: q = v / v_esc with density g(q) = q^2 (1 - q^2)^(7/2),
        // rejection against the bound 0.1 (max of g is about 0.092).
        double q = 0.0;
        for (;;) {
            const double u = s.uniform();
            const double y = 0.1 * s.uniform();
            if (y < u * u * std::pow(1.0 - u * u, 3.5)) {
                q = u;
                break;
            }
        }
        const double vesc = std::sqrt(2.0) * std::pow(1.0 + r * r / (a * a), -0.25) * vscale;
        isotropic(s, q * vesc, vel3 + 3 * i);
        mass[i] = m;
    }
    centre(n, pos3, vel3, mass);
}

void bh_ic_hernquist(uint64_t n, uint64_t seed, uint64_t stream, double total_mass,
                     double scale_radius, double r_max, double* pos3, double* vel3,
                     float* mass)
{
    const double a = scale_radius;
    const double vscale = std::sqrt(total_mass / a);
    const double xmax = std::pow(r_max / (r_max + a), 2.0);
    const float m = static_cast<float>(total_mass / static_cast<double>(n));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
        Stream s(seed, stream + 1, static_cast<uint64_t>(i));
        // M(<r) / M = r^2 / (r + a)^2  ->  r = a sqrt(X) / (1 - sqrt(X)).
        const double x = xmax * s.uniform();
        const double sx = std::sqrt(x);
        const double r = a * sx / (1.0 - sx);
        isotropic(s, r, pos3 + 3 * i);
        // Dimensionless potential psi = 1 / (1 + r/a); speed u = v / vscale in
        // [0, sqrt(2 psi)) with density u^2 f(psi - u^2 / 2).
        const double psi = 1.0 / (1.0 + r / a);
        const double umax = std::sqrt(2.0 * psi);
        double peak = 0.0;
        for (int k = 1; k <= 64; ++k) {
            const double u = umax * k / 65.0;
            peak = std::max(peak, u * u * hernquist_f(psi - 0.5 * u * u));
        }
        peak *= 1.5;
        double u = 0.0;
        for (int tries = 0; tries < 100000; ++tries) {
            const double c = umax * s.uniform();
            const double y = peak * s.uniform();
            const double g = c * c * hernquist_f(psi - 0.5 * c * c);
            if (g > peak)
                peak = g * 1.5;  // envelope too low here: raise it and keep sampling
            if (y < g) {
                u = c;
                break;
            }
        }
        isotropic(s, u * vscale, vel3 + 3 * i);
        mass[i] = m;
    }
    centre(n, pos3, vel3, mass);
}

void bh_ic_hernquist_merger(uint64_t n_per_halo, uint64_t seed, double separation,
                            double pericentre, double* pos3, double* vel3, float* mass)
{
    const uint64_t n = n_per_halo;
    bh_ic_hernquist(n, seed, 0, 1.0, 1.0, 200.0, pos3, vel3, mass);
    bh_ic_hernquist(n, seed, 1, 1.0, 1.0, 200.0, pos3 + 3 * n, vel3 + 3 * n, mass + n);
    // Relative orbit of two point masses (total mass 2) on a parabola:
    // |v| = sqrt(2 G M / d), angular momentum L = sqrt(2 G M r_p).
    const double mt = 2.0;
    const double d = separation;
    const double v = std::sqrt(2.0 * mt / d);
    const double vt = std::sqrt(2.0 * mt * pericentre) / d;
    const double vr = -std::sqrt(std::max(0.0, v * v - vt * vt));
    // Halo 1 at +x relative to halo 0; relative velocity (vr, vt, 0).
    const double off[2][3] = {{-0.5 * d, 0.0, 0.0}, {0.5 * d, 0.0, 0.0}};
    const double vel[2][3] = {{-0.5 * vr, -0.5 * vt, 0.0}, {0.5 * vr, 0.5 * vt, 0.0}};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(2 * n); ++i) {
        const int h = i < static_cast<int64_t>(n) ? 0 : 1;
        for (int a = 0; a < 3; ++a) {
            pos3[3 * i + a] += off[h][a];
            vel3[3 * i + a] += vel[h][a];
        }
    }
}

void bh_ic_uniform_cube(uint64_t n, uint64_t seed, double half_extent, double total_mass,
                        double* pos3, double* vel3, float* mass)
{
    const float m = static_cast<float>(total_mass / static_cast<double>(n));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
        Stream s(seed, 7, static_cast<uint64_t>(i));
        for (int a = 0; a < 3; ++a) {
            pos3[3 * i + a] = half_extent * (2.0 * s.uniform() - 1.0);
            vel3[3 * i + a] = 0.0;
        }
        mass[i] = m;
    }
}
}
