// TEST INFRASTRUCTURE ONLY -- the oracle. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this code, and only as the checker.
//
// FP64 direct sum, force-error metrics and FP64 multipole moments.
//
//   FP64 pairwise forces   engine/physics/treepm/model/TpmCpu.cpp:78-116
//                          (here with FP64 positions as well: the truth)
//   force-error metrics    engine/physics/fmm/metrics/FmmMetrics.cpp:11-45
//                          (relative L2, Linf, P99) plus the median
//   energy / momenta       engine/physics/fmm/metrics/FmmMetrics.cpp:47-83
//   order-2 moments        engine/physics/fmm/build/FmmBuild.cpp:254-292
//                          (raw moments about a centre; shifted to COM here)
//
// Parity status: direct sum pinned by closed forms (two-body, symmetric
// shells, Newton's third law sum_i m_i a_i = 0); metrics by hand-computed
// vectors; moments by the two-point closed form and parallel-axis theorem.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>
#include <omp.h>

extern "C" {

// Accelerations on `nt` targets from all n sources, FP64, skipping j == target.
// r2 = |d|^2 + eps^2; pairs with r2 <= min_d2 are skipped. G = 1.
void dx_direct_fp64(int n, const double* pos3, const double* mass, int nt, const int* targets,
                    double eps, double min_d2, double* acc3, int nthreads)
{
    const int threads = nthreads > 0 ? nthreads : omp_get_max_threads();
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int t = 0; t < nt; ++t) {
        const int i = targets[t];
        const double xi = pos3[3 * i], yi = pos3[3 * i + 1], zi = pos3[3 * i + 2];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const double dx = pos3[3 * j] - xi, dy = pos3[3 * j + 1] - yi,
                         dz = pos3[3 * j + 2] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz + e2;
            if (r2 <= min_d2)
                continue;
            const double inv = 1.0 / std::sqrt(r2);
            const double s = mass[j] * inv * inv * inv;
            ax += dx * s;
            ay += dy * s;
            az += dz * s;
        }
        acc3[3 * t] = ax;
        acc3[3 * t + 1] = ay;
        acc3[3 * t + 2] = az;
    }
}

// out[0..4] = relative L2, relative Linf, P99, median, max relative error.
// Per-sample relative error |a - r| / max(|r|, 1e-12).
void dx_force_error(int nt, const double* approx3, const double* ref3, double* out)
{
    std::vector<double> rel(static_cast<size_t>(nt));
    double r2 = 0.0, e2 = 0.0, rmax = 0.0, emax = 0.0;
    for (int i = 0; i < nt; ++i) {
        const double dx = approx3[3 * i] - ref3[3 * i], dy = approx3[3 * i + 1] - ref3[3 * i + 1],
                     dz = approx3[3 * i + 2] - ref3[3 * i + 2];
        const double rn = std::sqrt(ref3[3 * i] * ref3[3 * i] + ref3[3 * i + 1] * ref3[3 * i + 1] +
                                    ref3[3 * i + 2] * ref3[3 * i + 2]);
        const double en = std::sqrt(dx * dx + dy * dy + dz * dz);
        r2 += rn * rn;
        e2 += en * en;
        rmax = std::max(rmax, rn);
        emax = std::max(emax, en);
        rel[static_cast<size_t>(i)] = en / std::max(rn, 1.0e-12);
    }
    std::sort(rel.begin(), rel.end());
    const size_t p99 = std::min(rel.size() - 1u,
                                static_cast<size_t>(std::ceil(rel.size() * 0.99)) - 1u);
    const size_t mid = rel.size() / 2u;
    out[0] = std::sqrt(e2 / std::max(r2, 1.0e-24));
    out[1] = emax / std::max(rmax, 1.0e-12);
    out[2] = rel[p99];
    out[3] = (rel.size() % 2u == 1u) ? rel[mid] : 0.5 * (rel[mid - 1u] + rel[mid]);
    out[4] = rel.back();
}

// FP64 direct sum of the TreePM short-range force: the Newtonian pair force
// times w(r) = erfc(r / 2rs) + r / (rs sqrt(pi)) exp(-r^2 / 4rs^2), with r the
// softened distance (engine/physics/octree/force/OctreeForce.cpp:49-57), and
// pairs beyond the cutoff rc (unsoftened distance) dropped
// (OctreeTraversal.cpp:172-176).
// box > 0: periodic cube [0, box)^3, each pair at its nearest image
// (separation wrapped into [-box/2, box/2] per axis). The reference has no
// periodic short-range walk; this is the definition the GPU walk is held to.
void dx_direct_tpm_fp64_box(int n, const double* pos3, const double* mass, int nt,
                            const int* targets, double eps, double rs, double rc, double box,
                            double* acc3, int nthreads);

void dx_direct_tpm_fp64(int n, const double* pos3, const double* mass, int nt, const int* targets,
                        double eps, double rs, double rc, double* acc3, int nthreads)
{
    dx_direct_tpm_fp64_box(n, pos3, mass, nt, targets, eps, rs, rc, 0.0, acc3, nthreads);
}

void dx_direct_tpm_fp64_box(int n, const double* pos3, const double* mass, int nt,
                            const int* targets, double eps, double rs, double rc, double box,
                            double* acc3, int nthreads)
{
    const int threads = nthreads > 0 ? nthreads : omp_get_max_threads();
    const double e2 = eps * eps;
    const double rc2 = rc * rc;
    const double isp = 0.5641895835477563;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int t = 0; t < nt; ++t) {
        const int i = targets[t];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i)
                continue;
            double dx = pos3[3 * j] - pos3[3 * i], dy = pos3[3 * j + 1] - pos3[3 * i + 1],
                   dz = pos3[3 * j + 2] - pos3[3 * i + 2];
            if (box > 0.0) {
                dx -= box * std::nearbyint(dx / box);
                dy -= box * std::nearbyint(dy / box);
                dz -= box * std::nearbyint(dz / box);
            }
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (rc > 0.0 && d2 > rc2)
                continue;
            const double r = std::sqrt(d2 + e2);
            const double a = 0.5 * r / rs;
            const double w = std::erfc(a) + r * isp / rs * std::exp(-a * a);
            const double s = mass[j] * w / (r * r * r);
            ax += dx * s;
            ay += dy * s;
            az += dz * s;
        }
        acc3[3 * t] = ax;
        acc3[3 * t + 1] = ay;
        acc3[3 * t + 2] = az;
    }
}

// Kinetic energy, softened pair potential (i<j), linear and angular momentum.
void dx_energy(int n, const double* pos3, const double* vel3, const double* mass, double eps,
               double min_d2, double* out8, int nthreads)
{
    const int threads = nthreads > 0 ? nthreads : omp_get_max_threads();
    double ke = 0.0, pe = 0.0, p[3] = {0, 0, 0}, l[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const double* x = pos3 + 3 * i;
        const double* v = vel3 + 3 * i;
        ke += 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int a = 0; a < 3; ++a)
            p[a] += mass[i] * v[a];
        l[0] += mass[i] * (x[1] * v[2] - x[2] * v[1]);
        l[1] += mass[i] * (x[2] * v[0] - x[0] * v[2]);
        l[2] += mass[i] * (x[0] * v[1] - x[1] * v[0]);
    }
    const double e2 = eps * eps;
#pragma omp parallel for reduction(+ : pe) schedule(dynamic, 16) num_threads(threads)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = i + 1; j < n; ++j) {
            const double dx = pos3[3 * j] - pos3[3 * i], dy = pos3[3 * j + 1] - pos3[3 * i + 1],
                         dz = pos3[3 * j + 2] - pos3[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz + e2;
            if (r2 > min_d2)
                s -= mass[j] / std::sqrt(r2);
        }
        pe += mass[i] * s;
    }
    out8[0] = ke;
    out8[1] = pe;
    for (int a = 0; a < 3; ++a) {
        out8[2 + a] = p[a];
        out8[5 + a] = l[a];
    }
}

// FP64 monopole + traceless quadrupole of a particle set about its COM.
// out: mass, com(3), Q(6) = xx, xy, xz, yy, yz, zz with
// Q_ab = sum m (3 y_a y_b - |y|^2 delta_ab), y = x - com.
void dx_moments(int n, const double* pos3, const double* mass, double* out10)
{
    double m = 0.0, c[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        m += mass[i];
        for (int a = 0; a < 3; ++a)
            c[a] += mass[i] * pos3[3 * i + a];
    }
    for (int a = 0; a < 3; ++a)
        c[a] = m > 0.0 ? c[a] / m : 0.0;
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const double y[3] = {pos3[3 * i] - c[0], pos3[3 * i + 1] - c[1], pos3[3 * i + 2] - c[2]};
        const double r2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        q[0] += mass[i] * (3 * y[0] * y[0] - r2);
        q[1] += mass[i] * (3 * y[0] * y[1]);
        q[2] += mass[i] * (3 * y[0] * y[2]);
        q[3] += mass[i] * (3 * y[1] * y[1] - r2);
        q[4] += mass[i] * (3 * y[1] * y[2]);
        q[5] += mass[i] * (3 * y[2] * y[2] - r2);
    }
    out10[0] = m;
    for (int a = 0; a < 3; ++a)
        out10[1 + a] = c[a];
    for (int k = 0; k < 6; ++k)
        out10[4 + k] = q[k];
}

// Acceleration at point x from a (mass, com, Q) expansion with Plummer-softened
// radius r^2 = |d|^2 + eps^2, d = com - x:
//   a = M d / r^3 - Q d / r^5 + 5/2 (d.Q d) d / r^7.
void dx_multipole_accel(const double* m10, const double* x3, double eps, int order, double* a3)
{
    const double d[3] = {m10[1] - x3[0], m10[2] - x3[1], m10[3] - x3[2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps;
    const double ri = 1.0 / std::sqrt(r2);
    const double ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
    double a[3] = {m10[0] * d[0] * ri3, m10[0] * d[1] * ri3, m10[0] * d[2] * ri3};
    if (order >= 2) {
        const double* q = m10 + 4;
        const double qd[3] = {q[0] * d[0] + q[1] * d[1] + q[2] * d[2],
                              q[1] * d[0] + q[3] * d[1] + q[4] * d[2],
                              q[2] * d[0] + q[4] * d[1] + q[5] * d[2]};
        const double dqd = d[0] * qd[0] + d[1] * qd[1] + d[2] * qd[2];
        for (int k = 0; k < 3; ++k)
            a[k] += -qd[k] * ri5 + 2.5 * dqd * d[k] * ri7;
    }
    for (int k = 0; k < 3; ++k)
        a3[k] = a[k];
}
}
