// Dyadic block time steps: CPU restatement of the reference scheduler
// (TEST INFRASTRUCTURE ONLY).
//
// Follows engine/physics/octree/adaptive/OctAdaptive.cpp:17-133 and the GPU
// scheduler engine/physics/core/cuda/integration/CudOctreeGpuAdaptive.inl:
// 80-117 with the kernels of engine/physics/core/cuda/integration/
// CudAdaptive.inl:31-57 (stable dt, level choice), 147-172 (drift), 181-228
// (activity, correction):
//
//   levels L = min(max_level, 12), slices S = 2^L, quantum q = dt / S;
//   stable dt_i = min(dt, eta sqrt(eps / |a_i|) [if |a_i| > 1e-6],
//                 eta eps / |v_i| [if |v_i| > 1e-6 and the velocity
//                 criterion is on]), eps = max(softening, min_softening);
//   level_i = the smallest l in [0, L] with dt / 2^l <= stable dt_i, else L;
//   per slice (tick t -> t + 1):
//     all particles:  x += v q + a q^2 / 2;  v = clamp(v + a q)
//     active ((t + 1) mod 2^(L - level_i) == 0): new force a' at the
//       predicted positions; h = (t + 1 - last_i) q;
//       x += (a' - a) h^2 / 2;  v = clamp(v + (a' - a) h);  a = a';
//       last_i = t + 1;  level_i re-chosen from (a', v).
//
// The schedule is (re)initialised on the first step: forces of all particles,
// levels, last ticks = the current tick. (The reference sets the last ticks
// to 0 there, which gives a too-long first correction interval when it
// re-initialises after the tick counter has advanced; at tick 0 the two are
// the same.) Forces of active particles are computed after every particle has
// been drifted (the reference GPU kernel corrects a particle while other
// threads may still read its position).
//
// State is FP64; forces here are the FP64 Plummer-softened direct sum, so the
// restatement isolates the scheduler from the tree approximation.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

void direct_forces(int n, const double* pos, const double* mass, double eps2, const int* idx,
                   int ni, double* out)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < ni; ++k) {
        const int i = idx[k];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = 0; j < n; ++j) {
            const double dx = pos[3 * j] - pos[3 * i];
            const double dy = pos[3 * j + 1] - pos[3 * i + 1];
            const double dz = pos[3 * j + 2] - pos[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            if (j == i || r2 <= 0.0)
                continue;
            const double w = mass[j] / (r2 * std::sqrt(r2));
            ax += w * dx;
            ay += w * dy;
            az += w * dz;
        }
        out[3 * i] = ax;
        out[3 * i + 1] = ay;
        out[3 * i + 2] = az;
    }
}

int choose_level(const double* a, const double* v, double eta, double soft, double dt, int L,
                 int velocity_criterion)
{
    const double am = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double vm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double adt = am > 1e-6 ? eta * std::sqrt(soft / am) : dt;
    const double vdt = velocity_criterion && vm > 1e-6 ? eta * soft / vm : dt;
    const double stable = std::fmin(dt, std::fmin(adt, vdt));
    for (int l = 0; l <= L; ++l)
        if (dt / static_cast<double>(1u << l) <= stable)
            return l;
    return L;
}

void clamp_speed(double* v, double max_speed)
{
    if (max_speed <= 0.0)
        return;
    const double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (m > max_speed)
        for (int k = 0; k < 3; ++k)
            v[k] *= max_speed / m;
}

}  // namespace

extern "C" {

// One or more global steps of length dt. acc3/level/last_tick/tick carry the
// schedule between calls; reset = 1 initialises it (forces of all particles).
void ad_steps(int n, double* pos3, double* vel3, const double* mass, double* acc3,
              int32_t* level, uint64_t* last_tick, uint64_t* tick, double dt, int nsteps,
              int max_level, double eta, double soft, int velocity_criterion, double max_speed,
              double eps, int reset)
{
    const int L = max_level < 12 ? max_level : 12;
    const uint64_t S = 1ull << L;
    const double q = dt / static_cast<double>(S);
    const double eps2 = eps * eps;
    std::vector<int> all(n), active;
    for (int i = 0; i < n; ++i)
        all[i] = i;
    std::vector<double> anew(3 * static_cast<size_t>(n), 0.0);
    if (reset) {
        direct_forces(n, pos3, mass, eps2, all.data(), n, acc3);
        for (int i = 0; i < n; ++i) {
            level[i] = choose_level(acc3 + 3 * i, vel3 + 3 * i, eta, soft, dt, L,
                                    velocity_criterion);
            last_tick[i] = *tick;
        }
    }
    for (int s = 0; s < nsteps; ++s) {
        for (uint64_t sl = 0; sl < S; ++sl) {
            for (int i = 0; i < n; ++i) {
                for (int k = 0; k < 3; ++k) {
                    const double v = vel3[3 * i + k], a = acc3[3 * i + k];
                    pos3[3 * i + k] += v * q + a * (0.5 * q * q);
                    vel3[3 * i + k] = v + a * q;
                }
                clamp_speed(vel3 + 3 * i, max_speed);
            }
            const uint64_t t = *tick + 1;
            active.clear();
            for (int i = 0; i < n; ++i) {
                const int lv = level[i] < L ? level[i] : L;
                const uint64_t cadence = 1ull << (L - lv);
                if (t % cadence == 0)
                    active.push_back(i);
            }
            direct_forces(n, pos3, mass, eps2, active.data(), static_cast<int>(active.size()),
                          anew.data());
            for (int i : active) {
                const double h = static_cast<double>(t - last_tick[i]) * q;
                for (int k = 0; k < 3; ++k) {
                    const double c = anew[3 * i + k] - acc3[3 * i + k];
                    pos3[3 * i + k] += c * (0.5 * h * h);
                    vel3[3 * i + k] += c * h;
                    acc3[3 * i + k] = anew[3 * i + k];
                }
                clamp_speed(vel3 + 3 * i, max_speed);
                last_tick[i] = t;
                level[i] = choose_level(acc3 + 3 * i, vel3 + 3 * i, eta, soft, dt, L,
                                        velocity_criterion);
            }
            *tick = t;
        }
    }
}

int ad_choose_level(const double* a3, const double* v3, double eta, double soft, double dt,
                    int max_level, int velocity_criterion)
{
    return choose_level(a3, v3, eta, soft, dt, max_level < 12 ? max_level : 12,
                        velocity_criterion);
}

}  // extern "C"
