// bh-headless: the headless `--run --solver octree_hip` path of octree_hip.
//
// Mirrors the reference's headless runner (apps/headless/src/main.cpp:91-188)
// and batch runner (engine/batch/runner/BatRunner.cpp:126-215):
//   --config <file> (default simulation.ini), --inspect | --validate | --run,
//   --solver <name>, --target-steps <k>, --export-path <file>, --help
// load or generate the initial state -> configure the solver -> substeps =
// min(ceil(dt / target_dt), max_substeps) -> update loop -> export.
// Exit codes as the reference: 2 bad arguments, 3 configuration missing,
// 4 validation failed, 5 run faulted.
//
// The solver is reached only through the C ABI (include/bh_hip.h).
#include "../../../include/bh_hip.h"
#include "config.h"
#include "ic.h"
#include "snapshot.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct Args {
    std::string config = "simulation.ini";
    std::string command;  // inspect | validate | run
    std::string solver;
    std::string export_path;
    long long target_steps = 1;
    bool help = false;
    bool bad = false;
};

void usage(std::ostream& o, const char* prog)
{
    o << "usage: " << prog
      << " [--config <file>] (--inspect | --validate | --run) [--solver octree_hip]\n"
         "       [--target-steps <k>] [--export-path <file>]\n";
}

Args parse_args(int argc, char** argv)
{
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        auto next = [&](std::string& dst) {
            if (i + 1 >= argc) {
                a.bad = true;
                return;
            }
            dst = argv[++i];
        };
        if (s == "--help" || s == "-h")
            a.help = true;
        else if (s == "--config")
            next(a.config);
        else if (s == "--inspect" || s == "--validate" || s == "--run")
            a.command = s.substr(2);
        else if (s == "--solver")
            next(a.solver);
        else if (s == "--export-path")
            next(a.export_path);
        else if (s == "--target-steps") {
            std::string v;
            next(v);
            try {
                a.target_steps = std::stoll(v);
            } catch (...) {
                a.bad = true;
            }
            if (a.target_steps < 0)
                a.bad = true;
        } else {
            a.bad = true;
        }
    }
    return a;
}

bool generate(const bhh::CaseConfig& c, bhh::State& s, std::string& err)
{
    const uint64_t n = std::max<uint32_t>(2u, c.particle_count);
    std::string preset = c.preset;
    std::transform(preset.begin(), preset.end(), preset.begin(), ::tolower);
    if (preset == "plummer" || preset == "plummer_sphere") {
        s.pos.resize(3 * n);
        s.vel.resize(3 * n);
        s.mass.resize(n);
        bh_ic_plummer(n, c.seed, 1.0, c.scale_radius, s.pos.data(), s.vel.data(), s.mass.data());
    } else if (preset == "hernquist_merger" || preset == "merger") {
        const uint64_t h = n / 2;
        s.pos.resize(6 * h);
        s.vel.resize(6 * h);
        s.mass.resize(2 * h);
        bh_ic_hernquist_merger(h, c.seed, c.separation, c.pericentre, s.pos.data(), s.vel.data(),
                               s.mass.data());
    } else if (preset == "hernquist") {
        s.pos.resize(3 * n);
        s.vel.resize(3 * n);
        s.mass.resize(n);
        bh_ic_hernquist(n, c.seed, 0, 1.0, c.scale_radius, 200.0 * c.scale_radius, s.pos.data(),
                        s.vel.data(), s.mass.data());
    } else if (preset == "random_cloud" || preset == "uniform" || preset == "cube") {
        s.pos.resize(3 * n);
        s.vel.resize(3 * n);
        s.mass.resize(n);
        bh_ic_uniform_cube(n, c.seed, c.scale_radius, 1.0, s.pos.data(), s.vel.data(),
                           s.mass.data());
    } else {
        err = "scene preset '" + c.preset +
              "' has no octree_hip generator (plummer, hernquist, hernquist_merger, random_cloud)";
        return false;
    }
    s.temp.assign(s.mass.size(), 0.0f);
    return true;
}

}  // namespace

int main(int argc, char** argv)
{
    const Args a = parse_args(argc, argv);
    if (a.help) {
        usage(std::cout, argv[0]);
        return 0;
    }
    if (a.bad) {
        usage(std::cerr, argv[0]);
        return 2;
    }
    if (!std::filesystem::is_regular_file(a.config)) {
        std::cerr << "[headless] configuration file not found: "
                  << std::filesystem::absolute(a.config).string() << "\n";
        return 3;
    }
    bhh::CaseConfig cfg;
    if (!bhh::load_config(a.config, cfg, std::cerr)) {
        std::cerr << "[headless] cannot read configuration: " << a.config << "\n";
        return 3;
    }
    if (!a.solver.empty())
        cfg.solver = a.solver;
    if (a.command.empty()) {
        std::cerr << "[headless] no execution command selected; choose --inspect, --validate, or "
                     "--run\n";
        usage(std::cerr, argv[0]);
        return 2;
    }
    const std::filesystem::path case_dir = std::filesystem::absolute(a.config).parent_path();
    auto resolve_path = [&](std::string& p) {
        if (!p.empty() && std::filesystem::path(p).is_relative())
            p = (case_dir / p).lexically_normal().string();
    };
    resolve_path(cfg.input_file);
    resolve_path(cfg.export_directory);
    std::string export_path = a.export_path;
    resolve_path(export_path);

    const std::string solver = bhh::normalize_solver(cfg.solver);
    std::cout << "[headless] config path=" << std::filesystem::absolute(a.config).string() << "\n";
    std::cout << "[headless] effective configuration begin\n";
    bhh::write_config(std::cout, cfg);
    std::cout << "[headless] effective configuration end\n";

    std::vector<std::string> problems;
    if (solver != "octree_hip")
        problems.push_back("solver '" + cfg.solver +
                           "' is not provided by this build (octree_hip only; octree_cpu is the "
                           "reference's own CPU solver)");
    if (!(cfg.dt > 0.0f))
        problems.push_back("dt must be positive");
    for (const auto& p : problems)
        std::cout << "[validate] error: " << p << "\n";
    if (a.command == "inspect") {
        // the values the solver would run with (theta clamp of
        // SrvFormatAndTheta.cpp:54-69, then resolveForceLawPolicy)
        bh_hip_params_t in;
        bh_hip_default_params(&in);
        in.theta = bhh::resolve_theta(cfg);
        in.softening = cfg.softening;
        in.min_softening = cfg.min_softening;
        in.min_distance2 = cfg.min_distance2;
        in.min_theta = cfg.min_theta;
        bh_hip_params_t r;
        bh_hip_resolve_params(&in, &r);
        std::cout << "[headless] resolved theta=" << r.theta << " softening=" << r.softening
                  << " min_theta=" << r.min_theta << " min_distance2=" << r.min_distance2 << "\n";
        return 0;
    }
    if (a.command == "validate")
        return problems.empty() ? 0 : 4;
    if (!problems.empty())
        return 4;

    const auto t_init = std::chrono::steady_clock::now();
    bhh::State st;
    std::string err;
    if (cfg.scene_style == "file" || !cfg.input_file.empty()) {
        if (!bhh::read_snapshot(cfg.input_file, cfg.input_format, st, err) || st.size() < 2) {
            std::cerr << "[batch] failed to parse input state: " << cfg.input_file << " " << err
                      << "\n";
            if (!generate(cfg, st, err)) {
                std::cerr << "[batch] " << err << "\n";
                return 4;
            }
        }
    } else if (!generate(cfg, st, err)) {
        std::cerr << "[batch] " << err << "\n";
        return 4;
    }
    const uint64_t n = st.size();

    bh_hip_params_t p;
    bh_hip_default_params(&p);
    p.theta = bhh::resolve_theta(cfg);
    p.softening = cfg.softening;
    p.min_softening = cfg.min_softening;
    p.min_distance2 = cfg.min_distance2;
    p.min_theta = cfg.min_theta;
    p.max_acceleration = cfg.max_acceleration;
    p.max_speed = cfg.sph_max_speed;
    p.criterion = cfg.criterion == "bounds"     ? BH_HIP_CRITERION_BOUNDS
                  : cfg.criterion == "extent"   ? BH_HIP_CRITERION_EXTENT
                  : cfg.criterion == "relative" ? BH_HIP_CRITERION_RELATIVE
                                                : BH_HIP_CRITERION_COM;
    p.leaf_capacity = static_cast<int32_t>(cfg.leaf_capacity);
    p.bucket_capacity = static_cast<int32_t>(cfg.bucket);
    p.multipole_order = cfg.multipole_order;
    const bool euler = cfg.integrator != "leapfrog" && cfg.integrator != "kdk";
    p.integrator = euler ? BH_HIP_INTEGRATOR_EULER : BH_HIP_INTEGRATOR_LEAPFROG;
    p.treepm_split_scale = cfg.treepm_split_scale;
    p.treepm_cutoff = cfg.treepm_cutoff;
    p.far_alpha = cfg.far_alpha;
    if (cfg.relative_alpha > 0.0f)
        p.relative_alpha = cfg.relative_alpha;
    p.relative_mono = cfg.relative_mono;

    char cerr_buf[512] = {0};
    bh_hip_handle_t* h = nullptr;
    if (bh_hip_create(&p, n, &h, cerr_buf, sizeof(cerr_buf)) != BH_HIP_OK) {
        std::cerr << "[octree_hip] create failed: " << cerr_buf << "\n";
        return 5;
    }
    if (bh_hip_upload(h, st.pos.data(), st.vel.data(), st.mass.data()) != BH_HIP_OK) {
        bh_hip_get_last_error(h, cerr_buf, sizeof(cerr_buf));
        std::cerr << "[octree_hip] upload failed: " << cerr_buf << "\n";
        bh_hip_destroy(h);
        return 5;
    }
    // BatRunner.cpp:171-182
    const float target_dt = cfg.substep_target_dt > 0.0f ? cfg.substep_target_dt
                                                         : bhh::auto_target_substep_dt(euler, n);
    const uint32_t substeps = std::min<uint32_t>(
        std::max<uint32_t>(1u, static_cast<uint32_t>(std::ceil(cfg.dt / target_dt))),
        std::max<uint32_t>(1u, cfg.max_substeps));
    const double sub_dt = static_cast<double>(cfg.dt / static_cast<float>(substeps));
    const double init_ms = std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - t_init).count();
    std::cout << "[headless] batch runner=direct targetSteps=" << a.target_steps
              << " dt=" << cfg.dt << " solver=octree_hip integrator="
              << (euler ? "euler" : "leapfrog") << " substeps=" << substeps << " n=" << n
              << " theta=" << p.theta << "\n";

    const auto t_run = std::chrono::steady_clock::now();
    bool faulted = false;
    long long steps = 0;
    for (long long k = 0; k < a.target_steps && !faulted; ++k) {
        if (bh_hip_step(h, sub_dt, static_cast<int32_t>(substeps)) != BH_HIP_OK) {
            bh_hip_get_last_error(h, cerr_buf, sizeof(cerr_buf));
            std::cerr << "[batch] particle system update failed: " << cerr_buf << "\n";
            faulted = true;
            break;
        }
        ++steps;
    }
    const double run_ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t_run).count();
    double export_ms = 0.0;
    std::string out_path;
    if (!faulted) {
        const auto t_exp = std::chrono::steady_clock::now();
        if (bh_hip_download(h, st.pos.data(), st.vel.data()) != BH_HIP_OK) {
            faulted = true;
        } else {
            out_path = export_path.empty()
                           ? (std::filesystem::path(cfg.export_directory) /
                              ("snapshot_" + std::to_string(steps) + "." +
                               (cfg.export_format == "xyz" ? "xyz" : "bin")))
                                 .string()
                           : export_path;
            const std::string fmt = cfg.export_format == "xyz" ? "xyz" : "bin";
            if (!bhh::write_snapshot(out_path, export_path.empty() ? fmt : "auto", st,
                                     "solver=octree_hip step=" + std::to_string(steps), err)) {
                std::cerr << "[batch] could not write final state: " << err << "\n";
                faulted = true;
            }
        }
        export_ms = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - t_exp).count();
    }
    bh_hip_destroy(h);
    std::printf("[headless] completed steps=%lld particles=%llu solver=octree_hip faulted=%d "
                "init_ms=%.1f integrate_ms=%.1f export_ms=%.1f particle_steps_per_s=%.4g "
                "export=%s\n",
                steps, static_cast<unsigned long long>(n), faulted ? 1 : 0, init_ms, run_ms,
                export_ms, run_ms > 0 ? 1e3 * static_cast<double>(n) * steps * substeps / run_ms : 0.0,
                out_path.c_str());
    return faulted ? 5 : 0;
}
