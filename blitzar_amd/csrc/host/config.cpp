// Directive INI parsing and key mapping for the octree_hip headless runner.
#include "config.h"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <functional>
#include <map>
#include <ostream>
#include <sstream>

namespace bhh {
namespace {

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a])))
        ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1])))
        --b;
    return s.substr(a, b - a);
}

std::string unquote(const std::string& s)
{
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') ||
                          (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

bool push_arg(const std::string& raw, DirectiveArgs& args)
{
    const std::string e = trim(raw);
    if (e.empty())
        return true;
    const size_t eq = e.find('=');
    if (eq == std::string::npos || eq == 0)
        return false;
    args.emplace_back(trim(e.substr(0, eq)), unquote(trim(e.substr(eq + 1))));
    return true;
}

bool to_bool(const std::string& v)
{
    std::string s = v;
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
    return s == "1" || s == "true" || s == "on" || s == "yes";
}

}  // namespace

bool parse_directive(const std::string& line, std::string& name, DirectiveArgs& args)
{
    const std::string s = trim(line);
    const size_t open = s.find('(');
    const size_t close = s.rfind(')');
    if (open == std::string::npos || open == 0 || close != s.size() - 1)
        return false;
    name = trim(s.substr(0, open));
    args.clear();
    std::string tok;
    char quote = 0;
    for (size_t i = open + 1; i < close; ++i) {
        const char c = s[i];
        if (quote) {
            if (c == quote)
                quote = 0;
            tok.push_back(c);
            continue;
        }
        if (c == '"' || c == '\'') {
            quote = c;
            tok.push_back(c);
            continue;
        }
        if (c == ',') {
            if (!push_arg(tok, args))
                return false;
            tok.clear();
            continue;
        }
        tok.push_back(c);
    }
    return push_arg(tok, args);
}

bool load_config(const std::string& path, CaseConfig& c, std::ostream& err)
{
    std::ifstream in(path);
    if (!in)
        return false;
    using Set = std::function<void(const std::string&)>;
    auto f = [](float& dst) { return Set([&dst](const std::string& v) { dst = std::stof(v); }); };
    auto d = [](double& dst) { return Set([&dst](const std::string& v) { dst = std::stod(v); }); };
    auto u = [](uint32_t& dst) {
        return Set([&dst](const std::string& v) { dst = static_cast<uint32_t>(std::stoul(v)); });
    };
    auto i = [](int& dst) { return Set([&dst](const std::string& v) { dst = std::stoi(v); }); };
    auto s = [](std::string& dst) { return Set([&dst](const std::string& v) { dst = v; }); };
    auto b = [](bool& dst) { return Set([&dst](const std::string& v) { dst = to_bool(v); }); };
    // directive.argument -> field (CfgOptions.cpp:19-98 subset)
    const std::map<std::string, Set> keys = {
        {"simulation.particle_count", u(c.particle_count)},
        {"simulation.particles", u(c.particle_count)},
        {"simulation.dt", f(c.dt)},
        {"simulation.solver", s(c.solver)},
        {"simulation.integrator", s(c.integrator)},
        {"performance.substep_target_dt", f(c.substep_target_dt)},
        {"performance.max_substeps", u(c.max_substeps)},
        {"octree.theta", f(c.theta)},
        {"octree.softening", f(c.softening)},
        {"octree.criterion", s(c.criterion)},
        {"octree.theta_auto", b(c.theta_auto)},
        {"octree.theta_auto_min", f(c.theta_auto_min)},
        {"octree.theta_auto_max", f(c.theta_auto_max)},
        {"octree.leaf_capacity", u(c.leaf_capacity)},
        {"physics.max_acceleration", f(c.max_acceleration)},
        {"physics.min_softening", f(c.min_softening)},
        {"physics.min_distance2", f(c.min_distance2)},
        {"physics.min_theta", f(c.min_theta)},
        {"sph.max_speed", f(c.sph_max_speed)},
        {"generation.seed", u(c.seed)},
        {"generation.deterministic", b(c.deterministic)},
        {"scene.style", s(c.scene_style)},
        {"scene.preset", s(c.preset)},
        {"scene.file", s(c.input_file)},
        {"scene.format", s(c.input_format)},
        {"export.directory", s(c.export_directory)},
        {"export.format", s(c.export_format)},
        {"treepm.enabled", b(c.treepm_enabled)},
        {"octree_hip.bucket", u(c.bucket)},
        {"octree_hip.order", i(c.multipole_order)},
        {"octree_hip.scale_radius", d(c.scale_radius)},
        {"octree_hip.separation", d(c.separation)},
        {"octree_hip.pericentre", d(c.pericentre)},
        {"octree_hip.split_scale", f(c.treepm_split_scale)},
        {"octree_hip.cutoff", f(c.treepm_cutoff)},
        {"octree_hip.far_alpha", f(c.far_alpha)},
        {"octree_hip.relative_alpha", f(c.relative_alpha)},
        {"octree_hip.relative_mono", f(c.relative_mono)},
        {"octree_hip.ic", s(c.preset)},
    };
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#' || t[0] == ';')
            continue;
        std::string name;
        DirectiveArgs args;
        if (!parse_directive(t, name, args)) {
            err << "[config] " << path << ":" << lineno << ": malformed directive skipped\n";
            continue;
        }
        for (const auto& kv : args) {
            const auto it = keys.find(name + "." + kv.first);
            if (it == keys.end())
                continue;
            try {
                it->second(kv.second);
            } catch (const std::exception&) {
                err << "[config] " << path << ":" << lineno << ": bad value for " << name << "."
                    << kv.first << "\n";
            }
        }
    }
    return true;
}

void write_config(std::ostream& o, const CaseConfig& c)
{
    o << "simulation(particle_count=" << c.particle_count << ", dt=" << c.dt
      << ", solver=" << c.solver << ", integrator=" << c.integrator << ")\n";
    o << "performance(substep_target_dt=" << c.substep_target_dt
      << ", max_substeps=" << c.max_substeps << ")\n";
    o << "octree(theta=" << c.theta << ", softening=" << c.softening
      << ", criterion=" << c.criterion << ", theta_auto=" << (c.theta_auto ? "true" : "false")
      << ", theta_auto_min=" << c.theta_auto_min << ", theta_auto_max=" << c.theta_auto_max
      << ", leaf_capacity=" << c.leaf_capacity << ")\n";
    o << "physics(max_acceleration=" << c.max_acceleration << ", min_softening="
      << c.min_softening << ", min_distance2=" << c.min_distance2
      << ", min_theta=" << c.min_theta << ")\n";
    o << "sph(max_speed=" << c.sph_max_speed << ")\n";
    o << "generation(seed=" << c.seed << ", deterministic=" << (c.deterministic ? "true" : "false")
      << ")\n";
    o << "scene(style=" << c.scene_style << ", preset=" << c.preset << ", file=\""
      << c.input_file << "\", format=" << c.input_format << ")\n";
    o << "export(directory=" << c.export_directory << ", format=" << c.export_format << ")\n";
    o << "octree_hip(bucket=" << c.bucket << ", order=" << c.multipole_order
      << ", scale_radius=" << c.scale_radius << ", separation=" << c.separation
      << ", pericentre=" << c.pericentre << ", split_scale=" << c.treepm_split_scale
      << ", cutoff=" << c.treepm_cutoff << ", far_alpha=" << c.far_alpha
      << ", relative_alpha=" << c.relative_alpha << ", relative_mono=" << c.relative_mono
      << ")\n";
}

std::string normalize_solver(const std::string& name)
{
    std::string n = trim(name);
    std::transform(n.begin(), n.end(), n.begin(), [](unsigned char ch) {
        return ch == '-' ? '_' : static_cast<char>(std::tolower(ch));
    });
    if (n == "octree_hip" || n == "octree_gpu_hip" || n == "hip")
        return "octree_hip";
    if (n == "octree" || n == "cpu" || n == "octree_cpu")
        return "octree_cpu";
    if (n == "gpu" || n == "octree_gpu")
        return "octree_gpu";
    return n;
}

float resolve_theta(const CaseConfig& c)
{
    const float lo = std::clamp(c.theta_auto_min, 0.05f, 4.0f);
    const float hi = std::max(lo, std::clamp(c.theta_auto_max, 0.05f, 4.0f));
    return std::clamp(c.theta, lo, hi);
}

float auto_target_substep_dt(bool euler, std::size_t n)
{
    if (!euler)
        return 0.0025f;
    if (n >= 100000u)
        return 0.01f;
    if (n >= 20000u)
        return 0.005f;
    return 0.0025f;
}

}  // namespace bhh
