// Case configuration for the octree_hip headless runner: the directive INI
// grammar of the reference (engine/config/directive/parsing/CfgParser.cpp:38-90)
// and the subset of its keys the solver path reads
// (engine/config/directive/parsing/CfgOptions.cpp:19-98), so the reference's
// simulation.ini cases load unchanged. Unknown directives and keys are kept
// out of the way (ignored), like the reference does for options it does not
// own.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <string>
#include <utility>
#include <vector>

namespace bhh {

using DirectiveArgs = std::vector<std::pair<std::string, std::string>>;

// `name(k=v, k2="v 2")` -> name, args. Returns false on a malformed line.
bool parse_directive(const std::string& line, std::string& name, DirectiveArgs& args);

struct CaseConfig {
    // simulation(...)
    uint32_t particle_count = 10000;
    float dt = 0.01f;                 // FndConstants.hpp:17
    std::string solver = "pairwise_cuda";
    std::string integrator = "euler";
    // performance(...)
    float substep_target_dt = 0.0f;
    uint32_t max_substeps = 4;
    // octree(...)
    float theta = 1.2f;
    float softening = 2.5f;
    std::string criterion = "com";
    bool theta_auto = false;
    float theta_auto_min = 0.4f;
    float theta_auto_max = 1.2f;
    uint32_t leaf_capacity = 256;
    // physics(...)
    float max_acceleration = 64.0f;   // FndConstants.hpp:21
    float min_softening = 1.0e-4f;
    float min_distance2 = 1.0e-12f;
    float min_theta = 0.05f;
    // sph(...)
    float sph_max_speed = 120.0f;     // FndConstants.hpp:28
    // generation / scene / export
    uint32_t seed = 42;
    bool deterministic = true;
    std::string scene_style = "preset";
    std::string preset = "plummer";
    std::string input_file;
    std::string input_format = "auto";
    std::string export_directory = "exports";
    std::string export_format = "vtk";
    // treepm(...)
    bool treepm_enabled = false;
    float treepm_split_scale = 0.0f;  // octree_hip(split_scale=...)
    float treepm_cutoff = 0.0f;       // octree_hip(cutoff=...)
    float far_alpha = 0.0f;           // octree_hip(far_alpha=...): group far field
    float relative_alpha = 0.0f;      // octree_hip(relative_alpha=...): 0 = solver default
    float relative_mono = 0.0f;       // octree_hip(relative_mono=...): point-mass cells, 0 = off
    // octree_hip(...): this solver's own directive (ignored by the reference)
    uint32_t bucket = 16;
    int multipole_order = 2;
    double scale_radius = 1.0;
    double separation = 20.0;
    double pericentre = 2.0;
};

// Load a directive file into cfg. Returns false when the file cannot be read;
// malformed lines are reported on `err` and skipped.
bool load_config(const std::string& path, CaseConfig& cfg, std::ostream& err);

// Write the effective configuration in directive form.
void write_config(std::ostream& out, const CaseConfig& cfg);

// Solver name normalisation (engine/config/modes/normalization/CfgNormalize.cpp:46-67
// plus the octree_hip aliases).
std::string normalize_solver(const std::string& name);

// resolveOctreeTheta with auto-tune off
// (engine/server/simulation/runtime/SrvFormatAndTheta.cpp:54-69): the
// configured theta clamped into [theta_auto_min, theta_auto_max], each bound
// first clamped into [0.05, 4].
float resolve_theta(const CaseConfig& cfg);

// autoTargetSubstepDt (engine/server/simulation/runtime/SrvHelpers.cpp:280-302)
// for the octree_hip solver (treated like octree_gpu).
float auto_target_substep_dt(bool euler, std::size_t n);

}  // namespace bhh
