// Snapshot I/O for the octree_hip headless runner: the reference's NBSIMBIN
// v1 binary format (engine/server/simulation/persistence/SrvPersistenceInternal.hpp:21-39,
// engine/server/simulation/persistence/SrvIO.cpp:131-166) and its ASCII xyz
// format (SrvIO.cpp:167-177, tests/data/two_body_rest.xyz). There is no
// particle-count cap (the reference caps imports at 2,000,000,
// SrvConfigurationInternal.hpp:20).
#pragma once

#include <string>
#include <vector>

namespace bhh {

struct State {
    std::vector<double> pos;   // n x 3
    std::vector<double> vel;   // n x 3
    std::vector<float> mass;   // n
    std::vector<float> temp;   // n (carried through, not evolved)
    size_t size() const { return mass.size(); }
};

// "bin" (NBSIMBIN v1) or "xyz"; "auto" picks by extension.
bool read_snapshot(const std::string& path, const std::string& format, State& s, std::string& err);
bool write_snapshot(const std::string& path, const std::string& format, const State& s,
                    const std::string& label, std::string& err);
std::string guess_format(const std::string& path);

}  // namespace bhh
