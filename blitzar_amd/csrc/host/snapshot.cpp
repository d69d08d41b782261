// NBSIMBIN v1 and xyz snapshot I/O (see snapshot.h).
#include "snapshot.h"

#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

namespace bhh {
namespace {

constexpr char kMagic[8] = {'N', 'B', 'S', 'I', 'M', 'B', 'I', 'N'};
constexpr uint32_t kVersion = 1;
// Particle counts a snapshot may declare: at least 2, as the reference's
// import check (engine/server/simulation/runtime/SrvHelpers.cpp:207-210), and
// at most what one solver handle takes (2^31 - 1). The reference's 2M import
// cap (SrvConfigurationInternal.hpp:20) is not kept: 8M states are read.
constexpr uint64_t kMinCount = 2;
constexpr uint64_t kMaxCount = 0x7fffffffull;
// Shortest xyz row: "P 0 0 0 1" and a newline.
constexpr uint64_t kMinXyzRow = 10;

uint64_t file_size(const std::string& path)
{
    std::error_code ec;
    const auto n = std::filesystem::file_size(path, ec);
    return ec ? 0 : static_cast<uint64_t>(n);
}

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t count;
};
struct Record {
    float px, py, pz, vx, vy, vz, mass, temperature;
};
static_assert(sizeof(Header) == 16 && sizeof(Record) == 32, "NBSIMBIN v1 layout");

}  // namespace

std::string guess_format(const std::string& path)
{
    const std::string ext = std::filesystem::path(path).extension().string();
    if (ext == ".xyz")
        return "xyz";
    return "bin";
}

bool read_snapshot(const std::string& path, const std::string& format, State& s, std::string& err)
{
    const std::string fmt = format == "auto" ? guess_format(path) : format;
    if (fmt == "bin") {
        std::ifstream in(path, std::ios::binary);
        Header h{};
        if (!in.read(reinterpret_cast<char*>(&h), sizeof(h)) ||
            std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion) {
            err = "not an NBSIMBIN v1 file: " + path;
            return false;
        }
        // The declared count is checked against the bounds and the file size
        // before anything is allocated.
        const uint64_t need = sizeof(Header) + static_cast<uint64_t>(h.count) * sizeof(Record);
        if (h.count < kMinCount || h.count > kMaxCount || need > file_size(path)) {
            err = "NBSIMBIN header declares " + std::to_string(h.count) +
                  " particles, outside [2, 2^31 - 1] or beyond the file size: " + path;
            return false;
        }
        std::vector<Record> r(h.count);
        if (h.count && !in.read(reinterpret_cast<char*>(r.data()),
                                static_cast<std::streamsize>(r.size() * sizeof(Record)))) {
            err = "truncated NBSIMBIN file: " + path;
            return false;
        }
        s.pos.resize(3 * r.size());
        s.vel.resize(3 * r.size());
        s.mass.resize(r.size());
        s.temp.resize(r.size());
        for (size_t i = 0; i < r.size(); ++i) {
            s.pos[3 * i] = r[i].px;
            s.pos[3 * i + 1] = r[i].py;
            s.pos[3 * i + 2] = r[i].pz;
            s.vel[3 * i] = r[i].vx;
            s.vel[3 * i + 1] = r[i].vy;
            s.vel[3 * i + 2] = r[i].vz;
            s.mass[i] = r[i].mass;
            s.temp[i] = r[i].temperature;
        }
        return true;
    }
    if (fmt == "xyz") {
        std::ifstream in(path);
        size_t n = 0;
        std::string line;
        if (!(in >> n) || !std::getline(in, line) || !std::getline(in, line)) {
            err = "bad xyz header: " + path;
            return false;
        }
        if (n < kMinCount || n > kMaxCount || n > file_size(path) / kMinXyzRow) {
            err = "xyz header declares " + std::to_string(n) +
                  " particles, outside [2, 2^31 - 1] or beyond the file size: " + path;
            return false;
        }
        s.pos.assign(3 * n, 0.0);
        s.vel.assign(3 * n, 0.0);
        s.mass.assign(n, 0.0f);
        s.temp.assign(n, 0.0f);
        for (size_t i = 0; i < n; ++i) {
            if (!std::getline(in, line)) {
                err = "truncated xyz file: " + path;
                return false;
            }
            std::istringstream ls(line);
            std::string tag;
            double x, y, z, m, t = 0.0;
            if (!(ls >> tag >> x >> y >> z >> m)) {
                err = "bad xyz row in " + path;
                return false;
            }
            ls >> t;
            s.pos[3 * i] = x;
            s.pos[3 * i + 1] = y;
            s.pos[3 * i + 2] = z;
            s.mass[i] = static_cast<float>(m);
            s.temp[i] = static_cast<float>(t);
        }
        return true;
    }
    err = "unsupported snapshot format: " + fmt;
    return false;
}

bool write_snapshot(const std::string& path, const std::string& format, const State& s,
                    const std::string& label, std::string& err)
{
    const std::filesystem::path p(path);
    if (p.has_parent_path()) {
        std::error_code ec;
        std::filesystem::create_directories(p.parent_path(), ec);
    }
    const std::string fmt = format == "auto" ? guess_format(path) : format;
    const size_t n = s.size();
    if (fmt == "bin") {
        std::ofstream out(path, std::ios::binary | std::ios::trunc);
        Header h{};
        std::memcpy(h.magic, kMagic, 8);
        h.version = kVersion;
        h.count = static_cast<uint32_t>(n);
        out.write(reinterpret_cast<const char*>(&h), sizeof(h));
        std::vector<Record> r(n);
        for (size_t i = 0; i < n; ++i)
            r[i] = Record{static_cast<float>(s.pos[3 * i]), static_cast<float>(s.pos[3 * i + 1]),
                          static_cast<float>(s.pos[3 * i + 2]), static_cast<float>(s.vel[3 * i]),
                          static_cast<float>(s.vel[3 * i + 1]), static_cast<float>(s.vel[3 * i + 2]),
                          s.mass[i], s.temp.empty() ? 0.0f : s.temp[i]};
        out.write(reinterpret_cast<const char*>(r.data()),
                  static_cast<std::streamsize>(r.size() * sizeof(Record)));
        if (!out) {
            err = "write failed: " + path;
            return false;
        }
        return true;
    }
    if (fmt == "xyz") {
        std::ofstream out(path, std::ios::trunc);
        out << n << "\n" << label << "\n";
        out.precision(9);
        for (size_t i = 0; i < n; ++i)
            out << "P " << s.pos[3 * i] << " " << s.pos[3 * i + 1] << " " << s.pos[3 * i + 2]
                << " " << s.mass[i] << " " << (s.temp.empty() ? 0.0f : s.temp[i]) << "\n";
        if (!out) {
            err = "write failed: " + path;
            return false;
        }
        return true;
    }
    err = "unsupported export format: " + fmt;
    return false;
}

}  // namespace bhh
