// Engine configuration: defaults <- JSON file <- --set overrides, validated into a typed struct.
// Same keys, defaults, ranges and merge rules as the reference (engine/src/io/Config.cpp:153-469):
// unknown keys are an error, keys starting with '_' are comments, --set values parse as JSON or else
// as a string, and the merged document is what config_used.json records.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "json.hpp"

namespace eng {

class ConfigError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

enum class Neighborhood { VonNeumann = 0, Moore = 1, Pentagonal = 2, Hexagonal = 3 };
enum class TerrainType { Cone, File };
enum class VentPlacement { Crater, Random, Positions };
enum class OutputFormat { Txt, Bin, Both, None };

struct Dims {
    int nx = 1, ny = 1, nz = 1;
    uint64_t cells() const { return static_cast<uint64_t>(nx) * ny * nz; }
    int active_axes() const { return (nx > 1) + (ny > 1) + (nz > 1); }
    bool is2d() const { return active_axes() == 2; }
    bool operator==(const Dims& o) const { return nx == o.nx && ny == o.ny && nz == o.nz; }
    bool operator!=(const Dims& o) const { return !(*this == o); }
};

struct Config {
    Dims space{200, 1, 100};
    bool grid_matrix = false;
    bool periodic = false;
    Neighborhood neighborhood = Neighborhood::Moore;
    uint64_t seed = 42;
    struct {
        TerrainType type = TerrainType::Cone;
        double peak = 70, base = 90, crater_r = 6, crater_d = 4, noise = 0.05;
        std::string file;
    } terrain;
    struct {
        int count = 1;
        VentPlacement placement = VentPlacement::Crater;
        int eruption_steps = 300;
        double erupt_temp = 1200;
    } vents;
    struct {
        bool classic = false;
        int max_steps = 5000;
        double cooling_ambient = 0.3, cooling_lateral = 1.5, cooling_fall = 0.1, solidus = 800;
        bool ponding = true;
        bool front = true;
    } ca;
    struct {
        bool enabled = false;
        bool full = false;  // domain "full" (else "basalt")
        int Q = 50, mcs = 100;
        double J = 1.0, kT = 0.0;
        bool mobility_from_cooling = true;
        bool checkerboard = false;
        double mobility_min = 0.05;
        int stats_every = 10, snapshots_every = 0;
        bool full_only() const { return enabled && full; }
    } mc;
    struct {
        OutputFormat format = OutputFormat::Txt;
        int snapshots_every = 0;
    } output;
    struct {
        int threads = 1;
        std::string schedule = "static";
        int chunk = 0;
        std::string affinity = "none";
    } parallel;

    std::string source_path;
    bool space_from_user = false;
    std::vector<std::string> warnings;
    Json effective;

    static Json defaults();
    static Config load(const std::string& path, const std::vector<std::string>& overrides);
    static Config from_json(const Json& j);
    static void apply_override(Json& j, const std::string& key_value);
    static void merge(Json& base, const Json& patch, const std::string& prefix);
    std::vector<std::pair<int, int>> vent_columns() const;
};

const char* neighborhood_name(Neighborhood n);

}  // namespace eng
