// One engine run "config -> files" on the GPU: terrain -> craters -> lava CA (libvca) -> CA statistics
// -> Potts MC on the basalt (libvca) -> result files.  Mirrors the reference's Simulation
// (engine/src/app/Simulation.cpp:158-683) and McStage (engine/src/app/McStage.cpp:93-414); the CA step
// and the MC sweep run as HIP kernels through include/vca.h.
#pragma once

#include <chrono>
#include <cstdint>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "config.hpp"
#include "json.hpp"
#include "terrain.hpp"
#include "vca.h"
#include "writer.hpp"

namespace eng {

struct CliOptions {
    bool help = false, timing_detail = false, print_defaults = false, validate = false, progress = false;
    bool thread_stats = false, benchmark = false, quick = false;
    int repeats = 3;
    std::string config_path, out_dir, mc_from, batch_dir;
    std::vector<std::string> overrides;
};

class CliError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Clock {
public:
    Clock() : t0_(std::chrono::steady_clock::now()) {}
    void reset() { t0_ = std::chrono::steady_clock::now(); }
    double ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count(); }

private:
    std::chrono::steady_clock::time_point t0_;
};

// times.json: named sections in ms (insertion order), then run information (engine/src/io/TimingReport.cpp).
class Timing {
public:
    void add(const std::string& section, double ms);
    double get(const std::string& section) const;
    void set_info(const std::string& key, Json v) { info_[key] = std::move(v); }
    void add_file(const std::string& path, double ms);
    void record_files();
    Json to_json() const;
    const std::vector<std::pair<std::string, double>>& sections() const { return sections_; }

private:
    std::vector<std::pair<std::string, double>> sections_;
    Json info_ = Json::object();
    std::vector<std::pair<std::string, double>> files_;
};

// "PROGRESS <stage> <done>/<total>[ key=value]" lines for the GUI (engine/src/app/Progress.hpp:1-54).
class Progress {
public:
    explicit Progress(bool on) : on_(on) {}
    bool enabled() const { return on_; }
    void report(const char* stage, long long done, long long total, bool force = false, const char* key = nullptr,
                uint64_t value = 0);

private:
    bool on_;
    Clock since_;
    bool first_ = true;
};

// RAII owner of a libvca context on device 0.
class Gpu {
public:
    Gpu();
    ~Gpu();
    vca_ctx* ctx() const { return ctx_; }
    std::string device_name() const { return name_; }

private:
    vca_ctx* ctx_ = nullptr;
    std::string name_;
};
void vca_check(int rc, const char* what);

enum class Cooling { None, SolidOnly, Full };

struct McSummary {
    bool full = false;
    uint64_t domain_cells = 0;
    int mcs = 0, Q = 0;
    bool mobility_used = false;
    std::vector<McRow> history;
    std::vector<int32_t> crystals;
    Dims dims;
    Json stats;
    std::vector<std::string> warnings;
};

struct RunSummary {
    bool ca_ran = true;
    int steps = 0;
    int stop = 0;
    std::vector<Vent> vents;
    std::vector<uint64_t> active, changed;
    uint64_t tiles_visited = 0, changes = 0;
    int32_t tile_cells = 0;
    CaState result;
    bool mc_ran = false;
    McSummary mc;
};

const char* stop_text(int stop);
Json ca_stats(const Config& c, const RunSummary& s);
McSummary run_mc(const Config& cfg, const CaState* ca, Cooling cooling, Timing& timing, Progress& progress,
                 const std::string& out_dir, Gpu& gpu);
void save_mc(const Config& cfg, const McSummary& s, const std::string& out_dir, Timing& timing);
RunSummary run_simulation(const Config& cfg, const std::string& out_dir, Timing& timing, Progress& progress, Gpu& gpu);
int run_from_cli(const CliOptions& opt, const Clock& since_start);
int run_validate(const CliOptions& opt, std::ostream& out);
std::string usage();
CliOptions parse_cli(int argc, const char* const* argv);

}  // namespace eng
