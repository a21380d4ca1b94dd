// One engine run "config -> files" on the GPU: terrain -> craters -> lava CA (libvca) -> CA statistics
// -> Potts MC on the basalt (libvca) -> result files.  Mirrors the reference's Simulation
// (engine/src/app/Simulation.cpp:158-683) and McStage (engine/src/app/McStage.cpp:93-414); the CA step
// and the MC sweep run as HIP kernels through include/vca.h.
#pragma once

#include <chrono>
#include <filesystem>
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
    int timing_every = 1;             // --timing-sample N: --timing-detail on every N-th CA step / MC sweep
    bool thread_stats = false, benchmark = false, quick = false;
    bool gpu_batch = false;           // --gpu-batch: --batch in one process through the batched pipeline
    int steps_per_graph = 16;         // --graph-steps: hipGraph length of the batched CA steps (0 = direct)
    int max_jobs = 0;                 // --max-jobs: jobs admitted at once in the batched pipeline (0 = no limit)
    int autotune_steps = 0;           // --autotune: probe steps per launch configuration before the CA (0 = off)
    bool fast_forward = true;         // --no-fast-forward: every CA step launched (vca_ca_set_fast_forward)
    bool digests = false;             // --digests: result files replaced by <file>.sha256 (digest and size)
    std::string serve_path;           // --serve <jobs.jsonl>
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
// The quiescent fast-forward of every simulation this process creates (on unless --no-fast-forward).
void set_fast_forward(bool on);

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
    uint64_t tiles_visited = 0, tiles_staged = 0, changes = 0;
    int32_t ff_from = -1, ff_steps = 0;  // quiescent fast-forward: first step, steps completed
    int32_t tile_cells = 0;
    CaState result;               // the final state, when something on the host needs it (result files)
    // flowSums of the final state (per flowId, entry 0 unused), computed on the device
    bool flows_ready = false;
    std::vector<uint64_t> flow_lava, flow_basalt;
    std::vector<double> flow_reach;
    bool mc_ran = false;
    McSummary mc;
};

// MC inputs derived on the host from the CA result: descriptor, basalt domain, cooling mobility
// (CoolingMobility.cpp:8-129), warnings.
struct McPrep {
    bool full = false;
    Dims dims;
    vca_mc_desc desc{};
    std::vector<uint8_t> domain;
    std::vector<float> mobility;
    vca_ca* from_ca = nullptr;    // domain and cooling mobility taken on the device from this stopped CA
    bool dynamics = false, mobility_used = false;
    std::vector<std::string> warnings;
    bool has_mobility() const { return from_ca != nullptr || !mobility.empty(); }
};
McPrep mc_prepare(const Config& cfg, const CaState* ca, Cooling cooling);
// The MC of a run whose CA result is still on the device (Cooling::Full): nothing is downloaded.
McPrep mc_prepare_from_ca(const Config& cfg, vca_ca* ca);

struct ClassStat {
    uint64_t grains = 0, cells = 0, largest = 0;
    double mean() const { return grains > 0 ? double(cells) / double(grains) : 0.0; }
};

// One MC stage on the GPU (McStage::run, McStage.cpp:93-301): created with the measurement at MCS 0, advanced
// between measurement points (every statsEvery and the last sweep), then finished into the stats.json section
// and the crystal orientations.  The solo driver and the batched pipeline both use it.
class McRun {
public:
    McRun(const Config& cfg, McPrep prep, Gpu& gpu);
    ~McRun();
    McRun(const McRun&) = delete;
    McRun& operator=(const McRun&) = delete;
    int next_stats_point() const;
    void advance(int next);  // sweeps up to `next` (<= mcs), measuring when it is a measurement point
    bool done() const { return k_ >= cfg_.mc.mcs; }
    int sweeps_done() const { return k_; }
    const Dims& dims() const { return s_.dims; }
    uint64_t domain_cells() const { return s_.domain_cells; }
    void download(std::vector<int32_t>& out);
    void add_init_ms(double ms) { init_ms_ += ms; }
    McSummary finish(Timing& timing);

private:
    void measure(McRow& row, ClassStat* cls, uint64_t* largest, uint64_t* high);
    const Config& cfg_;
    McSummary s_;
    vca_mc* h_ = nullptr;
    int32_t colors_ = 0;
    bool has_classes_ = false;
    int k_ = 0;
    double init_ms_ = 0.0, sim_ms_ = 0.0, stats_ms_ = 0.0;
    ClassStat start_cls_[2], end_cls_[2];
    uint64_t largest_end_ = 0, high_cells_ = 0, att_before_ = 0, acc_before_ = 0;
};

// Pieces of one run, shared by the solo driver and the batched pipeline.
struct CaInputs {
    Heightmap hm, spill;
    std::vector<Vent> vents;
    double terrain_ms = 0.0, vents_ms = 0.0;
    bool tuned = false;               // tile shape and front mode chosen by vca_ca_autotune
    vca_tune_result tune{};
    double tune_ms = 0.0;
};
// Probes the launch configurations on this run's inputs (vca_ca_autotune) and keeps the fastest in `in`.
void autotune_ca(const Config& cfg, CaInputs& in, Gpu& gpu, int probe_steps);
Json autotune_json(const CaInputs& in);
vca_ca_desc make_ca_desc(const Config& cfg);
CaInputs ca_inputs(const Config& cfg);
vca_ca* create_ca(const Config& cfg, const CaInputs& in, Gpu& gpu);
// Whether anything on the host needs the final CA state (result files).
bool ca_state_needed(const Config& cfg);
// Steps, histories, work counters and (with_state) the final state of a stopped simulation that is out of any
// batch; reads only on the simulation's own stream (safe from a worker thread while a batch goes on).
void ca_collect(vca_ca* ca, const Config& cfg, RunSummary& s, bool with_state);
// flowSums (Simulation.cpp:408-466) on the device into s.flow_*.
void ca_flow_sums(vca_ca* ca, RunSummary& s);
void save_ca(const Config& cfg, const RunSummary& s, const std::string& out_dir, Timing& timing);
Json parallel_json(const Config& cfg);
void write_config_used(const Config& cfg, const std::string& out_dir, Timing& timing);
void run_info(Timing& timing, const Config& cfg, const RunSummary& s, const Gpu& gpu);

const char* stop_text(int stop);
Json ca_stats(const Config& c, const RunSummary& s);
McSummary run_mc(const Config& cfg, const CaState* ca, Cooling cooling, Timing& timing, Progress& progress,
                 const std::string& out_dir, Gpu& gpu);
McSummary run_mc(const Config& cfg, McPrep prep, double prep_ms, Timing& timing, Progress& progress,
                 const std::string& out_dir, Gpu& gpu);
void save_mc(const Config& cfg, const McSummary& s, const std::string& out_dir, Timing& timing);
RunSummary run_simulation(const Config& cfg, const std::string& out_dir, Timing& timing, Progress& progress, Gpu& gpu,
                          int autotune_steps = 0);
int run_from_cli(const CliOptions& opt, const Clock& since_start);

// --batch helpers (batch.cpp) and the batched pipeline (pipeline.cpp)
struct BatchItem {
    std::string name;
    std::filesystem::path config, out;
};
struct BatchOutcome {
    BatchItem item;
    int exit_code = -1;
    double wall_ms = 0.0;
    std::string error;
    bool ok() const { return exit_code == 0 && error.empty(); }
};
std::vector<BatchItem> batch_items(const CliOptions& opt, std::filesystem::path& out);
void write_batch_summary(const std::filesystem::path& csv, const std::vector<BatchOutcome>& outcomes);
int run_batch(const CliOptions& opt);
int run_batched(const CliOptions& opt);
int run_serve(const CliOptions& opt);
int run_validate(const CliOptions& opt, std::ostream& out);
std::string usage();
CliOptions parse_cli(int argc, const char* const* argv);

}  // namespace eng
