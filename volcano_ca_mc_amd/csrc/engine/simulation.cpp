#include "simulation.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>

namespace eng {

namespace fs = std::filesystem;

namespace {

double round3(double v) { return std::round(v * 1000.0) / 1000.0; }
double rounded(double v, int decimals) {
    const double f = std::pow(10.0, decimals);
    return std::round(v * f) / f;
}
std::string fixed(double v, int decimals) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.*f", decimals, v);
    return b;
}

std::string cpu_name() {
    std::ifstream in("/proc/cpuinfo");
    std::string line;
    while (std::getline(in, line))
        if (line.rfind("model name", 0) == 0) {
            const auto c = line.find(':');
            if (c != std::string::npos) return line.substr(c + 2);
        }
    return "nieznany";
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Timing / Progress / Gpu
// ------------------------------------------------------------------------------------------------
void Timing::add(const std::string& section, double ms) {
    for (auto& s : sections_)
        if (s.first == section) {
            s.second += ms;
            return;
        }
    sections_.emplace_back(section, ms);
}
double Timing::get(const std::string& section) const {
    for (const auto& s : sections_)
        if (s.first == section) return s.second;
    return 0.0;
}
void Timing::add_file(const std::string& path, double ms) {
    for (auto& f : files_)
        if (f.first == path) {
            f.second += ms;
            return;
        }
    files_.emplace_back(path, ms);
}
void Timing::record_files() {
    Json files = Json::object();
    for (const auto& f : files_) {
        std::error_code ec;
        uint64_t bytes = 0;
        if (fs::is_directory(f.first, ec)) {
            for (const auto& e : fs::directory_iterator(f.first, ec))
                if (e.is_regular_file(ec)) bytes += e.file_size(ec);
        } else {
            bytes = fs::file_size(f.first, ec);
            if (ec) bytes = 0;
        }
        Json e = Json::object();
        e["bytes"] = Json(static_cast<unsigned long long>(bytes));
        e["writeMs"] = Json(round3(f.second));
        files[fs::path(f.first).filename().string()] = e;
    }
    set_info("files", files);
}
Json Timing::to_json() const {
    Json j = Json::object();
    j["unit"] = "ms";
    for (const auto& s : sections_) j[s.first] = Json(round3(s.second));
    for (const auto& kv : info_.items()) j[kv.first] = kv.second;
    return j;
}

void Progress::report(const char* stage, long long done, long long total, bool force, const char* key, uint64_t value) {
    // the clock runs from construction, like the reference's reporter: an unforced line needs 500 ms since the
    // previous line (or since the start)
    if (!on_ || (!force && since_.ms() < 500.0)) return;
    std::cout << "PROGRESS " << stage << ' ' << done << '/' << total;
    if (key) std::cout << ' ' << key << '=' << value;
    std::cout << '\n' << std::flush;
    since_.reset();
}

namespace {
std::atomic<bool> g_fast_forward{true};
}  // namespace

void set_fast_forward(bool on) { g_fast_forward.store(on); }

void vca_check(int rc, const char* what) {
    if (rc != VCA_OK) {
        char buf[1024];
        vca_last_error(buf, sizeof(buf));
        throw std::runtime_error(std::string(what) + ": " + buf);
    }
}

Gpu::Gpu() {
    vca_check(vca_ctx_create(0, nullptr, &ctx_), "GPU (libvca)");
    name_ = "gfx950";
}
Gpu::~Gpu() {
    if (ctx_) vca_ctx_destroy(ctx_);
}

const char* stop_text(int stop) {  // engine/src/ca/CaRules.hpp:110-119
    switch (stop) {
        case 0: return "trwa";
        case 1: return "brak płynnej lawy";
        case 2: return "brak wolnych komórek (AIR)";
        case 3: return "brak zmian w kroku";
        case 4: return "osiągnięto ca.maxSteps";
    }
    return "?";
}

// ------------------------------------------------------------------------------------------------
// CA stage
// ------------------------------------------------------------------------------------------------
namespace {

struct CaHandle {
    vca_ca* h = nullptr;
    ~CaHandle() {
        if (h) vca_ca_destroy(h);
    }
};

void download(vca_ca* ca, CaState& s) {
    vca_check(vca_ca_download(ca, s.type.data(), s.flow.data(), s.temp.data(), s.solid.data(), s.arrival.data()),
              "pobranie wyniku CA");
}

}  // namespace

// flowSums + statsFromSums (engine/src/app/Simulation.cpp:408-524)
Json ca_stats(const Config& c, const RunSummary& s) {
    const size_t flows = s.vents.size();
    std::vector<uint64_t> lava(flows + 1, 0), basalt(flows + 1, 0);
    std::vector<double> reach(flows + 1, 0.0);
    if (s.flows_ready) {
        lava = s.flow_lava;
        basalt = s.flow_basalt;
        reach = s.flow_reach;
    } else {
        const Dims& d = s.result.dims;
        size_t i = 0;
        for (int z = 0; z < d.nz; ++z)
            for (int y = 0; y < d.ny; ++y)
                for (int x = 0; x < d.nx; ++x, ++i) {
                    const uint8_t t = s.result.type[i];
                    const int32_t f = s.result.flow[i];
                    if ((t != VCA_LAVA && t != VCA_BASALT) || f <= 0 || static_cast<size_t>(f) > flows) continue;
                    const size_t fi = static_cast<size_t>(f);
                    (t == VCA_LAVA ? lava : basalt)[fi] += 1;
                    const Vent& v = s.vents[fi - 1];
                    reach[fi] = std::max(reach[fi], std::hypot(double(x - v.x), double(y - v.y)));
                }
    }
    Json j = Json::object();
    j["mode"] = c.ca.classic ? "classic" : "lava";
    j["steps"] = s.steps;
    j["stopReason"] = stop_text(s.stop);
    uint64_t total = 0, total_lava = 0, min_size = std::numeric_limits<uint64_t>::max(), max_size = 0;
    double max_reach = 0.0;
    Json list = Json::array();
    for (size_t f = 1; f <= flows; ++f) {
        const uint64_t cells = lava[f] + basalt[f];
        total += cells;
        total_lava += lava[f];
        max_reach = std::max(max_reach, reach[f]);
        min_size = std::min(min_size, cells);
        max_size = std::max(max_size, cells);
        if (f <= 10000) {
            const Vent& v = s.vents[f - 1];
            Json e = Json::object();
            e["flowId"] = Json(static_cast<unsigned long long>(f));
            e["vent"] = Json::array({Json(v.x), Json(v.y), Json(v.z)});
            e["cells"] = Json(static_cast<unsigned long long>(cells));
            if (!c.ca.classic) {
                e["lava"] = Json(static_cast<unsigned long long>(lava[f]));
                e["basalt"] = Json(static_cast<unsigned long long>(basalt[f]));
                e["maxReach"] = Json(std::round(reach[f] * 100.0) / 100.0);
            }
            list.push_back(e);
        }
    }
    if (!c.ca.classic) {
        j["totalLavaVolume"] = Json(static_cast<unsigned long long>(total));
        j["liquidLava"] = Json(static_cast<unsigned long long>(total_lava));
        j["maxReach"] = Json(std::round(max_reach * 100.0) / 100.0);
    } else {
        j["grains"] = Json(static_cast<unsigned long long>(flows));
        j["meanGrainSize"] = Json(flows ? double(total) / double(flows) : 0.0);
        j["minGrainSize"] = Json(static_cast<unsigned long long>(flows ? min_size : 0));
        j["maxGrainSize"] = Json(static_cast<unsigned long long>(max_size));
    }
    j["flows"] = list;
    if (flows > 10000) j["flowsTruncated"] = true;
    Json a = Json::array(), ch = Json::array();
    for (uint64_t v : s.active) a.push_back(Json(static_cast<unsigned long long>(v)));
    for (uint64_t v : s.changed) ch.push_back(Json(static_cast<unsigned long long>(v)));
    j["activeCells"] = a;
    j["changedCells"] = ch;
    return j;
}

vca_ca_desc make_ca_desc(const Config& cfg) {
    vca_ca_desc d{};
    d.nx = cfg.space.nx;
    d.ny = cfg.space.ny;
    d.nz = cfg.space.nz;
    d.classic = cfg.ca.classic;
    d.neighborhood = static_cast<int32_t>(cfg.neighborhood);
    d.periodic = cfg.periodic;
    d.seed = cfg.seed;
    d.max_steps = cfg.ca.max_steps;
    d.eruption_steps = cfg.vents.eruption_steps;
    // CaParams narrows the double config values to float (Simulation.cpp:122-136)
    d.erupt_temp = static_cast<float>(cfg.vents.erupt_temp);
    d.solidus_temp = static_cast<float>(cfg.ca.solidus);
    d.cooling_ambient = static_cast<float>(cfg.ca.cooling_ambient);
    d.cooling_lateral = static_cast<float>(cfg.ca.cooling_lateral);
    d.cooling_fall = static_cast<float>(cfg.ca.cooling_fall);
    d.ponding = cfg.ca.ponding;
    d.dense = cfg.ca.front ? 0 : 1;
    return d;
}

CaInputs ca_inputs(const Config& cfg) {
    // terrain + spill levels, craters (host, once per run)
    CaInputs in;
    Clock t;
    in.hm = build_terrain(cfg);
    if (!in.hm.empty()) in.spill = fill_depressions(in.hm, cfg.periodic, cfg.periodic);
    in.terrain_ms = t.ms();
    t.reset();
    in.vents = build_vents(cfg, in.hm);
    in.vents_ms = t.ms();
    return in;
}

void autotune_ca(const Config& cfg, CaInputs& in, Gpu& gpu, int probe_steps) {
    Clock t;
    const vca_ca_desc d = make_ca_desc(cfg);
    std::vector<vca_vent> vv;
    for (const Vent& v : in.vents) vv.push_back(vca_vent{v.x, v.y, v.z, v.flow_id});
    vca_check(vca_ca_autotune(gpu.ctx(), &d, in.hm.empty() ? nullptr : in.hm.height.data(),
                              in.spill.empty() ? nullptr : in.spill.height.data(), vv.data(),
                              static_cast<int32_t>(vv.size()), probe_steps, &in.tune),
              "CA autotune");
    in.tuned = true;
    in.tune_ms = t.ms();
}

Json autotune_json(const CaInputs& in) {
    static const char* kShape[5] = {"64x4x4", "64x2x8", "64x8x2", "64x1x16", "64x16x1"};
    Json j = Json::object();
    j["tileShape"] = kShape[std::max(1, std::min(5, in.tune.tile_shape)) - 1];
    j["sweep"] = in.tune.dense ? "full" : "front";
    j["probeSteps"] = in.tune.probe_steps;
    Json c = Json::object();
    for (int dn = 0; dn < 2; ++dn)
        for (int sh = 0; sh < 5; ++sh)
            if (in.tune.ms_per_step[dn][sh] > 0.0f)
                c[std::string(dn ? "full " : "front ") + kShape[sh]] = Json(double(in.tune.ms_per_step[dn][sh]));
    j["msPerStep"] = c;
    j["ms"] = Json(in.tune_ms);
    return j;
}

vca_ca* create_ca(const Config& cfg, const CaInputs& in, Gpu& gpu) {
    // GPU state: allocation, rock columns, seeds, initial counts
    vca_ca_desc d = make_ca_desc(cfg);
    if (in.tuned) {
        d.tile_shape = in.tune.tile_shape;
        d.dense = in.tune.dense;
    }
    std::vector<vca_vent> vv;
    for (const Vent& v : in.vents) vv.push_back(vca_vent{v.x, v.y, v.z, v.flow_id});
    vca_ca* h = nullptr;
    vca_check(vca_ca_create(gpu.ctx(), &d, in.hm.empty() ? nullptr : in.hm.height.data(),
                            in.spill.empty() ? nullptr : in.spill.height.data(), vv.data(), static_cast<int32_t>(vv.size()),
                            &h),
              "CA");
    vca_check(vca_ca_set_fast_forward(h, g_fast_forward.load() ? 1 : 0), "CA");
    return h;
}

bool ca_state_needed(const Config& cfg) { return cfg.output.format != OutputFormat::None; }

void ca_collect(vca_ca* ca, const Config& cfg, RunSummary& s, bool with_state) {
    vca_ca_result r{};
    // histories of at most maxSteps entries (no status read on the context's stream, which may be busy with a
    // batch)
    const int32_t cap = std::max(1, cfg.ca.max_steps);
    s.active.assign(static_cast<size_t>(cap), 0);
    s.changed.assign(static_cast<size_t>(cap), 0);
    if (with_state) {
        s.result.resize(cfg.space);
        r.type = reinterpret_cast<uint8_t*>(s.result.type.data());
        r.flow = s.result.flow.data();
        r.temp = s.result.temp.data();
        r.solid = s.result.solid.data();
        r.arrival = s.result.arrival.data();
    }
    r.active = s.active.data();
    r.changed = s.changed.data();
    r.cap = cap;
    vca_check(vca_ca_collect(ca, &r), "CA");
    s.active.resize(static_cast<size_t>(std::min(r.steps, cap)));
    s.changed.resize(static_cast<size_t>(std::min(r.steps, cap)));
    s.steps = r.steps;
    s.stop = r.stop_reason;
    s.tiles_visited = r.tiles_visited;
    s.tiles_staged = r.tiles_staged;
    s.changes = r.changes;
    s.ff_from = r.ff_from;
    s.ff_steps = r.ff_steps;
    s.tile_cells = 1024;
}

namespace {
uint64_t isqrt(uint64_t v) {
    uint64_t r = static_cast<uint64_t>(std::sqrt(static_cast<double>(v)));
    while (r > 0 && r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// max hypot(x - vx, y - vy) over the flow's cells, from their largest squared distance d2: the lattice points
// (a, b), a^2 + b^2 = d2, all lie at the same true distance; the host libm's hypot may still differ between
// them in the last place, and then the device says which of them the flow reaches.
double exact_reach(vca_ca* ca, const Vent& v, int32_t flow, uint64_t d2) {
    std::vector<std::pair<uint64_t, uint64_t>> pts;
    const uint64_t r = isqrt(d2);
    for (uint64_t a = 0; a <= r; ++a) {
        const uint64_t rest = d2 - a * a, b = isqrt(rest);
        if (b * b == rest) pts.emplace_back(a, b);
    }
    if (pts.empty()) return 0.0;
    const double h0 = std::hypot(double(pts[0].first), double(pts[0].second));
    bool same = true;
    for (const auto& p : pts) same = same && std::hypot(double(p.first), double(p.second)) == h0;
    if (same) return h0;
    std::vector<uint8_t> seen(static_cast<size_t>(r) + 1, 0);
    const vca_vent vv{v.x, v.y, v.z, flow};
    vca_check(vca_ca_flow_points(ca, &vv, flow, d2, seen.data(), static_cast<int32_t>(seen.size())), "CA");
    double best = 0.0;
    for (const auto& p : pts)
        if (seen[static_cast<size_t>(p.first)]) best = std::max(best, std::hypot(double(p.first), double(p.second)));
    return best;
}
}  // namespace

void ca_flow_sums(vca_ca* ca, RunSummary& s) {
    const size_t flows = s.vents.size();
    std::vector<vca_vent> vv(flows);
    for (size_t f = 0; f < flows; ++f) vv[f] = vca_vent{s.vents[f].x, s.vents[f].y, s.vents[f].z, int32_t(f + 1)};
    s.flow_lava.assign(flows + 1, 0);
    s.flow_basalt.assign(flows + 1, 0);
    s.flow_reach.assign(flows + 1, 0.0);
    std::vector<uint64_t> d2(flows + 1, 0);
    vca_check(vca_ca_flow_stats(ca, vv.data(), static_cast<int32_t>(flows), s.flow_lava.data(), s.flow_basalt.data(),
                                d2.data()),
              "CA");
    for (size_t f = 1; f <= flows; ++f)
        if (s.flow_lava[f] + s.flow_basalt[f] > 0)
            s.flow_reach[f] = exact_reach(ca, s.vents[f - 1], static_cast<int32_t>(f), d2[f]);
    s.flows_ready = true;
}

void save_ca(const Config& cfg, const RunSummary& s, const std::string& out_dir, Timing& timing) {
    const fs::path out(out_dir);
    Clock tf;
    if (cfg.output.format == OutputFormat::Txt || cfg.output.format == OutputFormat::Both) {
        tf.reset();
        write_ca_txt((out / "ca_result.txt").string(), s.result);
        timing.add_file((out / "ca_result.txt").string(), tf.ms());
    }
    if (cfg.output.format == OutputFormat::Bin || cfg.output.format == OutputFormat::Both) {
        tf.reset();
        write_ca_bin((out / "ca_result.bin").string(), s.result);
        timing.add_file((out / "ca_result.bin").string(), tf.ms());
    }
}

Json parallel_json(const Config& cfg) {
    Json par = Json::object();
    par["threads"] = cfg.parallel.threads;
    par["schedule"] = cfg.parallel.schedule;
    par["chunk"] = cfg.parallel.chunk;
    par["affinity"] = cfg.parallel.affinity;
    return par;
}

void write_config_used(const Config& cfg, const std::string& out_dir, Timing& timing) {
    const fs::path out(out_dir);
    Clock t;
    write_json((out / "config_used.json").string(), cfg.effective);
    timing.add_file((out / "config_used.json").string(), t.ms());
}

RunSummary run_simulation(const Config& cfg, const std::string& out_dir, Timing& timing, Progress& progress, Gpu& gpu,
                          int autotune_steps) {
    RunSummary s;
    const fs::path out(out_dir);
    write_config_used(cfg, out_dir, timing);
    CaHandle ca;  // kept until the MC has taken its domain from the final state
    if (cfg.mc.full_only()) {
        s.ca_ran = false;
    } else {
        CaInputs in = ca_inputs(cfg);
        s.vents = in.vents;
        if (autotune_steps > 0) {
            autotune_ca(cfg, in, gpu, autotune_steps);
            timing.add("autotune", in.tune_ms);
            timing.set_info("autotune", autotune_json(in));
        }
        Clock t;
        ca.h = create_ca(cfg, in, gpu);
        timing.add("allocate", t.ms());
        timing.add("terrain", in.terrain_ms);
        timing.add("vents", in.vents_ms);

        // steps; snapshots (output.snapshotsEvery) and progress between chunks, outside caSim
        const int every = cfg.output.snapshots_every;
        int32_t steps = 0, stop = 0;
        vca_check(vca_ca_status(ca.h, &steps, &stop), "CA");
        double ca_ms = 0.0, snap_ms = 0.0;
        const fs::path snap_dir = out / "snapshots";
        CaState snap;
        auto write_snapshot = [&](int k) {
            Clock ts;
            if (snap.type.empty()) snap.resize(cfg.space);
            download(ca.h, snap);
            char name[32];
            std::snprintf(name, sizeof(name), "step_%06d.bin", k);
            write_ca_bin((snap_dir / name).string(), snap);
            snap_ms += ts.ms();
        };
        auto report = [&](bool force) {
            if (!progress.enabled()) return;
            uint64_t active = 0;
            if (steps > 0) {
                std::vector<uint64_t> a(static_cast<size_t>(steps)), c(static_cast<size_t>(steps));
                int32_t n = 0;
                vca_check(vca_ca_histories(ca.h, a.data(), c.data(), steps, &n), "CA");
                active = a.back();
            }
            progress.report("ca", steps, cfg.ca.max_steps, force, "active", active);
        };
        if (every > 0) {
            fs::create_directories(snap_dir);
            write_snapshot(0);
        }
        // the step loop with snapshots reports only after steps (Simulation.cpp:226-240): no line for step 0
        if (every <= 0) report(true);
        while (stop == VCA_RUNNING) {
            int chunk = progress.enabled() ? 64 : cfg.ca.max_steps;
            if (every > 0) chunk = std::min(chunk, every - steps % every);
            Clock tc;
            vca_check(vca_ca_run(ca.h, chunk, &steps, &stop), "CA");
            ca_ms += tc.ms();
            report(stop != VCA_RUNNING);
            if (every > 0 && (steps % every == 0 || stop != VCA_RUNNING)) write_snapshot(steps);
        }
        timing.add("caSim", ca_ms);
        if (every > 0) {
            timing.add("snapshots", snap_ms);
            timing.add_file(snap_dir.string(), snap_ms);
        }
        if (progress.enabled()) progress.report("save", 0, 1, true);

        t.reset();
        ca_collect(ca.h, cfg, s, ca_state_needed(cfg));
        save_ca(cfg, s, out_dir, timing);
        timing.add("save", t.ms());
    }
    Clock ts;
    if (s.ca_ran) ca_flow_sums(ca.h, s);
    Json stats = s.ca_ran ? ca_stats(cfg, s) : Json::object();
    if (s.ca_ran) timing.add("caStats", ts.ms());
    stats["parallel"] = parallel_json(cfg);
    if (cfg.mc.enabled) {
        Clock tp;
        McPrep prep = s.ca_ran ? mc_prepare_from_ca(cfg, ca.h) : mc_prepare(cfg, nullptr, Cooling::Full);
        s.mc = run_mc(cfg, std::move(prep), tp.ms(), timing, progress, out_dir, gpu);
        s.mc_ran = true;
        if (progress.enabled()) progress.report("save", 0, 1, true);
        save_mc(cfg, s.mc, out_dir, timing);
        stats["mc"] = s.mc.stats;
    }
    Clock tw;
    write_json((out / "stats.json").string(), stats);
    timing.add_file((out / "stats.json").string(), tw.ms());
    timing.add(s.mc_ran ? "mcSave" : "save", tw.ms());
    return s;
}

namespace {

Json class_json(uint64_t cells, const ClassStat& a, const ClassStat& b) {
    Json j = Json::object();
    j["cells"] = Json(static_cast<unsigned long long>(cells));
    j["grainsStart"] = Json(static_cast<unsigned long long>(a.grains));
    j["grainsEnd"] = Json(static_cast<unsigned long long>(b.grains));
    j["meanGrainSizeStart"] = Json(rounded(a.mean(), 3));
    j["meanGrainSizeEnd"] = Json(rounded(b.mean(), 3));
    j["largestGrainEnd"] = Json(static_cast<unsigned long long>(b.largest));
    return j;
}

}  // namespace

McPrep mc_prepare_from_ca(const Config& cfg, vca_ca* ca) {
    McPrep p = mc_prepare(cfg, nullptr, Cooling::Full);  // descriptor of the full domain ...
    p.full = false;                                       // ... then the basalt domain's periodicity (x, y only)
    p.desc.periodic[2] = 0;
    p.from_ca = ca;
    p.dynamics = cfg.mc.mobility_from_cooling;
    p.mobility_used = cfg.mc.mobility_from_cooling;
    return p;
}

McPrep mc_prepare(const Config& cfg, const CaState* ca, Cooling cooling) {
    McPrep p;
    p.full = ca == nullptr;
    p.dims = p.full ? cfg.space : ca->dims;
    const Dims& dims = p.dims;
    vca_mc_desc& d = p.desc;
    d = vca_mc_desc{};
    d.nx = dims.nx;
    d.ny = dims.ny;
    d.nz = dims.nz;
    d.neighborhood = static_cast<int32_t>(cfg.neighborhood);
    // full domain wraps every axis, basalt only x and y (McStage.cpp:102-106; PottsMC.cpp:18)
    d.periodic[0] = cfg.periodic && dims.nx > 1;
    d.periodic[1] = cfg.periodic && dims.ny > 1;
    d.periodic[2] = cfg.periodic && p.full && dims.nz > 1;
    d.seed = cfg.seed;
    d.Q = cfg.mc.Q;
    d.J = cfg.mc.J;
    d.kT = cfg.mc.kT;
    d.checkerboard = cfg.mc.checkerboard;
    if (!p.full) {
        const size_t n = ca->type.size();
        p.domain.resize(n);
        for (size_t i = 0; i < n; ++i) p.domain[i] = (ca->type[i] == VCA_BASALT || ca->type[i] == VCA_LAVA) ? 1 : 0;
        if (cooling != Cooling::None) {
            // CoolingMobility::fromCooling (CoolingMobility.cpp:8-129)
            uint32_t end = 0;
            for (size_t i = 0; i < n; ++i)
                if (p.domain[i]) end = std::max<uint32_t>(end, std::max(ca->solid[i], ca->arrival[i]));
            std::vector<uint32_t> time(n, 0);
            uint32_t tmin = 0xFFFFFFFFu, tmax = 0;
            for (size_t i = 0; i < n; ++i) {
                if (!p.domain[i]) continue;
                const uint32_t sd = ca->type[i] == VCA_BASALT ? ca->solid[i] : end;
                time[i] = sd > ca->arrival[i] ? sd - ca->arrival[i] : 0;
                tmin = std::min(tmin, time[i]);
                tmax = std::max(tmax, time[i]);
            }
            p.mobility.assign(n, 0.0f);
            for (size_t i = 0; i < n; ++i) {
                if (!p.domain[i]) continue;
                if (tmax == tmin) {
                    p.mobility[i] = 1.0f;
                    continue;
                }
                const double norm = double(time[i] - tmin) / double(tmax - tmin);
                p.mobility[i] = static_cast<float>(cfg.mc.mobility_min + (1.0 - cfg.mc.mobility_min) * norm);
            }
            if (cfg.mc.mobility_from_cooling) {
                p.dynamics = true;
                p.mobility_used = true;
            }
        }
        if (cooling == Cooling::None && cfg.mc.mobility_from_cooling)
            p.warnings.push_back("wynik CA bez solidStep/arrivalStep (plik txt): mobilność ze stygnięcia wyłączona, "
                                 "MC jak dla mobilityFromCooling = false");
        if (cooling == Cooling::SolidOnly)
            p.warnings.push_back("ca_result.bin w wersji 1 nie zawiera arrivalStep, czas w stanie płynnym liczony od "
                                 "kroku 0 (= solidStep)");
    }
    return p;
}

McRun::McRun(const Config& cfg, McPrep prep, Gpu& gpu) : cfg_(cfg) {
    s_.full = prep.full;
    s_.mcs = cfg.mc.mcs;
    s_.Q = cfg.mc.Q;
    s_.dims = prep.dims;
    s_.mobility_used = prep.mobility_used;
    s_.warnings = std::move(prep.warnings);
    has_classes_ = prep.has_mobility();
    Clock t;
    if (prep.from_ca)
        vca_check(vca_mc_create_from_ca(gpu.ctx(), &prep.desc, prep.from_ca, prep.dynamics ? 1 : 0,
                                        cfg.mc.mobility_min, &h_),
                  "MC");
    else
        vca_check(vca_mc_create(gpu.ctx(), &prep.desc, s_.full ? nullptr : prep.domain.data(),
                                prep.mobility.empty() ? nullptr : prep.mobility.data(), prep.dynamics ? 1 : 0, &h_),
                  "MC");
    uint64_t dom_cells = 0;
    vca_check(vca_mc_info(h_, &colors_, &dom_cells), "MC");
    s_.domain_cells = dom_cells;
    if (dom_cells == 0)
        s_.warnings.push_back(s_.full ? "domena MC jest pusta" : "domena MC jest pusta: wynik CA nie zawiera bazaltu");
    init_ms_ = t.ms();
    t.reset();
    McRow row0;
    measure(row0, start_cls_, &largest_end_, &high_cells_);
    s_.history.push_back(row0);
    for (int c = 0; c < 2; ++c) end_cls_[c] = start_cls_[c];
    stats_ms_ += t.ms();
}

McRun::~McRun() {
    if (h_) vca_mc_destroy(h_);
}

void McRun::measure(McRow& row, ClassStat* cls, uint64_t* largest, uint64_t* high) {
    vca_mc_stats st{};
    vca_mc_class c[2]{};
    uint64_t hi = 0;
    vca_check(vca_mc_measure(h_, &st, has_classes_ ? c : nullptr, &hi), "MC");
    row.mcs = st.mcs;
    row.grains = st.grains;
    row.mean_grain_size = st.mean_grain_size;
    row.energy = st.energy;
    if (cls) {
        cls[0] = ClassStat{c[0].grains, c[0].grain_cells, c[0].largest};
        cls[1] = ClassStat{c[1].grains, c[1].grain_cells, c[1].largest};
    }
    if (largest) *largest = st.largest;
    if (high) *high = hi;
}

int McRun::next_stats_point() const {
    return std::min(cfg_.mc.mcs, (k_ / cfg_.mc.stats_every + 1) * cfg_.mc.stats_every);
}

void McRun::advance(int next) {
    next = std::min(next, cfg_.mc.mcs);
    if (next <= k_) return;
    vca_mc_counts counts{};
    Clock tc;
    vca_check(vca_mc_sweep(h_, next - k_, &counts), "MC");
    sim_ms_ += tc.ms();
    k_ = next;
    if (k_ % cfg_.mc.stats_every == 0 || k_ == cfg_.mc.mcs) {
        Clock tm;
        McRow row;
        const bool last = k_ == cfg_.mc.mcs;
        measure(row, last ? end_cls_ : nullptr, last ? &largest_end_ : nullptr, nullptr);
        const uint64_t att = counts.attempts - att_before_, acc = counts.accepted - acc_before_;
        row.accept_rate = att > 0 ? double(acc) / double(att) : 0.0;
        att_before_ = counts.attempts;
        acc_before_ = counts.accepted;
        s_.history.push_back(row);
        stats_ms_ += tm.ms();
    }
}

void McRun::download(std::vector<int32_t>& out) {
    out.resize(s_.dims.cells());
    vca_check(vca_mc_download(h_, out.data()), "MC");
}

McSummary McRun::finish(Timing& timing) {
    // stats.json "mc" section (McStage::statsJson, McStage.cpp:326-383)
    Clock t;
    vca_mc_counts total{};
    vca_check(vca_mc_counts_get(h_, &total), "MC");
    const McRow& first = s_.history.front();
    const McRow& last = s_.history.back();
    Json j = Json::object();
    j["domain"] = s_.full ? "full" : "basalt";
    j["domainCells"] = Json(static_cast<unsigned long long>(s_.domain_cells));
    j["space"] = Json::array({Json(s_.dims.nx), Json(s_.dims.ny), Json(s_.dims.nz)});
    j["Q"] = cfg_.mc.Q;
    j["mcs"] = cfg_.mc.mcs;
    j["J"] = Json(cfg_.mc.J);
    j["kT"] = Json(cfg_.mc.kT);
    j["update"] = cfg_.mc.checkerboard ? "checkerboard" : "random";
    if (cfg_.mc.checkerboard) j["colors"] = colors_;
    j["neighborhood"] = neighborhood_name(cfg_.neighborhood);
    j["boundary"] = cfg_.periodic ? "periodic" : "absorbing";
    j["grainConnectivity"] = cfg_.neighborhood != Neighborhood::VonNeumann ? "moore" : "von_neumann";
    j["mobility"] = s_.mobility_used ? "cooling" : "off";
    if (s_.mobility_used) j["mobilityMin"] = Json(cfg_.mc.mobility_min);
    j["grainsStart"] = Json(static_cast<unsigned long long>(first.grains));
    j["grainsEnd"] = Json(static_cast<unsigned long long>(last.grains));
    j["meanGrainSizeStart"] = Json(rounded(first.mean_grain_size, 3));
    j["meanGrainSizeEnd"] = Json(rounded(last.mean_grain_size, 3));
    j["largestGrainEnd"] = Json(static_cast<unsigned long long>(largest_end_));
    j["energyStart"] = Json(first.energy);
    j["energyEnd"] = Json(last.energy);
    j["attempts"] = Json(static_cast<unsigned long long>(total.attempts));
    j["accepted"] = Json(static_cast<unsigned long long>(total.accepted));
    j["acceptedUphill"] = Json(static_cast<unsigned long long>(total.uphill));
    if (has_classes_) {
        Json bm = Json::object();
        bm["threshold"] = Json(0.5);
        bm["mobilityMin"] = Json(cfg_.mc.mobility_min);
        bm["rule"] = "ziarno należy do klasy, w której jest większość jego komórek";
        bm["high"] = class_json(high_cells_, start_cls_[0], end_cls_[0]);
        bm["low"] = class_json(s_.domain_cells - high_cells_, start_cls_[1], end_cls_[1]);
        j["byMobility"] = bm;
    }
    s_.stats = j;
    // the orientations go to the host only for mc_result.bin / .txt
    if (cfg_.output.format != OutputFormat::None) download(s_.crystals);
    timing.add("mcInit", init_ms_);
    timing.add("mcSim", sim_ms_);
    timing.add("mcStats", stats_ms_ + t.ms());
    return std::move(s_);
}

McSummary run_mc(const Config& cfg, const CaState* ca, Cooling cooling, Timing& timing, Progress& progress,
                 const std::string& out_dir, Gpu& gpu) {
    Clock tp;
    McPrep prep = mc_prepare(cfg, ca, cooling);
    return run_mc(cfg, std::move(prep), tp.ms(), timing, progress, out_dir, gpu);
}

McSummary run_mc(const Config& cfg, McPrep prep, double prep_ms, Timing& timing, Progress& progress,
                 const std::string& out_dir, Gpu& gpu) {
    McRun run(cfg, std::move(prep), gpu);
    run.add_init_ms(prep_ms);
    const int snap_every = out_dir.empty() ? 0 : cfg.mc.snapshots_every;
    const fs::path snap_dir = fs::path(out_dir) / "mc_snapshots";
    double snap_ms = 0.0;
    std::vector<int32_t> snap;
    auto write_snapshot = [&](int k) {
        Clock ts;
        run.download(snap);
        char name[32];
        std::snprintf(name, sizeof(name), "mcs_%06d.bin", k);
        write_crystal_bin((snap_dir / name).string(), run.dims(), snap);
        snap_ms += ts.ms();
    };
    if (snap_every > 0) {
        fs::create_directories(snap_dir);
        write_snapshot(0);
    }
    progress.report("mc", 0, cfg.mc.mcs, true);
    while (!run.done()) {
        // sweep up to the next measurement / snapshot point
        int next = run.next_stats_point();
        const int k = run.sweeps_done();
        if (snap_every > 0) next = std::min(next, (k / snap_every + 1) * snap_every);
        if (progress.enabled()) next = std::min(next, k + 16);
        run.advance(next);
        if (snap_every > 0 && (run.sweeps_done() % snap_every == 0 || run.done())) write_snapshot(run.sweeps_done());
        progress.report("mc", run.sweeps_done(), cfg.mc.mcs, run.done());
    }
    McSummary s = run.finish(timing);
    if (snap_every > 0) {
        timing.add("mcSnapshots", snap_ms);
        timing.add_file(snap_dir.string(), snap_ms);
    }
    return s;
}

void save_mc(const Config& cfg, const McSummary& s, const std::string& out_dir, Timing& timing) {
    Clock t;
    const fs::path out(out_dir);
    Clock tf;
    if (cfg.output.format == OutputFormat::Txt || cfg.output.format == OutputFormat::Both) {
        tf.reset();
        write_crystal_txt((out / "mc_result.txt").string(), s.dims, s.crystals);
        timing.add_file((out / "mc_result.txt").string(), tf.ms());
    }
    if (cfg.output.format == OutputFormat::Bin || cfg.output.format == OutputFormat::Both) {
        tf.reset();
        write_crystal_bin((out / "mc_result.bin").string(), s.dims, s.crystals);
        timing.add_file((out / "mc_result.bin").string(), tf.ms());
    }
    tf.reset();
    write_mc_stats_csv((out / "mc_stats.csv").string(), s.history);
    timing.add_file((out / "mc_stats.csv").string(), tf.ms());
    timing.add("mcSave", t.ms());
}

// ------------------------------------------------------------------------------------------------
// CLI driver (engine/src/app/Simulation.cpp:526-683)
// ------------------------------------------------------------------------------------------------
namespace {

void common_info(Timing& timing, const Config& c, const Dims& d) {
    timing.set_info("cells", Json(static_cast<unsigned long long>(d.cells())));
    timing.set_info("space", Json::array({Json(d.nx), Json(d.ny), Json(d.nz)}));
    timing.set_info("neighborhood", neighborhood_name(c.neighborhood));
    timing.set_info("boundary", c.periodic ? "periodic" : "absorbing");
}
void mc_info(Timing& timing, const McSummary& mc) {
    timing.set_info("mcDomain", mc.full ? "full" : "basalt");
    timing.set_info("mcs", mc.mcs);
    timing.set_info("domainCells", Json(static_cast<unsigned long long>(mc.domain_cells)));
    timing.set_info("Q", mc.Q);
    timing.set_info("mcMobility", mc.mobility_used);
}
void env_info(Timing& timing, const Gpu& gpu) {
    timing.set_info("cpu", cpu_name());
    timing.set_info("device", gpu.device_name());
    timing.set_info("compiler", std::string("GCC ") + __VERSION__ + " + hipcc (gfx950)");
    timing.set_info("build", "Release");
}
}  // namespace

void run_info(Timing& timing, const Config& cfg, const RunSummary& s, const Gpu& gpu) {
    common_info(timing, cfg, cfg.space);
    if (s.ca_ran) {
        timing.set_info("steps", s.steps);
        timing.set_info("grid", cfg.grid_matrix ? "matrix" : "vector");
        timing.set_info("mode", cfg.ca.classic ? "classic" : "lava");
        timing.set_info("stopReason", stop_text(s.stop));
    } else {
        timing.set_info("mode", "mc_full");
    }
    if (s.mc_ran) mc_info(timing, s.mc);
    env_info(timing, gpu);
}

namespace {

std::string times_line(const Timing& timing) {
    std::string line = "Czasy [ms]     ";
    const char* sep = "";
    for (const auto& s : timing.sections()) {
        line += sep + s.first + ' ' + fixed(s.second, 2);
        sep = " | ";
    }
    return line + '\n';
}

}  // namespace

int run_from_cli(const CliOptions& opt, const Clock& since_start) {
    Timing timing;
    Progress progress(opt.progress);
    Clock t;
    Config cfg = Config::load(opt.config_path, opt.overrides);
    timing.add("readConfig", t.ms());
    timing.set_info("threads", cfg.parallel.threads);
    timing.set_info("schedule", cfg.parallel.schedule);
    timing.set_info("chunk", cfg.parallel.chunk);
    timing.set_info("affinity", cfg.parallel.affinity);
    for (const std::string& w : cfg.warnings) std::cerr << "Ostrzeżenie: " << w << '\n';
    std::string out_dir = opt.out_dir;
    if (out_dir.empty()) {
        const std::string stem = opt.config_path.empty() ? "default" : fs::path(opt.config_path).stem().string();
        out_dir = (fs::path("results") / stem).string();
    }
    fs::create_directories(out_dir);
    Gpu gpu;
    if (opt.timing_detail) vca_check(vca_ctx_set_timing(gpu.ctx(), std::max(1, opt.timing_every)), "GPU");
    // --timing-detail: device time of every kernel (HIP events on the engine's stream) and CA work counters
    auto detail = [&](const RunSummary* s) {
        if (!opt.timing_detail) return;
        Json d = Json::object();
        d["unit"] = "ms";
        d["sampleEvery"] = std::max(1, opt.timing_every);  // launches of every N-th CA step / MC sweep timed
        Json k = Json::object();
        for (const char* name : {"ca_front", "ca_descent", "ca_rules", "ca_finish", "ca_step", "ca_ff",
                                  "mc_color", "mc_sweep", "mc_sweeps"}) {
            vca_kernel_time kt{};
            vca_check(vca_ctx_kernel_time(gpu.ctx(), name, &kt), "GPU");
            if (kt.launches == 0) continue;
            Json e = Json::object();
            e["ms"] = Json(kt.ms);
            e["launches"] = Json(static_cast<unsigned long long>(kt.launches));
            k[name] = e;
        }
        d["kernels"] = k;
        if (s && s->ca_ran) {
            Json w = Json::object();
            w["tilesVisited"] = Json(static_cast<unsigned long long>(s->tiles_visited));
            w["tilesStaged"] = Json(static_cast<unsigned long long>(s->tiles_staged));
            w["tileCells"] = s->tile_cells;
            w["changes"] = Json(static_cast<unsigned long long>(s->changes));
            // steps the step kernels ran (the rest completed by the quiescent fast-forward)
            w["stepsLaunched"] = s->steps - s->ff_steps;
            w["fastForwardFrom"] = s->ff_from;
            w["fastForwardSteps"] = s->ff_steps;
            uint64_t lava = 0;
            for (uint64_t a : s->active) lava += a;
            w["lavaCellSteps"] = Json(static_cast<unsigned long long>(lava));
            d["caWork"] = w;
        }
        timing.set_info("detail", d);
    };

    if (!opt.mc_from.empty()) {
        // engine --mc-from: MC on a saved CA result (Simulation.cpp:334-402)
        if (cfg.mc.full)
            throw ConfigError("--mc-from krystalizuje wynik CA z pliku, więc wymaga mc.domain = \"basalt\" (podano \"full\")");
        std::vector<std::string> warnings;
        if (!cfg.mc.enabled) {
            warnings.push_back("mc.enabled = false, a w trybie --mc-from Monte Carlo jest uruchamiane zawsze");
            cfg.mc.enabled = true;
        }
        Clock tr;
        CaSource src;
        const CaState ca = read_ca_folder(opt.mc_from, &src);
        timing.add("readInput", tr.ms());
        if (ca.dims != cfg.space && cfg.space_from_user)
            warnings.push_back("wymiary z pliku " + std::to_string(ca.dims.nx) + " x " + std::to_string(ca.dims.ny) + " x " +
                               std::to_string(ca.dims.nz) + " różnią się od space z configu; obowiązują wymiary z pliku");
        cfg.space = ca.dims;
        if (cfg.neighborhood == Neighborhood::Hexagonal && !cfg.space.is2d())
            throw ConfigError("Sąsiedztwo „hexagonal” jest zdefiniowane tylko w 2D, a wynik CA z pliku jest 3D");
        Json sp = Json::object();
        sp["nx"] = ca.dims.nx;
        sp["ny"] = ca.dims.ny;
        sp["nz"] = ca.dims.nz;
        cfg.effective["space"] = sp;
        cfg.effective["mc"]["enabled"] = true;
        cfg.effective["_mcFrom"] = fs::absolute(fs::path(opt.mc_from)).lexically_normal().string();
        const fs::path out(out_dir);
        tr.reset();
        write_json((out / "config_used.json").string(), cfg.effective);
        timing.add_file((out / "config_used.json").string(), tr.ms());
        const Cooling cooling = src.has_arrival ? Cooling::Full : src.has_solid ? Cooling::SolidOnly : Cooling::None;
        McSummary mc = run_mc(cfg, &ca, cooling, timing, progress, out_dir, gpu);
        mc.warnings.insert(mc.warnings.begin(), warnings.begin(), warnings.end());
        progress.report("save", 0, 1, true);
        save_mc(cfg, mc, out_dir, timing);
        tr.reset();
        Json stats = Json::object();
        Json source = Json::object();
        source["folder"] = cfg.effective["_mcFrom"];
        source["file"] = src.path;
        source["format"] = src.format;
        source["version"] = src.version;
        source["coolingData"] = cooling == Cooling::Full ? "solidStep+arrivalStep" : cooling == Cooling::SolidOnly ? "solidStep" : "brak";
        stats["source"] = source;
        stats["parallel"] = parallel_json(cfg);
        stats["mc"] = mc.stats;
        Clock tw;
        write_json((out / "stats.json").string(), stats);
        timing.add_file((out / "stats.json").string(), tw.ms());
        timing.add("mcSave", tr.ms());
        timing.add("total", since_start.ms());
        common_info(timing, cfg, mc.dims);
        timing.set_info("mode", "mc_from");
        timing.set_info("mcFrom", opt.mc_from);
        mc_info(timing, mc);
        env_info(timing, gpu);
        detail(nullptr);
        timing.record_files();
        write_json((out / "times.json").string(), timing.to_json());
        for (const std::string& w : mc.warnings) std::cerr << "Ostrzeżenie: " << w << '\n';
        std::cout << "=== Monte Carlo na wyniku CA z pliku zakończone ===\n"
                  << "Wejście        " << opt.mc_from << '\n'
                  << "Przestrzeń     " << mc.dims.nx << " x " << mc.dims.ny << " x " << mc.dims.nz << '\n'
                  << times_line(timing) << "Wyniki         " << out_dir << '\n';
        return 0;
    }

    const RunSummary s = run_simulation(cfg, out_dir, timing, progress, gpu, opt.autotune_steps);
    timing.add("total", since_start.ms());
    const Dims& d = cfg.space;
    run_info(timing, cfg, s, gpu);
    detail(&s);
    timing.record_files();
    write_json((fs::path(out_dir) / "times.json").string(), timing.to_json());
    for (const std::string& w : s.mc.warnings) std::cerr << "Ostrzeżenie: " << w << '\n';
    std::ostringstream os;
    os << "=== Symulacja zakończona ===\n";
    os << "Przestrzeń     " << d.nx << " x " << d.ny << " x " << d.nz << " (" << (d.is2d() ? "2D" : "3D") << ", " << d.cells()
       << " komórek)\n";
    if (s.ca_ran) {
        os << "Tryb / siatka  " << (cfg.ca.classic ? "classic" : "lava") << (s.mc_ran ? " + MC" : "") << " / GPU\n";
        os << "Kroki          " << s.steps << " (koniec: " << stop_text(s.stop) << ")\n";
    } else {
        os << "Tryb           Monte Carlo na całej przestrzeni (bez terenu i CA)\n";
    }
    if (s.mc_ran)
        os << "Monte Carlo    domena " << (s.mc.full ? "full" : "basalt") << ": " << s.mc.domain_cells << " komórek, Q "
           << s.mc.Q << ", " << s.mc.mcs << " MCS\n";
    os << times_line(timing) << "Wyniki         " << out_dir << '\n';
    std::cout << os.str();
    return 0;
}

// engine --validate (engine/src/app/ConfigCheck.cpp:43-72)
int run_validate(const CliOptions& opt, std::ostream& out) {
    Json errors = Json::array(), warnings = Json::array();
    try {
        const Config cfg = Config::load(opt.config_path, opt.overrides);
        for (const std::string& w : cfg.warnings) warnings.push_back(Json(w));
        if (!opt.mc_from.empty()) {
            if (cfg.mc.full)
                errors.push_back(Json("--mc-from krystalizuje wynik CA z pliku, więc wymaga mc.domain = \"basalt\" "
                                      "(podano \"full\")"));
            if (!cfg.mc.enabled)
                warnings.push_back(Json("mc.enabled = false, a w trybie --mc-from Monte Carlo jest uruchamiane zawsze"));
            const fs::path dir(opt.mc_from);
            if (!fs::is_directory(dir))
                errors.push_back(Json("Folder wyniku CA (--mc-from) nie istnieje: " + opt.mc_from));
            else if (!fs::exists(dir / "ca_result.bin") && !fs::exists(dir / "ca_result.txt"))
                errors.push_back(Json("W folderze " + opt.mc_from + " nie ma ca_result.bin ani ca_result.txt"));
        } else if (!cfg.mc.full_only() && !cfg.ca.classic && cfg.terrain.type == TerrainType::File) {
            (void)build_terrain(cfg);
        }
    } catch (const std::exception& e) {
        errors.push_back(Json(std::string(e.what())));
    }
    Json r = Json::object();
    r["ok"] = errors.size() == 0;
    r["errors"] = errors;
    r["warnings"] = warnings;
    out << r.dump(2) << '\n';
    return errors.size() == 0 ? 0 : 1;
}

}  // namespace eng
