// Batched pipeline: many engine jobs in one process.  Their CA phases advance together (one batched launch
// per kernel and step for every running job, vca_batch_* over hipGraph-captured steps); host stages (config,
// terrain, spill levels, result files, statistics) run on worker threads; MC phases are interleaved on the
// GPU thread between CA chunks.  Jobs are admitted as they arrive and as device memory allows (estimated
// footprint against the free HBM), so a job arriving while others run joins the running batch.  Every job
// writes exactly the files a solo `engine --config` run writes.
//
//   engine --batch <folder> --gpu-batch --out <folder>     the --batch series, in-process (batch_summary.csv)
//   engine --serve <jobs.jsonl> --out <folder>             a stream: one {"at": s, "config": path[, "name": n]}
//                                                          per line, admitted from `at` seconds after start;
//                                                          serve_summary.json with p50/p99 latency, jobs/s
#include <algorithm>
#include <atomic>
#include <cmath>
#include <deque>
#include <filesystem>
#include <fstream>
#include <future>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <thread>

#include "simulation.hpp"

namespace eng {

namespace fs = std::filesystem;

namespace {

enum class Stage { Waiting, Loading, Ready, Creating, Created, Ca, PostCa, McReady, Mc, Writing, Done, Failed };

struct Job {
    std::string name;
    fs::path config, out;
    double at = 0.0;  // arrival, seconds after the start
    Stage stage = Stage::Waiting;
    std::future<void> work;
    Config cfg;
    Timing timing;
    CaInputs in;
    RunSummary s;
    McPrep prep;
    double prep_ms = 0.0;
    vca_ca* ca = nullptr;
    int slot = -1;
    int mc_stream = -1;
    Json stats = Json::object();
    uint64_t fp_ca = 0, fp_mc = 0;
    Clock ca_clock;
    double latency_ms = 0.0;
    double done_s = 0.0;
    std::string error;
    int exit_code = 0;    // what `engine --config` would have exited with (1 config, 2 arguments, 3 other)
};

uint64_t mc_footprint(const Config& cfg, const Dims& d) {
    // spins (<= 4 B), domain upload, union-find parent/size/hi (12 B) per cell; list + threshold per domain cell
    const uint64_t n = d.cells();
    const int spin = cfg.mc.Q <= 255 ? 1 : cfg.mc.Q <= 65535 ? 2 : 4;
    return n * (spin + 1 + 12 + 4) + n * 12 + (16u << 20);
}

class Pipeline {
public:
    Pipeline(const CliOptions& opt, std::vector<std::unique_ptr<Job>>& jobs) : opt_(opt), jobs_(jobs) {
        uint64_t free_b = 0, total_b = 0;
        vca_check(vca_device_memory(gpu_.ctx(), &free_b, &total_b), "GPU");
        budget_ = static_cast<uint64_t>(static_cast<double>(free_b) * 0.8);
        vca_check(vca_batch_create(gpu_.ctx(), kSlots, (1u << 22) - 1u, &batch_), "GPU");
        vca_check(vca_batch_set_graph(batch_, std::max(0, opt.steps_per_graph)), "GPU");
        workers_ = std::max(2u, std::min(16u, std::thread::hardware_concurrency()));
    }
    ~Pipeline() {
        for (auto& j : jobs_) {
            if (j->work.valid()) j->work.wait();
            if (j->ca) {
                if (j->slot >= 0) vca_batch_remove(batch_, j->ca);
                vca_ca_destroy(j->ca);
            }
        }
        vca_batch_destroy(batch_);
    }

    void run() {
        start_.reset();
        const Clock& start = start_;
        size_t finished = 0;
        while (finished < jobs_.size()) {
            const double now = start.ms() / 1000.0;
            bool busy = false;
            Clock ph;
            start_loading(now);
            spawn_ms_ += ph.ms();
            harvest();
            phase_ms_[0] += ph.ms();
            ph.reset();
            busy |= admit();
            phase_ms_[1] += ph.ms();
            ph.reset();
            busy |= step_ca();
            phase_ms_[2] += ph.ms();
            ph.reset();
            busy |= step_mc();
            phase_ms_[3] += ph.ms();
            finished = 0;
            for (auto& j : jobs_) {
                if ((j->stage == Stage::Done || j->stage == Stage::Failed) && j->done_s == 0.0) {
                    j->done_s = start.ms() / 1000.0;
                    report(*j);
                }
                finished += j->stage == Stage::Done || j->stage == Stage::Failed;
            }
            if (!busy && finished < jobs_.size()) std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
        makespan_s_ = start.ms() / 1000.0;
    }
    double makespan_s() const { return makespan_s_; }
    // host time of the GPU thread per phase: host stages bookkeeping, admission (CA creation), CA chunks
    // (enqueue, stop rules, result collection), MC intervals (creation, sweeps, measurements)
    Json phases() const {
        Json j = Json::object();
        j["harvestMs"] = Json(phase_ms_[0]);
        j["admitMs"] = Json(phase_ms_[1]);
        j["caMs"] = Json(phase_ms_[2]);
        j["caCollectMs"] = Json(collect_ms_);
        j["loadSpawnMs"] = Json(spawn_ms_);
        j["releaseMs"] = Json(release_ms_);
        j["batchAddMs"] = Json(create_ms_);
        j["caEnqueueMs"] = Json(enqueue_ms_);
        j["caStatusMs"] = Json(status_ms_);
        j["mcMs"] = Json(phase_ms_[3]);
        return j;
    }

private:
    static constexpr int kSlots = 1024;

    // latency: from the job's arrival (not from when a worker picked it up)
    double since_arrival(const Job& j) const { return start_.ms() - j.at * 1000.0; }

    int in_flight() const {
        int n = 0;
        for (auto& j : jobs_) n += j->work.valid();
        return n;
    }

    // host stage 1: config, output folder, config_used.json, terrain / spill levels / craters
    void start_loading(double now) {
        int free_workers = static_cast<int>(workers_) - in_flight();
        int loaded = 0;  // bound the host state of admitted-but-waiting jobs
        for (auto& j : jobs_) loaded += j->stage == Stage::Ready;
        for (auto& jp : jobs_) {
            Job& j = *jp;
            if (free_workers <= 0 || loaded >= 2 * static_cast<int>(workers_)) break;
            if (j.stage != Stage::Waiting || j.at > now) continue;
            j.stage = Stage::Loading;
            --free_workers;
            ++loaded;
            j.work = std::async(std::launch::async, [this, &j] {
                Clock t;
                j.cfg = Config::load(j.config.string(), opt_.overrides);
                j.timing.add("readConfig", t.ms());
                j.timing.set_info("threads", j.cfg.parallel.threads);
                j.timing.set_info("schedule", j.cfg.parallel.schedule);
                j.timing.set_info("chunk", j.cfg.parallel.chunk);
                j.timing.set_info("affinity", j.cfg.parallel.affinity);
                fs::create_directories(j.out);
                write_config_used(j.cfg, j.out.string(), j.timing);
                if (!j.cfg.mc.full_only()) {
                    j.in = ca_inputs(j.cfg);
                    j.s.vents = j.in.vents;
                    j.fp_ca = 0;
                    const vca_ca_desc d = make_ca_desc(j.cfg);
                    vca_check(vca_ca_footprint(&d, &j.fp_ca), "CA");
                } else {
                    j.s.ca_ran = false;
                }
                j.fp_mc = j.cfg.mc.enabled ? mc_footprint(j.cfg, j.cfg.space) : 0;
            });
        }
    }

    // finished host stages move their job on
    void harvest() {
        for (auto& jp : jobs_) {
            Job& j = *jp;
            if (!j.work.valid() || j.work.wait_for(std::chrono::seconds(0)) != std::future_status::ready) continue;
            try {
                j.work.get();
            } catch (const std::exception& e) {
                if (j.stage == Stage::Creating) used_ -= j.fp_ca;
                if (j.stage == Stage::Mc) {
                    mc_busy_[static_cast<size_t>(j.mc_stream)] = 0;
                    used_ -= j.fp_mc;
                }
                fail(j, e);
                continue;
            }
            switch (j.stage) {
                case Stage::Loading:
                    j.stage = j.s.ca_ran ? Stage::Ready : (j.cfg.mc.enabled ? Stage::McReady : Stage::Writing);
                    if (j.stage == Stage::McReady) {
                        Clock tp;
                        // MC on the whole space (no CA): no host preparation beyond the descriptor
                        j.prep = mc_prepare(j.cfg, nullptr, Cooling::Full);
                        j.prep_ms = tp.ms();
                    }
                    if (j.stage == Stage::Writing) write_async(j);
                    break;
                case Stage::Creating:
                    j.stage = Stage::Created;
                    break;
                case Stage::PostCa:
                    // the MC takes its domain from the simulation's final state on the device: the simulation is
                    // freed once the MC stage is done
                    if (!j.cfg.mc.enabled) release_ca(j);
                    j.stage = j.cfg.mc.enabled ? Stage::McReady : Stage::Writing;
                    if (j.stage == Stage::Writing) write_async(j);
                    break;
                case Stage::Mc:
                    mc_busy_[static_cast<size_t>(j.mc_stream)] = 0;
                    used_ -= j.fp_mc;
                    release_ca(j);
                    j.stage = Stage::Done;
                    j.latency_ms = since_arrival(j);
                    break;
                case Stage::Writing:
                    j.stage = Stage::Done;
                    j.latency_ms = since_arrival(j);
                    break;
                default:
                    break;
            }
        }
    }

    // on the GPU thread, between chunks
    void release_ca(Job& j) {
        if (!j.ca) return;
        Clock t;
        vca_ca_destroy(j.ca);
        release_ms_ += t.ms();
        j.ca = nullptr;
        used_ -= j.fp_ca;
    }

    // The job's error as a child `engine --config` reports it (its last "Błąd..." line and its exit code, which
    // engine/src/app/Batch.cpp:100-114 copies into batch_summary.csv).
    void fail(Job& j, const std::exception& e) {
        if (dynamic_cast<const ConfigError*>(&e)) {
            j.exit_code = 1;
            fail(j, std::string("Błąd konfiguracji: ") + e.what());
        } else if (dynamic_cast<const CliError*>(&e)) {
            j.exit_code = 2;
            fail(j, std::string("Błąd argumentów: ") + e.what());
        } else {
            j.exit_code = 3;
            fail(j, std::string("Błąd: ") + e.what());
        }
    }
    void fail(Job& j, const std::string& what) {
        if (j.exit_code == 0) j.exit_code = 3;
        j.error = what;
        if (j.ca) {
            if (j.slot >= 0) vca_batch_remove(batch_, j.ca);
            vca_ca_destroy(j.ca);
            j.ca = nullptr;
            j.slot = -1;
            used_ -= j.fp_ca;
        }
        j.stage = Stage::Failed;
        j.latency_ms = since_arrival(j);
    }

    int admitted() const {
        int n = 0;
        for (auto& j : jobs_)
            n += j->stage == Stage::Creating || j->stage == Stage::Created || j->stage == Stage::Ca ||
                 j->stage == Stage::PostCa || j->stage == Stage::McReady || j->stage == Stage::Mc;
        return n;
    }

    // arrival order; a job waits while its footprint does not fit next to the running ones (or while
    // --max-jobs jobs are between admission and their result files).  Simulations are created on worker threads
    // (allocation, rock columns, seeds on a stream of their own) and join the batch here, between chunks.
    bool admit() {
        bool any = false;
        for (auto& jp : jobs_) {
            Job& j = *jp;
            if (j.stage != Stage::Created) continue;
            Clock t;
            int32_t slot = -1;
            if (vca_batch_add(batch_, j.ca, &slot) != VCA_OK) break;  // no slot / tile range: retry later
            create_ms_ += t.ms();
            j.slot = slot;
            j.stage = Stage::Ca;
            j.ca_clock.reset();
            any = true;
        }
        int free_workers = static_cast<int>(workers_) - in_flight();
        for (auto& jp : jobs_) {
            Job& j = *jp;
            if (j.stage != Stage::Ready || j.work.valid()) continue;
            if (opt_.max_jobs > 0 && admitted() >= opt_.max_jobs) break;
            if (used_ + j.fp_ca > budget_ && used_ > 0) break;
            if (free_workers <= 0) break;
            try {
                if (opt_.autotune_steps > 0) tuned(j);
            } catch (const std::exception& e) {
                fail(j, e);
                continue;
            }
            used_ += j.fp_ca;
            j.stage = Stage::Creating;
            --free_workers;
            any = true;
            j.work = std::async(std::launch::async, [this, &j] {
                Clock t;
                j.ca = create_ca(j.cfg, j.in, gpu_);
                j.timing.add("allocate", t.ms());
                j.timing.add("terrain", j.in.terrain_ms);
                j.timing.add("vents", j.in.vents_ms);
                j.in = CaInputs{};  // host terrain no longer needed
            });
        }
        return any;
    }

    // --autotune: one probe per class of runs (dims, neighbourhood, mode, boundary), reused by the others
    void tuned(Job& j) {
        const Config& c = j.cfg;
        std::ostringstream key;
        key << c.space.nx << 'x' << c.space.ny << 'x' << c.space.nz << '/' << static_cast<int>(c.neighborhood) << '/'
            << c.ca.classic << '/' << c.periodic;
        auto it = tune_cache_.find(key.str());
        if (it == tune_cache_.end()) {
            autotune_ca(j.cfg, j.in, gpu_, opt_.autotune_steps);
            j.timing.add("autotune", j.in.tune_ms);
            tune_cache_[key.str()] = j.in.tune;
        } else {
            j.in.tuned = true;
            j.in.tune = it->second;
        }
        j.timing.set_info("autotune", autotune_json(j.in));
    }

    // one chunk of batched CA steps, then the stop rules of every member in one read
    bool step_ca() {
        int running = 0;
        for (auto& j : jobs_) running += j->stage == Stage::Ca;
        if (running == 0) return false;
        Clock te;
        vca_check(vca_batch_enqueue(batch_, kChunk), "CA");
        enqueue_ms_ += te.ms();
        te.reset();
        std::vector<int32_t> st(2 * kSlots);
        vca_check(vca_batch_status(batch_, st.data(), kSlots), "CA");
        status_ms_ += te.ms();
        for (auto& jp : jobs_) {
            Job& j = *jp;
            if (j.stage != Stage::Ca || st[2 * static_cast<size_t>(j.slot) + 1] == VCA_RUNNING) continue;
            try {
                j.timing.add("caSim", j.ca_clock.ms());
                Clock t;
                vca_check(vca_batch_remove(batch_, j.ca), "CA");
                j.slot = -1;
                collect_ms_ += t.ms();
                // the final state is read on the simulation's own stream while the batch goes on; the GPU thread
                // frees the simulation when this stage is done
                j.stage = Stage::PostCa;
                j.work = std::async(std::launch::async, [&j] {
                    Clock ts;
                    ca_collect(j.ca, j.cfg, j.s, ca_state_needed(j.cfg));
                    save_ca(j.cfg, j.s, j.out.string(), j.timing);
                    j.timing.add("save", ts.ms());
                    ts.reset();
                    ca_flow_sums(j.ca, j.s);
                    j.stats = ca_stats(j.cfg, j.s);
                    j.timing.add("caStats", ts.ms());
                    j.stats["parallel"] = parallel_json(j.cfg);
                    j.s.result = CaState{};  // written; the MC reads the state on the device
                    if (j.cfg.mc.enabled) j.prep = mc_prepare_from_ca(j.cfg, j.ca);
                });
            } catch (const std::exception& e) {
                fail(j, e);
            }
        }
        return true;
    }

    // MC stages run whole on worker threads, each on one of kMcStreams contexts of its own (so their small
    // colour launches overlap each other and the batched CA steps); then the job's files are written there.
    bool step_mc() {
        bool any = false;
        for (auto& jp : jobs_) {
            Job& j = *jp;
            if (j.stage != Stage::McReady || j.work.valid()) continue;
            if (used_ + j.fp_mc > budget_ && used_ > 0) continue;
            int k = -1;
            for (int c = 0; c < kMcStreams; ++c)
                if (!mc_busy_[static_cast<size_t>(c)]) {
                    k = c;
                    break;
                }
            if (k < 0) break;
            if (!mc_gpus_[static_cast<size_t>(k)]) mc_gpus_[static_cast<size_t>(k)] = std::make_unique<Gpu>();
            mc_busy_[static_cast<size_t>(k)] = 1;
            j.mc_stream = k;
            used_ += j.fp_mc;
            if (!j.s.ca_ran) j.stats["parallel"] = parallel_json(j.cfg);
            j.stage = Stage::Mc;
            any = true;
            Gpu& g = *mc_gpus_[static_cast<size_t>(k)];
            j.work = std::async(std::launch::async, [this, &j, &g] {
                Clock tc;
                {
                    McRun run(j.cfg, std::move(j.prep), g);
                    run.add_init_ms(j.prep_ms);
                    j.prep = McPrep{};
                    while (!run.done()) run.advance(run.next_stats_point());
                    j.s.mc = run.finish(j.timing);
                }
                j.s.mc_ran = true;
                j.stats["mc"] = j.s.mc.stats;
                write_files(j);
            });
        }
        return any;
    }

    // host stage: MC files, stats.json, times.json
    void write_files(Job& j) {
        const fs::path out = j.out;
        j.s.result = CaState{};  // the CA state is written and no longer needed
        if (j.s.mc_ran) save_mc(j.cfg, j.s.mc, out.string(), j.timing);
        Clock tw;
        write_json((out / "stats.json").string(), j.stats);
        j.timing.add_file((out / "stats.json").string(), tw.ms());
        j.timing.add(j.s.mc_ran ? "mcSave" : "save", tw.ms());
        j.timing.add("total", since_arrival(j));
        run_info(j.timing, j.cfg, j.s, gpu_);
        j.timing.set_info("pipeline", "gpu-batch");
        j.timing.record_files();
        write_json((out / "times.json").string(), j.timing.to_json());
        j.s.mc.crystals.clear();
    }
    void write_async(Job& j) {
        j.stage = Stage::Writing;
        j.work = std::async(std::launch::async, [this, &j] { write_files(j); });
    }

    void report(const Job& j) {
        ++reported_;
        std::cout << '[' << reported_ << '/' << jobs_.size() << "] " << j.name << ' '
                  << (j.stage == Stage::Done ? "ok" : "BŁĄD  " + j.error) << " (" << static_cast<long long>(j.latency_ms)
                  << " ms)\n"
                  << std::flush;
    }

    static constexpr int kChunk = 32;
    static constexpr int kMcStreams = 4;  // the device's hardware queues
    std::unique_ptr<Gpu> mc_gpus_[kMcStreams];
    int mc_busy_[kMcStreams] = {0, 0, 0, 0};
    const CliOptions& opt_;
    std::vector<std::unique_ptr<Job>>& jobs_;
    Gpu gpu_;
    vca_batch* batch_ = nullptr;
    uint64_t budget_ = 0, used_ = 0;
    unsigned workers_ = 4;
    size_t reported_ = 0;
    double makespan_s_ = 0.0;
    Clock start_;
    double phase_ms_[4] = {0.0, 0.0, 0.0, 0.0};
    std::map<std::string, vca_tune_result> tune_cache_;
    double collect_ms_ = 0.0;
    double spawn_ms_ = 0.0, release_ms_ = 0.0, create_ms_ = 0.0, enqueue_ms_ = 0.0, status_ms_ = 0.0;
};

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t k = static_cast<size_t>(std::lround(p * static_cast<double>(v.size() - 1)));
    return v[std::min(k, v.size() - 1)];
}

// Throughput and latency of a finished pipeline run (serve_summary.json / pipeline_summary.json).
Json pipeline_summary(const std::vector<std::unique_ptr<Job>>& jobs, double makespan, std::vector<double>& lat) {
    double cell_updates = 0.0, spin_attempts = 0.0;
    size_t ok = 0;
    Json list = Json::array();
    for (const auto& j : jobs) {
        Json r = Json::object();
        r["name"] = j->name;
        r["at"] = Json(j->at);
        r["latencyMs"] = Json(j->latency_ms);
        r["status"] = j->stage == Stage::Done ? "ok" : "error";
        if (!j->error.empty()) r["error"] = j->error;
        if (j->stage == Stage::Done) {
            ++ok;
            lat.push_back(j->latency_ms);
            if (j->s.ca_ran) cell_updates += static_cast<double>(j->cfg.space.cells()) * j->s.steps;
            if (j->s.mc_ran) spin_attempts += static_cast<double>(j->s.mc.domain_cells) * j->s.mc.mcs;
        }
        list.push_back(r);
    }
    Json sum = Json::object();
    sum["jobs"] = Json(static_cast<unsigned long long>(jobs.size()));
    sum["ok"] = Json(static_cast<unsigned long long>(ok));
    sum["makespanS"] = Json(makespan);
    sum["jobsPerS"] = Json(makespan > 0 ? static_cast<double>(ok) / makespan : 0.0);
    sum["p50LatencyMs"] = Json(percentile(lat, 0.50));
    sum["p99LatencyMs"] = Json(percentile(lat, 0.99));
    sum["cellUpdates"] = Json(cell_updates);
    sum["cellUpdatesPerS"] = Json(makespan > 0 ? cell_updates / makespan : 0.0);
    sum["spinAttemptsPerS"] = Json(makespan > 0 ? spin_attempts / makespan : 0.0);
    sum["jobsList"] = list;
    return sum;
}

}  // namespace

int run_batched(const CliOptions& opt) {
    fs::path out;
    const std::vector<BatchItem> items = batch_items(opt, out);
    fs::create_directories(out);
    std::vector<std::unique_ptr<Job>> jobs;
    for (const BatchItem& it : items) {
        auto j = std::make_unique<Job>();
        j->name = it.name;
        j->config = it.config;
        j->out = it.out;
        jobs.push_back(std::move(j));
    }
    std::cout << "=== Seria (--batch --gpu-batch): " << jobs.size() << " configów, jeden proces, wspólne kroki GPU -> "
              << out.generic_string() << " ===\n"
              << std::flush;
    double makespan = 0.0;
    Json phases;
    {
        Pipeline p(opt, jobs);
        p.run();
        makespan = p.makespan_s();
        phases = p.phases();
    }
    std::vector<BatchOutcome> outcomes;
    size_t ok = 0;
    for (const auto& j : jobs) {
        BatchOutcome o;
        o.item = BatchItem{j->name, j->config, j->out};
        o.exit_code = j->stage == Stage::Done ? 0 : j->exit_code;
        o.error = j->error;
        o.wall_ms = j->latency_ms;
        ok += o.ok();
        outcomes.push_back(o);
    }
    const fs::path summary = out / "batch_summary.csv";
    write_batch_summary(summary, outcomes);
    std::vector<double> lat;
    Json ps = pipeline_summary(jobs, makespan, lat);
    ps["maxJobs"] = opt.max_jobs;
    ps["gpuThread"] = phases;
    ps["graphSteps"] = opt.steps_per_graph;
    write_json((out / "pipeline_summary.json").string(), ps);
    std::cout << "\nRazem: " << outcomes.size() << " przebiegów: " << ok << " ok, " << (outcomes.size() - ok)
              << " z błędem, " << makespan << " s\nPodsumowanie   " << summary.generic_string() << '\n';
    return ok == outcomes.size() ? 0 : 5;
}

int run_serve(const CliOptions& opt) {
    std::ifstream in(opt.serve_path);
    if (!in) throw CliError("--serve: nie można otworzyć listy zadań: " + opt.serve_path);
    const fs::path base = fs::path(opt.serve_path).parent_path();
    const fs::path out = opt.out_dir.empty() ? fs::path("results") / fs::path(opt.serve_path).stem() : fs::path(opt.out_dir);
    fs::create_directories(out);
    std::vector<std::unique_ptr<Job>> jobs;
    std::string line;
    while (std::getline(in, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        const Json e = Json::parse(line);
        auto j = std::make_unique<Job>();
        j->at = e.at("at").get_double();
        fs::path cfg = e.at("config").get_string();
        if (cfg.is_relative()) cfg = base / cfg;
        j->config = cfg;
        j->name = e.contains("name") ? e.at("name").get_string() : cfg.stem().string() + "_" + std::to_string(jobs.size());
        j->out = out / j->name;
        jobs.push_back(std::move(j));
    }
    std::stable_sort(jobs.begin(), jobs.end(), [](const auto& a, const auto& b) { return a->at < b->at; });
    std::cout << "=== Serwowanie (--serve): " << jobs.size() << " zadań ze strumienia " << opt.serve_path << " ===\n"
              << std::flush;
    double makespan = 0.0;
    Json phases;
    {
        Pipeline p(opt, jobs);
        p.run();
        makespan = p.makespan_s();
        phases = p.phases();
    }
    std::vector<double> lat;
    Json sum = pipeline_summary(jobs, makespan, lat);
    const unsigned long long ok = sum.at("ok").get_uint();
    sum["gpuThread"] = phases;
    write_json((out / "serve_summary.json").string(), sum);
    std::cout << "\nRazem: " << ok << "/" << jobs.size() << " zadań ok, " << makespan << " s, p50 "
              << percentile(lat, 0.50) << " ms, p99 " << percentile(lat, 0.99) << " ms\nPodsumowanie   "
              << (out / "serve_summary.json").generic_string() << '\n';
    return ok == jobs.size() ? 0 : 5;
}

}  // namespace eng
