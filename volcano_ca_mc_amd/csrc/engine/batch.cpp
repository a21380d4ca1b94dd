// engine --batch <folder> --out <folder>: every *.json config of a folder (name order; series.json and
// _*.json skipped) runs in its own child engine process into <out>/<name>/ with engine.log;
// batch_summary.csv is rewritten after every item; exit code 5 when any item failed
// (engine/src/app/Batch.cpp:147-418).  The parent never touches the GPU; each child owns it in turn.
// With --gpu-batch the whole series runs in this process through the batched pipeline (pipeline.cpp).
#include <spawn.h>
#include <sys/wait.h>
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>

#include "simulation.hpp"

extern char** environ;

namespace eng {

namespace fs = std::filesystem;

namespace {

using Item = BatchItem;
using Outcome = BatchOutcome;

const char* const kSections[] = {"readConfig", "readInput", "allocate", "terrain", "vents", "caSim", "snapshots",
                                 "save", "mcInit", "mcSim", "mcStats", "mcSave", "total"};

Json read_json_file(const fs::path& p) {
    std::ifstream in(p, std::ios::binary);
    if (!in) return Json();
    std::stringstream ss;
    ss << in.rdbuf();
    try {
        return Json::parse(ss.str());
    } catch (...) {
        return Json();
    }
}
std::string field(const Json& j, const std::string& key) {
    if (!j.is_object() || !j.contains(key)) return "";
    const Json& v = j.at(key);
    if (v.is_string()) return v.get_string();
    if (v.is_number() || v.is_bool()) return v.dump();
    return "";
}
std::string csv_cell(const std::string& s) {
    if (s.find_first_of(",\"\n\r") == std::string::npos) return s;
    std::string out = "\"";
    for (char c : s) {
        if (c == '"') out += '"';
        out += c;
    }
    return out + '"';
}
std::string now_text() {
    char buf[32];
    const std::time_t now = std::time(nullptr);
    std::tm local{};
    localtime_r(&now, &local);
    std::strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S", &local);
    return buf;
}
std::string last_error_line(const fs::path& log) {
    std::ifstream in(log);
    std::string line, last;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.rfind("Błąd", 0) == 0) last = line;
    }
    return last;
}

std::vector<std::string> summary_columns() {
    std::vector<std::string> cols = {"name", "status", "exitCode", "error", "mode", "nx", "ny", "nz", "cells", "grid",
                                     "neighborhood", "boundary", "seeds", "steps", "stopReason", "mcs", "Q", "domainCells"};
    for (const char* s : kSections) cols.emplace_back(s);
    for (const char* c : {"wallMs", "peakRssBytes", "peakPrivateBytes", "modelPeakBytes", "caResultTxtBytes",
                          "caResultBinBytes", "mcResultTxtBytes", "mcResultBinBytes", "outputBytes", "compiler", "build",
                          "config", "outDir"})
        cols.emplace_back(c);
    return cols;
}

std::vector<std::string> summary_row(const Outcome& o) {
    const Json times = o.ok() ? read_json_file(o.item.out / "times.json") : Json();
    const Json cfg = o.ok() ? read_json_file(o.item.out / "config_used.json") : Json();
    const std::string mode = field(times, "mode");
    std::vector<std::string> row = {o.item.name, o.ok() ? "ok" : "error", std::to_string(o.exit_code), o.error, mode};
    for (size_t a = 0; a < 3; ++a) {
        const bool has = times.is_object() && times.contains("space") && times.at("space").size() == 3;
        row.push_back(has ? times.at("space")[a].dump() : "");
    }
    row.push_back(field(times, "cells"));
    row.push_back(field(times, "grid"));
    row.push_back(field(times, "neighborhood"));
    row.push_back(field(times, "boundary"));
    const bool ca_ran = mode == "lava" || mode == "classic";
    row.push_back(ca_ran && cfg.is_object() && cfg.contains("vents") ? field(cfg.at("vents"), "count") : "");
    row.push_back(field(times, "steps"));
    row.push_back(field(times, "stopReason"));
    row.push_back(field(times, "mcs"));
    row.push_back(field(times, "Q"));
    row.push_back(field(times, "domainCells"));
    for (const char* s : kSections) row.push_back(field(times, s));
    char w[64];
    std::snprintf(w, sizeof(w), "%.3f", o.wall_ms);
    row.push_back(w);
    row.push_back("");  // peakRssBytes: not modelled by the GPU engine
    row.push_back("");
    row.push_back("");
    uint64_t total = 0;
    bool any = false;
    for (const char* f : {"ca_result.txt", "ca_result.bin", "mc_result.txt", "mc_result.bin"}) {
        std::string v;
        if (times.is_object() && times.contains("files") && times.at("files").contains(f))
            v = field(times.at("files").at(f), "bytes");
        row.push_back(v);
    }
    if (times.is_object() && times.contains("files"))
        for (const auto& kv : times.at("files").items())
            if (kv.second.contains("bytes")) {
                total += kv.second.at("bytes").get_uint();
                any = true;
            }
    row.push_back(any ? std::to_string(total) : "");
    row.push_back(field(times, "compiler"));
    row.push_back(field(times, "build"));
    row.push_back(o.item.config.generic_string());
    row.push_back(o.item.out.generic_string());
    return row;
}

}  // namespace

void write_batch_summary(const fs::path& csv, const std::vector<BatchOutcome>& outcomes) {
    std::ofstream f(csv, std::ios::trunc);
    if (!f) throw std::runtime_error("Nie można zapisać podsumowania serii: " + csv.string());
    const auto cols = summary_columns();
    for (size_t i = 0; i < cols.size(); ++i) f << (i ? "," : "") << cols[i];
    f << '\n';
    for (const Outcome& o : outcomes) {
        const auto row = summary_row(o);
        for (size_t i = 0; i < row.size(); ++i) f << (i ? "," : "") << csv_cell(row[i]);
        f << '\n';
    }
}

namespace {

// Child engine process with stdout / stderr appended to log; returns its exit code (-1: did not start).
int run_child(const fs::path& exe, const std::vector<std::string>& args, const fs::path& log, std::string& err) {
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_addopen(&fa, 1, log.c_str(), O_WRONLY | O_APPEND | O_CREAT, 0644);
    posix_spawn_file_actions_adddup2(&fa, 1, 2);
    std::vector<std::string> all = {exe.string()};
    all.insert(all.end(), args.begin(), args.end());
    std::vector<char*> argv;
    for (auto& a : all) argv.push_back(const_cast<char*>(a.c_str()));
    argv.push_back(nullptr);
    pid_t pid = 0;
    const int rc = posix_spawn(&pid, exe.c_str(), &fa, nullptr, argv.data(), environ);
    posix_spawn_file_actions_destroy(&fa);
    if (rc != 0) {
        err = std::string("przebieg nie wystartował: ") + std::strerror(rc);
        return -1;
    }
    int status = 0;
    while (waitpid(pid, &status, 0) < 0) {
    }
    if (WIFEXITED(status)) return WEXITSTATUS(status);
    err = "silnik zakończony sygnałem " + std::to_string(WTERMSIG(status));
    return 128 + WTERMSIG(status);
}

}  // namespace

std::vector<BatchItem> batch_items(const CliOptions& opt, fs::path& out) {
    fs::path folder(opt.batch_dir);
    if (!fs::is_directory(folder)) throw CliError("--batch: folder z configami nie istnieje: " + opt.batch_dir);
    if (folder.filename().empty()) folder = folder.parent_path();
    out = opt.out_dir.empty() ? fs::path("results") / folder.filename() : fs::path(opt.out_dir);
    std::vector<Item> items;
    for (const auto& e : fs::directory_iterator(folder)) {
        if (!e.is_regular_file()) continue;
        const fs::path& p = e.path();
        std::string ext = p.extension().string();
        std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
        const std::string stem = p.stem().string();
        if (ext != ".json" || stem.empty() || stem[0] == '_' || p.filename() == "series.json") continue;
        items.push_back(Item{stem, p, out / p.stem()});
    }
    std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.name < b.name; });
    if (items.empty())
        throw ConfigError("W folderze " + opt.batch_dir + " nie ma plików konfiguracyjnych *.json (series.json i pliki _*.json są pomijane)");
    return items;
}

int run_batch(const CliOptions& opt) {
    if (opt.gpu_batch) return run_batched(opt);
    fs::path out;
    const std::vector<Item> items = batch_items(opt, out);
    const fs::path folder = fs::path(opt.batch_dir).filename().empty() ? fs::path(opt.batch_dir).parent_path()
                                                                         : fs::path(opt.batch_dir);
    const fs::path exe = fs::read_symlink("/proc/self/exe");
    fs::create_directories(out);
    std::vector<std::string> extra;
    for (const std::string& o : opt.overrides) {
        extra.push_back("--set");
        extra.push_back(o);
    }
    if (opt.timing_detail) extra.push_back("--timing-detail");
    if (opt.thread_stats) extra.push_back("--thread-stats");
    if (opt.digests) extra.push_back("--digests");
    Progress progress(opt.progress);
    const fs::path summary = out / "batch_summary.csv";
    std::cout << "=== Seria (--batch): " << items.size() << " configów z " << folder.generic_string() << " -> "
              << out.generic_string() << " ===\n"
              << std::flush;
    std::vector<Outcome> outcomes;
    const long long n = static_cast<long long>(items.size());
    progress.report("batch", 0, n, true);
    for (size_t i = 0; i < items.size(); ++i) {
        const Item& it = items[i];
        std::error_code ec;
        fs::create_directories(it.out, ec);
        const fs::path log = it.out / "engine.log";
        {
            std::ofstream head(log, std::ios::trunc);
            head << "# " << now_text() << " engine --batch: " << it.name << '\n';
            head << "# polecenie: " << exe.generic_string() << " --config " << it.config.generic_string() << " --out "
                 << it.out.generic_string();
            for (const auto& a : extra) head << ' ' << a;
            head << '\n';
        }
        std::vector<std::string> args = {"--config", it.config.string(), "--out", it.out.string()};
        args.insert(args.end(), extra.begin(), extra.end());
        Clock t;
        Outcome o;
        o.item = it;
        std::string err;
        o.exit_code = run_child(exe, args, log, err);
        o.wall_ms = t.ms();
        if (!err.empty()) {
            o.error = err;
        } else if (o.exit_code != 0) {
            o.error = last_error_line(log);
            if (o.error.empty()) o.error = "silnik zakończył się kodem " + std::to_string(o.exit_code);
        }
        {
            std::ofstream tail(log, std::ios::app);
            char b[64];
            std::snprintf(b, sizeof(b), "%.2f", o.wall_ms / 1000.0);
            tail << "# koniec: kod wyjścia " << o.exit_code << ", czas " << b << " s\n";
        }
        outcomes.push_back(o);
        write_batch_summary(summary, outcomes);
        std::cout << '[' << (i + 1) << '/' << n << "] " << it.name << ' ' << (o.ok() ? "ok" : "BŁĄD  " + o.error) << '\n'
                  << std::flush;
        progress.report("batch", static_cast<long long>(i + 1), n, true);
    }
    size_t ok = 0;
    for (const Outcome& o : outcomes) ok += o.ok() ? 1 : 0;
    std::cout << "\nRazem: " << outcomes.size() << " przebiegów: " << ok << " ok, " << (outcomes.size() - ok)
              << " z błędem\nPodsumowanie   " << summary.generic_string() << '\n';
    return ok == outcomes.size() ? 0 : 5;
}

}  // namespace eng
