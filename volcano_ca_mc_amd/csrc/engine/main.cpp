// engine — the drop-in command line of the GPU engine (engine/src/main.cpp:20-56, engine/src/app/Cli.cpp:40-193).
// Exit codes: 0 ok, 1 configuration error, 2 argument error, 3 other error, 5 batch with failed items.
#include <iostream>

#include "simulation.hpp"

namespace eng {

std::string usage() {
    return "engine: symulacja erupcji wulkanu na GPU (MI355X): spływ lawy jako automat komórkowy i krystalizacja\n"
           "zastygłej lawy metodą Monte Carlo\n"
           "\n"
           "Użycie:\n"
           "  engine --config <plik.json> --out <folder> [--set klucz=wartość]... [--progress]\n"
           "  engine --mc-from <folder_wyniku_CA> --config <mc.json> --out <folder> [--set klucz=wartość]...\n"
           "  engine --batch <folder_z_configami> --out <folder_wyników> [--set klucz=wartość]... [--progress]\n"
           "  engine --batch <folder_z_configami> --gpu-batch --out <folder_wyników> [--set klucz=wartość]...\n"
           "  engine --serve <zadania.jsonl> --out <folder_wyników> [--set klucz=wartość]...\n"
           "  engine --validate [--config <plik.json>] [--set klucz=wartość]... [--mc-from <folder>]\n"
           "  engine --print-defaults\n"
           "  engine --help\n"
           "\n"
           "Opcje:\n"
           "  --config <plik>    konfiguracja JSON; brakujące pola mają wartości domyślne\n"
           "  --out <folder>     folder wyników (domyślnie results/<nazwa configu>)\n"
           "  --set k=w          nadpisanie parametru ścieżką z kropkami, np. --set space.nx=1000\n"
           "  --grid <g>         skrót dla --set grid=<g> (na GPU bez znaczenia dla wyniku)\n"
           "  --threads N        skrót dla --set parallel.threads=N (akceptowane dla zgodności)\n"
           "  --mc-from <f>      samo Monte Carlo na wyniku CA z folderu f (ca_result.bin lub .txt)\n"
           "  --batch <f>        seria: wszystkie configi *.json z folderu f, każdy w osobnym procesie\n"
           "  --gpu-batch        z --batch: cała seria w jednym procesie, kroki CA wszystkich przebiegów\n"
           "                     wspólnymi wywołaniami GPU (wyniki identyczne jak przebiegi pojedyncze)\n"
           "  --serve <plik>     strumień zadań: linie {\"at\": s, \"config\": ścieżka[, \"name\": n]}, przyjmowane\n"
           "                     od chwili at, wspólne kroki GPU; serve_summary.json z opóźnieniami p50/p99\n"
           "  --graph-steps N    długość grafu HIP kroków wsadowych (domyślnie 16, 0 = bez grafów)\n"
           "  --max-jobs N       najwyżej N zadań naraz w potoku wsadowym (1 = po jednym; domyślnie bez limitu)\n"
           "  --timing-sample N  jak --timing-detail, ale czasy kerneli z co N-tego kroku CA i przemiatania MC\n"
           "  --digests          zamiast plików wyników (ca_result.*, mc_result.*, migawki) zapisuje <plik>.sha256:\n"
           "                     skrót SHA-256 i rozmiar tych samych bajtów (porównania przebiegów większych niż dysk)\n"
           "  --autotune N       przed przebiegiem CA: N kroków próbnych na każdy kształt kafla i tryb frontu,\n"
           "                     przebieg z najszybszym (wyniki bez zmian; wybór w times.json)\n"
           "  --no-fast-forward  każdy krok CA uruchamiany na GPU, także po ustaniu przepływów (gdy lawa już tylko\n"
           "                     stygnie); wyniki bez zmian\n"
           "  --progress         linie PROGRESS ca/mc/save/batch na stdout (dla GUI)\n"
           "  --validate         tylko walidacja configu; JSON {\"ok\", \"errors\", \"warnings\"} na stdout\n"
           "  --print-defaults   pełny domyślny config JSON\n"
           "\n"
           "Wyniki w folderze --out: config_used.json, ca_result.txt|bin, times.json, stats.json,\n"
           "przy mc.enabled także mc_result.txt|bin i mc_stats.csv.\n";
}

CliOptions parse_cli(int argc, const char* const* argv) {
    CliOptions opt;
    auto value = [&](int& i, const std::string& name) {
        if (i + 1 >= argc) throw CliError("opcja " + name + " wymaga wartości");
        return std::string(argv[++i]);
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-h") opt.help = true;
        else if (a == "--config") opt.config_path = value(i, a);
        else if (a == "--out") opt.out_dir = value(i, a);
        else if (a == "--mc-from") opt.mc_from = value(i, a);
        else if (a == "--batch") opt.batch_dir = value(i, a);
        else if (a == "--gpu-batch") opt.gpu_batch = true;
        else if (a == "--serve") opt.serve_path = value(i, a);
        else if (a == "--max-jobs") {
            const std::string v = value(i, a);
            try {
                opt.max_jobs = std::stoi(v);
            } catch (...) {
                throw CliError("--max-jobs wymaga liczby, podano: " + v);
            }
            if (opt.max_jobs < 0) throw CliError("--max-jobs nie może być ujemne");
        } else if (a == "--autotune") {
            const std::string v = value(i, a);
            try {
                opt.autotune_steps = std::stoi(v);
            } catch (...) {
                throw CliError("--autotune wymaga liczby, podano: " + v);
            }
            if (opt.autotune_steps < 1 || opt.autotune_steps > 65535) throw CliError("--autotune poza zakresem 1..65535");
        } else if (a == "--graph-steps") {
            const std::string v = value(i, a);
            try {
                opt.steps_per_graph = std::stoi(v);
            } catch (...) {
                throw CliError("--graph-steps wymaga liczby, podano: " + v);
            }
            if (opt.steps_per_graph < 0 || opt.steps_per_graph > 4096) throw CliError("--graph-steps poza zakresem 0..4096");
        }
        else if (a == "--set") {
            const std::string kv = value(i, a);
            if (kv.find('=') == std::string::npos || kv.front() == '=')
                throw CliError("--set wymaga postaci klucz=wartość (np. space.nx=100), podano: " + kv);
            opt.overrides.push_back(kv);
        } else if (a == "--grid") opt.overrides.push_back("grid=" + value(i, a));
        else if (a == "--threads") {
            const std::string n = value(i, a);
            if (n.empty() || n.find_first_not_of("0123456789") != std::string::npos)
                throw CliError("opcja --threads wymaga liczby całkowitej >= 0, podano: " + n);
            opt.overrides.push_back("parallel.threads=" + n);
        } else if (a == "--thread-stats") opt.thread_stats = true;
        else if (a == "--high-priority") opt.overrides.push_back("parallel.priority=high");
        else if (a == "--benchmark") opt.benchmark = true;
        else if (a == "--timing-detail") opt.timing_detail = true;
        else if (a == "--digests") opt.digests = true;
        else if (a == "--no-fast-forward") opt.fast_forward = false;
        else if (a == "--timing-sample") {
            const std::string v = value(i, a);
            try {
                opt.timing_every = std::stoi(v);
            } catch (...) {
                throw CliError("--timing-sample wymaga liczby, podano: " + v);
            }
            if (opt.timing_every < 1) throw CliError("--timing-sample musi być >= 1");
            opt.timing_detail = true;
        }
        else if (a == "--print-defaults") opt.print_defaults = true;
        else if (a == "--validate") opt.validate = true;
        else if (a == "--progress") opt.progress = true;
        else throw CliError("nieznany argument: " + a + " (zobacz --help)");
    }
    if (opt.print_defaults && (opt.validate || opt.benchmark || opt.progress || opt.timing_detail || opt.thread_stats ||
                               !opt.config_path.empty() || !opt.out_dir.empty() || !opt.mc_from.empty() ||
                               !opt.batch_dir.empty() || !opt.overrides.empty()))
        throw CliError("--print-defaults nie przyjmuje innych opcji");
    if (!opt.batch_dir.empty() && (!opt.config_path.empty() || !opt.mc_from.empty() || opt.benchmark || opt.validate))
        throw CliError("--batch nie łączy się z --config, --mc-from, --benchmark ani --validate");
    if (opt.gpu_batch && opt.batch_dir.empty()) throw CliError("--gpu-batch działa tylko z --batch");
    if (!opt.serve_path.empty() && (!opt.batch_dir.empty() || !opt.config_path.empty() || !opt.mc_from.empty() ||
                                    opt.validate))
        throw CliError("--serve nie łączy się z --batch, --config, --mc-from ani --validate");
    if (opt.validate && (opt.benchmark || opt.progress || opt.timing_detail || opt.thread_stats))
        throw CliError("--validate nie łączy się z --benchmark, --progress, --timing-detail ani --thread-stats");
    if (opt.benchmark)
        throw CliError("--benchmark (porównanie vector vs macierz CPU) nie dotyczy silnika GPU; wydajność mierzy bench.py");
    return opt;
}

}  // namespace eng

int main(int argc, char** argv) {
    using namespace eng;
    const Clock since_start;
    try {
        const CliOptions opt = parse_cli(argc, argv);
        if (opt.help) {
            std::cout << usage();
            return 0;
        }
        if (opt.print_defaults) {
            std::cout << Config::defaults().dump(2) << '\n';
            return 0;
        }
        if (opt.validate) return run_validate(opt, std::cout);
        if (opt.digests) set_result_digests(true);
        set_fast_forward(opt.fast_forward);
        if (!opt.batch_dir.empty()) return run_batch(opt);
        if (!opt.serve_path.empty()) return run_serve(opt);
        return run_from_cli(opt, since_start);
    } catch (const CliError& e) {
        std::cerr << "Błąd argumentów: " << e.what() << "\n\n" << usage();
        return 2;
    } catch (const ConfigError& e) {
        std::cerr << "Błąd konfiguracji: " << e.what() << '\n';
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "Błąd: " << e.what() << '\n';
        return 3;
    }
}
