#include "config.hpp"

#include <cmath>
#include <fstream>
#include <initializer_list>
#include <limits>
#include <sstream>

namespace eng {

namespace {

std::string q(const std::string& path) { return "„" + path + "”"; }

std::vector<std::string> split_path(const std::string& path) {
    std::vector<std::string> parts;
    std::stringstream ss(path);
    std::string part;
    while (std::getline(ss, part, '.')) parts.push_back(part);
    return parts;
}

const Json& field(const Json& j, const std::string& path) {
    const Json* cur = &j;
    for (const std::string& part : split_path(path)) {
        if (!cur->is_object() || !cur->contains(part)) throw ConfigError("Brak parametru " + q(path));
        cur = &cur->at(part);
    }
    return *cur;
}

long long get_integer(const Json& j, const std::string& path, long long lo, long long hi) {
    const Json& v = field(j, path);
    if (!v.is_number_integer() || (v.is_number_unsigned() && v.get_uint() > static_cast<uint64_t>(hi))) {
        if (v.is_number_unsigned())
            throw ConfigError("Parametr " + q(path) + " jest za duży (maks. " + std::to_string(hi) + "), podano: " + v.dump());
        throw ConfigError("Parametr " + q(path) + " musi być liczbą całkowitą, podano: " + v.dump());
    }
    const long long x = v.get_int();
    if (x < lo || x > hi)
        throw ConfigError("Parametr " + q(path) + " musi być z zakresu " + std::to_string(lo) + ".." + std::to_string(hi) +
                          ", podano: " + std::to_string(x));
    return x;
}
int get_int(const Json& j, const std::string& path, int lo, int hi) {
    return static_cast<int>(get_integer(j, path, lo, hi));
}
double get_number(const Json& j, const std::string& path, double lo, double hi) {
    const Json& v = field(j, path);
    if (!v.is_number()) throw ConfigError("Parametr " + q(path) + " musi być liczbą, podano: " + v.dump());
    const double x = v.get_double();
    if (!std::isfinite(x) || x < lo || x > hi) {
        std::ostringstream os;
        os << "Parametr " << q(path) << " musi być z zakresu " << lo << ".." << hi << ", podano: " << x;
        throw ConfigError(os.str());
    }
    return x;
}
bool get_bool(const Json& j, const std::string& path) {
    const Json& v = field(j, path);
    if (!v.is_bool()) throw ConfigError("Parametr " + q(path) + " musi mieć wartość true lub false, podano: " + v.dump());
    return v.get_bool();
}
template <class E>
E get_enum(const Json& j, const std::string& path, std::initializer_list<std::pair<const char*, E>> options) {
    const Json& v = field(j, path);
    std::string allowed;
    for (const auto& o : options) allowed += (allowed.empty() ? "" : ", ") + std::string(o.first);
    if (v.is_string())
        for (const auto& o : options)
            if (v.get_string() == o.first) return o.second;
    throw ConfigError("Parametr " + q(path) + " ma nieprawidłową wartość " + v.dump() + "; dozwolone: " + allowed);
}

std::vector<std::pair<int, int>> get_columns(const Json& j, const std::string& path, const Dims& d) {
    const Json& list = field(j, path);
    if (!list.is_array())
        throw ConfigError("Parametr " + q(path) + " musi być listą par [x, y], np. [[80, 75], [82, 70]], podano: " +
                          list.dump());
    std::vector<std::pair<int, int>> out;
    for (size_t i = 0; i < list.size(); ++i) {
        const Json& p = list[i];
        const std::string where = path + "[" + std::to_string(i) + "]";
        auto is_int = [](const Json& v) {
            return v.is_number_integer() && (!v.is_number_unsigned() || v.get_uint() <= 100000ULL);
        };
        if (!p.is_array() || p.size() != 2 || !is_int(p[0]) || !is_int(p[1]))
            throw ConfigError("Parametr " + q(where) + " musi być parą liczb całkowitych [x, y], podano: " + p.dump());
        const long long x = p[0].get_int(), y = p[1].get_int();
        if (x < 0 || x >= d.nx || y < 0 || y >= d.ny)
            throw ConfigError("Parametr " + q(where) + " = " + p.dump() + " leży poza siatką: x musi być z zakresu 0.." +
                              std::to_string(d.nx - 1) + ", y z zakresu 0.." + std::to_string(d.ny - 1) +
                              " (space.nx = " + std::to_string(d.nx) + ", space.ny = " + std::to_string(d.ny) + ")");
        for (const auto& c : out)
            if (c.first == x && c.second == y)
                throw ConfigError("Parametr " + q(where) + " = " + p.dump() +
                                  " powtarza kolumnę z wcześniejszego elementu listy; w jednej kolumnie może być "
                                  "najwyżej jeden krater");
        out.emplace_back(static_cast<int>(x), static_cast<int>(y));
    }
    return out;
}

}  // namespace

const char* neighborhood_name(Neighborhood n) {
    switch (n) {
        case Neighborhood::VonNeumann: return "von_neumann";
        case Neighborhood::Moore: return "moore";
        case Neighborhood::Pentagonal: return "pentagonal";
        case Neighborhood::Hexagonal: return "hexagonal";
    }
    return "?";
}

// The configuration surface of the engine: the same keys and default values as the reference
// (engine/src/io/Config.cpp:153-176), in the same order (config_used.json keeps it).
Json Config::defaults() {
    static const char* kDefaults = R"({
      "space":        { "nx": 200, "ny": 1, "nz": 100 },
      "grid":         "vector",
      "boundary":     "absorbing",
      "neighborhood": "moore",
      "rngSeed":      42,
      "terrain": { "type": "cone", "peakHeight": 70, "baseRadius": 90,
                   "craterRadius": 6, "craterDepth": 4, "noise": 0.05, "file": null },
      "vents":   { "count": 1, "placement": "crater", "positions": [], "eruptionSteps": 300,
                   "eruptTemp": 1200 },
      "ca":      { "mode": "lava", "maxSteps": 5000, "coolingAmbient": 0.3,
                   "coolingLateral": 1.5, "coolingFall": 0.1, "solidusTemp": 800,
                   "ponding": true, "sweep": "front" },
      "mc":      { "enabled": false, "domain": "basalt", "Q": 50, "mcs": 100, "J": 1.0, "kT": 0.0,
                   "mobilityFromCooling": true, "mobilityMin": 0.05, "statsEvery": 10,
                   "snapshotsEvery": 0, "update": "random" },
      "output":  { "format": "txt", "snapshotsEvery": 0 },
      "parallel": { "threads": 1, "schedule": "static", "chunk": 0, "affinity": "none", "priority": "normal" },
      "mpi":     { "exchange": "always", "pack": "manual", "bcast": "json", "io": "gather" }
    })";
    return Json::parse(kDefaults);
}

void Config::merge(Json& base, const Json& patch, const std::string& prefix) {
    for (const auto& kv : patch.items()) {
        const std::string& key = kv.first;
        if (!key.empty() && key[0] == '_') continue;
        const std::string path = prefix.empty() ? key : prefix + "." + key;
        if (!base.contains(key)) {
            std::string allowed;
            for (const auto& b : base.items()) allowed += (allowed.empty() ? "" : ", ") + b.first;
            throw ConfigError("Nieznany parametr " + q(path) + "; dozwolone w tej sekcji: " + allowed);
        }
        Json& target = base[key];
        if (target.is_object()) {
            if (!kv.second.is_object())
                throw ConfigError("Parametr " + q(path) + " jest sekcją: oczekiwano obiektu { ... }, podano: " +
                                  kv.second.dump());
            merge(target, kv.second, path);
        } else {
            target = kv.second;
        }
    }
}

void Config::apply_override(Json& j, const std::string& key_value) {
    const size_t eq = key_value.find('=');
    if (eq == std::string::npos || eq == 0) throw ConfigError("Nadpisanie musi mieć postać klucz=wartość, podano: " + key_value);
    const std::string key = key_value.substr(0, eq);
    const std::string text = key_value.substr(eq + 1);
    Json value;
    try {
        value = Json::parse(text, false);
    } catch (const JsonError&) {
        value = Json(text);
    }
    const std::vector<std::string> parts = split_path(key);
    for (const std::string& p : parts)
        if (p.empty()) throw ConfigError("Nieprawidłowa ścieżka parametru w nadpisaniu: " + key);
    Json patch = value;
    for (auto it = parts.rbegin(); it != parts.rend(); ++it) {
        Json obj = Json::object();
        obj[*it] = patch;
        patch = obj;
    }
    merge(j, patch, "");
}

Config Config::load(const std::string& path, const std::vector<std::string>& overrides) {
    Json j = defaults();
    bool space_given = false;
    if (!path.empty()) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw ConfigError("Nie można otworzyć pliku konfiguracyjnego: " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        Json file;
        try {
            file = Json::parse(ss.str(), true);
        } catch (const JsonError& e) {
            throw ConfigError("Błąd składni JSON w pliku " + path + ": " + e.what());
        }
        if (!file.is_object()) throw ConfigError("Plik konfiguracyjny " + path + " musi zawierać obiekt JSON { ... }");
        merge(j, file, "");
        space_given = file.contains("space");
    }
    for (const std::string& o : overrides) {
        apply_override(j, o);
        space_given = space_given || o.rfind("space", 0) == 0;
    }
    Config c = from_json(j);
    c.source_path = path;
    c.space_from_user = space_given;
    return c;
}

// Validation with the reference's rules and ranges (engine/src/io/Config.cpp:268-437).
Config Config::from_json(const Json& j) {
    Config c;
    c.space.nx = get_int(j, "space.nx", 1, 100000);
    c.space.ny = get_int(j, "space.ny", 1, 100000);
    c.space.nz = get_int(j, "space.nz", 1, 100000);
    if (c.space.active_axes() < 2)
        throw ConfigError("Przestrzeń musi być 2D (dokładnie jeden wymiar równy 1) lub 3D; podano " +
                          std::to_string(c.space.nx) + " x " + std::to_string(c.space.ny) + " x " + std::to_string(c.space.nz));
    if (c.space.cells() > 2000000000ULL)
        throw ConfigError("Przestrzeń ma za dużo komórek (" + std::to_string(c.space.cells()) + "); maksimum to 2 000 000 000");
    c.grid_matrix = get_enum<bool>(j, "grid", {{"vector", false}, {"matrix", true}});
    c.periodic = get_enum<bool>(j, "boundary", {{"absorbing", false}, {"periodic", true}});
    c.neighborhood = get_enum<Neighborhood>(j, "neighborhood",
                                            {{"von_neumann", Neighborhood::VonNeumann},
                                             {"moore", Neighborhood::Moore},
                                             {"pentagonal", Neighborhood::Pentagonal},
                                             {"hexagonal", Neighborhood::Hexagonal}});
    c.seed = static_cast<uint64_t>(get_integer(j, "rngSeed", 0, std::numeric_limits<long long>::max()));

    c.mc.enabled = get_bool(j, "mc.enabled");
    c.mc.full = get_enum<bool>(j, "mc.domain", {{"basalt", false}, {"full", true}});
    c.mc.Q = get_int(j, "mc.Q", 2, 1000000000);
    c.mc.mcs = get_int(j, "mc.mcs", 0, 10000000);
    c.mc.J = get_number(j, "mc.J", 0.0, 1e6);
    if (c.mc.J <= 0.0)
        throw ConfigError("Parametr „mc.J” (energia granicy ziarna) musi być większy od 0, podano: " + field(j, "mc.J").dump());
    c.mc.kT = get_number(j, "mc.kT", 0.0, 1e6);
    c.mc.mobility_from_cooling = get_bool(j, "mc.mobilityFromCooling");
    c.mc.mobility_min = get_number(j, "mc.mobilityMin", 0.0, 1.0);
    c.mc.stats_every = get_int(j, "mc.statsEvery", 1, 1000000000);
    c.mc.snapshots_every = get_int(j, "mc.snapshotsEvery", 0, 1000000000);
    c.mc.checkerboard = get_enum<bool>(j, "mc.update", {{"random", false}, {"checkerboard", true}});
    const bool runs_ca = !c.mc.full_only();

    c.ca.classic = get_enum<bool>(j, "ca.mode", {{"lava", false}, {"classic", true}});
    c.ca.max_steps = get_int(j, "ca.maxSteps", 0, 65535);
    c.ca.cooling_ambient = get_number(j, "ca.coolingAmbient", 0.0, 1e6);
    c.ca.cooling_lateral = get_number(j, "ca.coolingLateral", 0.0, 1e6);
    c.ca.cooling_fall = get_number(j, "ca.coolingFall", 0.0, 1e6);
    c.ca.solidus = get_number(j, "ca.solidusTemp", -1e6, 1e6);
    c.ca.ponding = get_bool(j, "ca.ponding");
    c.ca.front = get_enum<bool>(j, "ca.sweep", {{"front", true}, {"full", false}});
    if (runs_ca && !c.ca.classic && c.space.is2d() && c.space.ny != 1)
        throw ConfigError("W trybie lava przestrzeń 2D to przekrój pionowy x–z: wymagane ny = 1 (podano nx = " +
                          std::to_string(c.space.nx) + ", ny = " + std::to_string(c.space.ny) + ", nz = " +
                          std::to_string(c.space.nz) + ")");
    if (c.neighborhood == Neighborhood::Hexagonal && !c.space.is2d())
        throw ConfigError("Sąsiedztwo „hexagonal” jest zdefiniowane tylko w 2D; dla przestrzeni 3D "
                          "wybierz von_neumann, moore lub pentagonal");

    c.terrain.type = get_enum<TerrainType>(j, "terrain.type", {{"cone", TerrainType::Cone}, {"file", TerrainType::File}});
    c.terrain.peak = get_number(j, "terrain.peakHeight", 0.0, 1e6);
    c.terrain.base = get_number(j, "terrain.baseRadius", 1e-9, 1e6);
    c.terrain.crater_r = get_number(j, "terrain.craterRadius", 0.0, 1e6);
    c.terrain.crater_d = get_number(j, "terrain.craterDepth", 0.0, 1e6);
    c.terrain.noise = get_number(j, "terrain.noise", 0.0, 1.0);
    const Json& file = field(j, "terrain.file");
    if (!file.is_null() && !file.is_string())
        throw ConfigError("Parametr „terrain.file” musi być ścieżką (tekst) albo null, podano: " + file.dump());
    c.terrain.file = file.is_string() ? file.get_string() : std::string();
    if (runs_ca && !c.ca.classic && c.terrain.type == TerrainType::File && c.terrain.file.empty())
        throw ConfigError("Dla terrain.type = \"file\" trzeba podać ścieżkę w „terrain.file”");

    c.vents.count = get_int(j, "vents.count", 1, 100000000);
    c.vents.placement = get_enum<VentPlacement>(j, "vents.placement",
                                                {{"crater", VentPlacement::Crater},
                                                 {"random", VentPlacement::Random},
                                                 {"positions", VentPlacement::Positions}});
    const size_t listed = get_columns(j, "vents.positions", c.space).size();
    const bool lava_ca = runs_ca && !c.ca.classic;
    if (lava_ca && c.vents.placement == VentPlacement::Positions) {
        if (listed == 0)
            throw ConfigError("vents.placement = \"positions\" wymaga listy kraterów w „vents.positions”, "
                              "np. [[80, 75]] (w GUI: kliknięcie na mapie wysokości)");
        if (static_cast<size_t>(c.vents.count) != listed)
            throw ConfigError("vents.count (" + std::to_string(c.vents.count) +
                              ") musi być równe liczbie kraterów w „vents.positions” (" + std::to_string(listed) +
                              "), bo przy placement = \"positions\" każdy krater ma podane położenie");
    } else if (lava_ca && listed > 0) {
        c.warnings.push_back("vents.positions jest ignorowane, bo vents.placement = \"" +
                             std::string(c.vents.placement == VentPlacement::Crater ? "crater" : "random") +
                             "\" (położenia losowane); kratery w zadanych miejscach: placement = \"positions\"");
    } else if (runs_ca && c.ca.classic && c.vents.placement == VentPlacement::Positions) {
        c.warnings.push_back("w trybie classic zarodki są losowane w całej przestrzeni, więc vents.placement = "
                             "\"positions\" i vents.positions są ignorowane");
    }
    c.vents.eruption_steps = get_int(j, "vents.eruptionSteps", 0, 65535);
    c.vents.erupt_temp = get_number(j, "vents.eruptTemp", -1e6, 1e6);
    if (runs_ca && !c.ca.classic && c.vents.erupt_temp < c.ca.solidus)
        throw ConfigError("vents.eruptTemp musi być nie mniejsze niż ca.solidusTemp, inaczej lawa "
                          "zastygłaby od razu w kraterze");
    if (runs_ca && c.ca.classic && static_cast<uint64_t>(c.vents.count) > c.space.cells())
        throw ConfigError("vents.count (" + std::to_string(c.vents.count) + ") nie może przekraczać liczby komórek (" +
                          std::to_string(c.space.cells()) + ")");

    if (c.mc.enabled && !c.mc.full && c.ca.classic)
        throw ConfigError("mc.domain = \"basalt\" wymaga ca.mode = \"lava\" (w trybie classic nie ma bazaltu); "
                          "klasyczny rozrost ziaren MC to mc.domain = \"full\"");
    if (c.mc.full_only() && c.mc.mobility_from_cooling)
        c.warnings.push_back("mc.mobilityFromCooling dotyczy tylko mc.domain = \"basalt\" (czas stygnięcia lawy z CA); "
                             "w trybie full jest ignorowane");

    c.output.format = get_enum<OutputFormat>(j, "output.format",
                                             {{"txt", OutputFormat::Txt},
                                              {"bin", OutputFormat::Bin},
                                              {"both", OutputFormat::Both},
                                              {"none", OutputFormat::None}});
    c.output.snapshots_every = get_int(j, "output.snapshotsEvery", 0, 65535);

    // CPU threading keys stay valid (configs written for the reference keep loading); on the GPU they
    // only select nothing and are reported back as given.
    c.parallel.threads = get_int(j, "parallel.threads", 0, 256);
    c.parallel.schedule = field(j, "parallel.schedule").is_string() ? field(j, "parallel.schedule").get_string() : "";
    get_enum<int>(j, "parallel.schedule", {{"static", 0}, {"dynamic", 1}, {"guided", 2}});
    c.parallel.chunk = get_int(j, "parallel.chunk", 0, 1000000);
    c.parallel.affinity = get_enum<std::string>(j, "parallel.affinity", {{"none", "none"}, {"pcores", "pcores"}});
    // the sequential MC update runs as one lane on the device as well; the warning text is the reference's
    if (c.mc.enabled && !c.mc.checkerboard && c.parallel.threads != 1)
        c.warnings.push_back("mc.update = \"random\" jest z natury sekwencyjne (każda próba widzi wynik poprzednich), "
                             "więc pętla MC idzie w jednym wątku; równoległe MC: mc.update = \"checkerboard\"");
    get_enum<bool>(j, "parallel.priority", {{"normal", false}, {"high", true}});
    get_enum<bool>(j, "mpi.exchange", {{"always", false}, {"changed", true}});
    get_enum<bool>(j, "mpi.pack", {{"manual", false}, {"datatype", true}});
    get_enum<bool>(j, "mpi.bcast", {{"json", false}, {"struct", true}});
    get_enum<bool>(j, "mpi.io", {{"gather", false}, {"mpiio", true}});
    c.effective = j;
    return c;
}

std::vector<std::pair<int, int>> Config::vent_columns() const { return get_columns(effective, "vents.positions", space); }

}  // namespace eng
