#include "terrain.hpp"

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <random>

namespace eng {

namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
constexpr uint64_t kSaltTerrain = 0x5445525241494E31ULL;  // engine/src/core/Rng.hpp:24
constexpr uint64_t kSaltVents = 0x56454E5453454544ULL;    // engine/src/core/Rng.hpp:25

// hashUniform (engine/src/core/Rng.hpp:46-59)
double hash_uniform(uint64_t seed, uint64_t step, uint64_t idx, uint64_t salt) {
    uint64_t h = splitmix64(seed ^ 0xA0761D6478BD642FULL);
    h = splitmix64(h ^ step);
    h = splitmix64(h ^ idx);
    h = splitmix64(h ^ salt);
    return static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
}

double lattice(uint64_t seed, int octave, int i, int j) {
    const uint64_t key = static_cast<uint64_t>(static_cast<uint32_t>(i)) | (static_cast<uint64_t>(static_cast<uint32_t>(j)) << 32);
    return 2.0 * hash_uniform(seed, static_cast<uint64_t>(octave), key, kSaltTerrain) - 1.0;
}
double smooth(double t) { return t * t * (3.0 - 2.0 * t); }
double value_noise(uint64_t seed, int octave, int spacing, int x, int y) {
    const int i = x / spacing, j = y / spacing;
    const double fx = smooth(static_cast<double>(x % spacing) / spacing);
    const double fy = smooth(static_cast<double>(y % spacing) / spacing);
    const double v00 = lattice(seed, octave, i, j), v10 = lattice(seed, octave, i + 1, j);
    const double v01 = lattice(seed, octave, i, j + 1), v11 = lattice(seed, octave, i + 1, j + 1);
    const double a = v00 + (v10 - v00) * fx;
    const double b = v01 + (v11 - v01) * fx;
    return a + (b - a) * fy;
}
double smooth_noise(uint64_t seed, int x, int y) {
    const double n = 1.0 * value_noise(seed, 0, 16, x, y) + 0.5 * value_noise(seed, 1, 8, x, y) +
                     0.25 * value_noise(seed, 2, 4, x, y);
    return n / 1.75;
}
double cone_height(const Config& c, double r) { return c.terrain.peak * std::max(0.0, 1.0 - r / c.terrain.base); }

// Fields of a CSV line: runs of characters between separators (',' ';' space tab CR), as [begin, end) offsets.
void split_fields(const std::string& line, std::vector<std::pair<size_t, size_t>>& out) {
    out.clear();
    size_t b = 0;
    const size_t n = line.size();
    for (size_t i = 0; i <= n; ++i) {
        const char ch = i < n ? line[i] : ',';
        if (ch == ',' || ch == ';' || ch == ' ' || ch == '\t' || ch == '\r') {
            if (i > b) out.emplace_back(b, i);
            b = i + 1;
        }
    }
}

}  // namespace

double distance_from_center(const Dims& d, int x, int y) {
    const double cx = (d.nx - 1) / 2.0, cy = (d.ny - 1) / 2.0;
    return std::hypot(x - cx, y - cy);
}

Heightmap cone_terrain(const Config& c) {
    const Dims& d = c.space;
    Heightmap hm;
    hm.nx = d.nx;
    hm.ny = d.ny;
    hm.height.resize(static_cast<size_t>(d.nx) * d.ny);
    const double amplitude = c.terrain.noise * c.terrain.peak;
    const double floor_h = cone_height(c, c.terrain.crater_r) - c.terrain.crater_d;
    for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x) {
            const double r = distance_from_center(d, x, y);
            double h = (r < c.terrain.crater_r) ? floor_h : cone_height(c, r);
            h += amplitude * smooth_noise(c.seed, x, y);
            const int cells = static_cast<int>(std::ceil(h));
            hm.height[static_cast<size_t>(x) + static_cast<size_t>(d.nx) * y] = std::clamp(cells, 0, d.nz);
        }
    return hm;
}

Heightmap load_heightmap_csv(const std::string& path, const Dims& d) {
    std::ifstream in(path);
    if (!in) throw ConfigError("Nie można otworzyć pliku mapy wysokości: " + path);
    Heightmap hm;
    hm.nx = d.nx;
    hm.ny = d.ny;
    hm.height.reserve(static_cast<size_t>(d.nx) * d.ny);
    std::string line;
    std::vector<std::pair<size_t, size_t>> f;
    int line_no = 0, row = 0;
    while (std::getline(in, line)) {
        ++line_no;
        split_fields(line, f);
        if (f.empty() || line[f[0].first] == '#') continue;
        if (row >= d.ny) throw ConfigError("Mapa wysokości " + path + " ma więcej wierszy danych niż ny = " + std::to_string(d.ny));
        if (static_cast<int>(f.size()) != d.nx)
            throw ConfigError("Mapa wysokości " + path + ", wiersz " + std::to_string(line_no) + ": " +
                              std::to_string(f.size()) + " wartości, oczekiwano nx = " + std::to_string(d.nx));
        for (size_t col = 0; col < f.size(); ++col) {
            // parsed in place: a separator ends any number, so the field must be consumed up to its end exactly
            const char* b = line.c_str() + f[col].first;
            const char* e = line.c_str() + f[col].second;
            char* end = nullptr;
            const double h = std::strtod(b, &end);
            if (end != e || !std::isfinite(h) || h < 0.0)
                throw ConfigError("Mapa wysokości " + path + ", wiersz " + std::to_string(line_no) + ", kolumna " +
                                  std::to_string(col + 1) + ": oczekiwano nieujemnej liczby, jest \"" +
                                  std::string(b, e) + "\"");
            hm.height.push_back(std::min(static_cast<int>(std::ceil(h)), d.nz));
        }
        ++row;
    }
    if (row != d.ny)
        throw ConfigError("Mapa wysokości " + path + " ma " + std::to_string(row) + " wierszy danych, oczekiwano ny = " +
                          std::to_string(d.ny));
    return hm;
}

// Priority-flood: outlets on non-periodic edges (or the lowest columns when both axes wrap), then each
// neighbour is flooded to max(own height, level of the column it was reached from).  The spill surface
// is unique, so the visiting order of equal levels does not matter.
Heightmap fill_depressions(const Heightmap& hm, bool px, bool py) {
    // Levels are integers and a column is pushed at a level no lower than the one being drained: a bucket queue
    // (one LIFO list per level, drained in increasing level), O(n + levels) instead of a heap's O(n log n).
    const int nx = hm.nx, ny = hm.ny;
    const size_t n = hm.height.size();
    Heightmap out = hm;
    if (n == 0) return out;
    const auto [mn, mx] = std::minmax_element(hm.height.begin(), hm.height.end());
    const int lo = *mn;
    const size_t levels = static_cast<size_t>(static_cast<int64_t>(*mx) - lo) + 1;
    std::vector<uint8_t> closed(n, 0);
    std::vector<int64_t> head(levels, -1);
    std::vector<int64_t> next(n, -1);
    size_t cur = levels;  // lowest level holding columns
    auto push = [&](size_t i, int level) {
        const size_t l = static_cast<size_t>(level - lo);
        next[i] = head[l];
        head[l] = static_cast<int64_t>(i);
        cur = std::min(cur, l);
    };
    auto seed = [&](size_t i) {
        if (!closed[i]) {
            closed[i] = 1;
            push(i, hm.height[i]);
        }
    };
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const bool ex = !px && nx > 1 && (x == 0 || x == nx - 1);
            const bool ey = !py && ny > 1 && (y == 0 || y == ny - 1);
            if (ex || ey) seed(static_cast<size_t>(x) + static_cast<size_t>(nx) * y);
        }
    if (cur == levels) {
        for (size_t i = 0; i < n; ++i)
            if (hm.height[i] == lo) seed(i);
    }
    const int dx[4] = {-1, 1, 0, 0}, dy[4] = {0, 0, -1, 1};
    for (; cur < levels; ++cur) {
        while (head[cur] >= 0) {
            const size_t i = static_cast<size_t>(head[cur]);
            head[cur] = next[i];
            const int level = lo + static_cast<int>(cur);
            const int x = static_cast<int>(i % static_cast<size_t>(nx)), y = static_cast<int>(i / static_cast<size_t>(nx));
            for (int k = 0; k < 4; ++k) {
                int sx = x + dx[k], sy = y + dy[k];
                if (sx < 0 || sx >= nx) {
                    if (!px || nx == 1) continue;
                    sx = (sx + nx) % nx;
                }
                if (sy < 0 || sy >= ny) {
                    if (!py || ny == 1) continue;
                    sy = (sy + ny) % ny;
                }
                const size_t j = static_cast<size_t>(sx) + static_cast<size_t>(nx) * sy;
                if (closed[j]) continue;
                closed[j] = 1;
                out.height[j] = std::max(hm.height[j], level);
                push(j, out.height[j]);  // at a level >= cur: drained in this pass or a later one
            }
        }
    }
    return out;
}

Heightmap build_terrain(const Config& c) {
    if (c.ca.classic) return Heightmap{};
    if (c.terrain.type == TerrainType::File) {
        namespace fs = std::filesystem;
        fs::path p(c.terrain.file);
        if (!p.is_absolute() && !fs::exists(p) && !c.source_path.empty()) {
            const fs::path next = fs::path(c.source_path).parent_path() / p;
            if (fs::exists(next)) p = next;
        }
        return load_heightmap_csv(p.string(), c.space);
    }
    return cone_terrain(c);
}

std::vector<Vent> build_vents(const Config& c, const Heightmap& hm) {
    const Dims& d = c.space;
    std::vector<Vent> vents;
    if (c.ca.classic) {
        // distinct random cells (VentPlacer.cpp:83-104): std::mt19937_64 is fully specified by the standard
        const uint64_t n = d.cells();
        std::mt19937_64 gen(splitmix64(c.seed ^ kSaltVents));
        std::vector<bool> taken(n, false);
        while (static_cast<int>(vents.size()) < c.vents.count) {
            const uint64_t idx = gen() % n;
            if (taken[idx]) continue;
            taken[idx] = true;
            vents.push_back(Vent{static_cast<int>(idx % static_cast<uint64_t>(d.nx)),
                                 static_cast<int>((idx / static_cast<uint64_t>(d.nx)) % static_cast<uint64_t>(d.ny)),
                                 static_cast<int>(idx / (static_cast<uint64_t>(d.nx) * d.ny)),
                                 static_cast<int32_t>(vents.size() + 1)});
        }
        return vents;
    }
    if (c.vents.placement == VentPlacement::Positions) {
        const auto cols = c.vent_columns();
        for (size_t i = 0; i < cols.size(); ++i) {
            const int cx = cols[i].first, cy = cols[i].second;
            const int h = hm.empty() ? 0 : hm.at(cx, cy);
            if (h >= d.nz)
                throw ConfigError("vents.positions[" + std::to_string(i) + "] = [" + std::to_string(cx) + ", " +
                                  std::to_string(cy) + "]: kolumna jest w całości skałą (wysokość terenu " +
                                  std::to_string(h) + " >= space.nz = " + std::to_string(d.nz) + ")");
            vents.push_back(Vent{cx, cy, h, static_cast<int32_t>(i + 1)});
        }
        return vents;
    }
    // crater / random: candidate columns in (y, x) order, partial Fisher-Yates (VentPlacer.cpp:11-58)
    std::vector<std::pair<int, int>> cand;
    for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x) {
            const int h = hm.at(x, y);
            if (h >= d.nz) continue;
            const double r = distance_from_center(d, x, y);
            bool ok;
            if (c.vents.placement == VentPlacement::Crater)
                ok = r <= c.terrain.crater_r;
            else
                ok = r > c.terrain.crater_r && h >= 1 && (c.terrain.type != TerrainType::Cone || r < c.terrain.base);
            if (ok) cand.emplace_back(x, y);
        }
    if (static_cast<int>(cand.size()) < c.vents.count)
        throw ConfigError(std::string("Za mało miejsc na kratery (placement = ") +
                          (c.vents.placement == VentPlacement::Crater ? "crater" : "random") + "): dostępnych " +
                          std::to_string(cand.size()) + ", a vents.count = " + std::to_string(c.vents.count));
    std::mt19937_64 gen(splitmix64(c.seed ^ kSaltVents));
    for (int i = 0; i < c.vents.count; ++i) {
        const size_t remaining = cand.size() - static_cast<size_t>(i);
        const size_t j = static_cast<size_t>(i) + static_cast<size_t>(gen() % remaining);
        std::swap(cand[static_cast<size_t>(i)], cand[j]);
        const auto [x, y] = cand[static_cast<size_t>(i)];
        vents.push_back(Vent{x, y, hm.at(x, y), i + 1});
    }
    return vents;
}

}  // namespace eng
