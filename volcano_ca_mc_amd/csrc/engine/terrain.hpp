// Host-side inputs of the lava CA: heightmap (cone generator or CSV), spill levels, craters / seeds.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "config.hpp"

namespace eng {

// Rock cells per column (x + nx*y); column h has ROCK for z < h.
struct Heightmap {
    int nx = 0, ny = 0;
    std::vector<int32_t> height;
    int at(int x, int y) const { return height[static_cast<size_t>(x) + static_cast<size_t>(nx) * y]; }
    bool empty() const { return height.empty(); }
};

struct Vent {
    int x, y, z;
    int32_t flow_id;
};

// Cone with a flat crater floor and 3-octave value noise (engine/src/terrain/TerrainGenerator.cpp:53-84).
Heightmap cone_terrain(const Config& c);
// CSV heightmap: ny rows of nx values, ceil(h) clamped to nz (engine/src/terrain/HeightmapLoader.cpp:37-85).
Heightmap load_heightmap_csv(const std::string& path, const Dims& d);
// Spill level per column by priority-flood (engine/src/terrain/TerrainGenerator.cpp:86-160).
Heightmap fill_depressions(const Heightmap& hm, bool periodic_x, bool periodic_y);
// Terrain of a config (cone or file, file resolved next to the config when not found as given).
Heightmap build_terrain(const Config& c);
// Craters (lava) or grain seeds (classic) (engine/src/terrain/VentPlacer.cpp:11-104).
std::vector<Vent> build_vents(const Config& c, const Heightmap& hm);
double distance_from_center(const Dims& d, int x, int y);

}  // namespace eng
