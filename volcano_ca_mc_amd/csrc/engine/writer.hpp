// Result files.  Byte formats are the reference's (engine/src/io/ResultWriter.hpp:4-32):
//   ca_result.bin  "VOLC", u32 version 2, i32 nx ny nz, u8 type[N], i32 flowId[N], f32 temp[N],
//                  u16 solidStep[N], u16 arrivalStep[N]   (little-endian, no frame, x fastest)
//   ca_result.txt  "nx ny nz" / "x y z type flowId temp" header, one line per cell, temp as %.1f
//   mc_result.bin  "VMC1", i32 nx ny nz, i32 crystal[N]
//   mc_result.txt  "nx ny nz" / "x y z crystal" header, one line per cell
//   mc_stats.csv   mcs,grains,meanGrainSize,energy,acceptRate  ("%d,%llu,%.4f,%.3f,%.6f")
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"
#include "json.hpp"

namespace eng {

struct CaState {
    Dims dims;
    std::vector<uint8_t> type;
    std::vector<int32_t> flow;
    std::vector<float> temp;
    std::vector<uint16_t> solid;
    std::vector<uint16_t> arrival;
    void resize(const Dims& d) {
        dims = d;
        const size_t n = d.cells();
        type.assign(n, 0);
        flow.assign(n, 0);
        temp.assign(n, 0.0f);
        solid.assign(n, 0);
        arrival.assign(n, 0);
    }
};

struct McRow {
    int mcs = 0;
    uint64_t grains = 0;
    double mean_grain_size = 0.0, energy = 0.0, accept_rate = 0.0;
};

// Result digests (engine --digests): the writers below record each result file's SHA-256 and size in <path>.sha256
// instead of writing the file.  Process-wide; set before any job starts.
void set_result_digests(bool on);

void write_ca_bin(const std::string& path, const CaState& s);
void write_ca_txt(const std::string& path, const CaState& s);
void write_crystal_bin(const std::string& path, const Dims& d, const std::vector<int32_t>& crystal);
void write_crystal_txt(const std::string& path, const Dims& d, const std::vector<int32_t>& crystal);
void write_mc_stats_csv(const std::string& path, const std::vector<McRow>& rows);
void write_json(const std::string& path, const Json& j);

// ca_result.bin (v1 without arrivalStep, v2) or ca_result.txt of a folder (engine/src/io/ResultReader.cpp).
struct CaSource {
    std::string path, format;
    int version = 0;
    bool has_solid = false, has_arrival = false;
};
CaState read_ca_folder(const std::string& folder, CaSource* src);

}  // namespace eng
