#include "writer.hpp"

#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <atomic>

#include "sha256.hpp"

namespace eng {

std::atomic<bool> g_digests{false};

namespace {

std::ofstream open_out(const std::string& path, bool binary) {
    std::ofstream out(path, binary ? std::ios::binary : std::ios::out);
    if (!out) throw std::runtime_error("Nie można utworzyć pliku: " + path);
    return out;
}

void finish(std::ofstream& out, const std::string& path) {
    out.flush();
    if (!out) throw std::runtime_error("Błąd zapisu pliku: " + path);
}

// One result file: written to disk, or -- with result digests on -- only hashed, its SHA-256 and size going to
// <path>.sha256 ("<hex> <bytes>\n") instead.  Either way it receives exactly the bytes of the file.
class ResultOut {
public:
    ResultOut(const std::string& path, bool binary) : path_(path), digest_(g_digests.load()) {
        if (!digest_) out_ = open_out(path, binary);
    }
    void write(const char* p, size_t n) {
        if (digest_)
            sha_.update(p, n);
        else
            out_.write(p, static_cast<std::streamsize>(n));
    }
    template <class T>
    void put(const std::vector<T>& v) {
        write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
    }
    void close() {
        if (!digest_) {
            finish(out_, path_);
            return;
        }
        std::ofstream d = open_out(path_ + ".sha256", false);
        const uint64_t n = sha_.bytes();
        d << sha_.hex() << ' ' << n << '\n';
        finish(d, path_ + ".sha256");
    }

private:
    std::string path_;
    bool digest_;
    std::ofstream out_;
    Sha256 sha_;
};

// Buffered text writer: appends into a ~1 MB buffer and flushes it in one write.
class Text {
public:
    explicit Text(ResultOut& out) : out_(out) { buf_.reserve((1 << 20) + 512); }
    ~Text() { flush(); }
    void add(const char* s, size_t n) {
        buf_.append(s, n);
        if (buf_.size() >= (1u << 20)) flush();
    }
    void add(const std::string& s) { add(s.data(), s.size()); }
    void flush() {
        if (!buf_.empty()) out_.write(buf_.data(), buf_.size());
        buf_.clear();
    }

private:
    ResultOut& out_;
    std::string buf_;
};

}  // namespace

void set_result_digests(bool on) { g_digests.store(on); }

void write_ca_bin(const std::string& path, const CaState& s) {
    ResultOut out(path, true);
    const char magic[4] = {'V', 'O', 'L', 'C'};
    const uint32_t version = 2;
    const int32_t dims[3] = {s.dims.nx, s.dims.ny, s.dims.nz};
    out.write(magic, 4);
    out.write(reinterpret_cast<const char*>(&version), 4);
    out.write(reinterpret_cast<const char*>(dims), 12);
    out.put(s.type);
    out.put(s.flow);
    out.put(s.temp);
    out.put(s.solid);
    out.put(s.arrival);
    out.close();
}

void write_ca_txt(const std::string& path, const CaState& s) {
    ResultOut out(path, true);
    {
        Text t(out);
        const Dims& d = s.dims;
        char line[192];
        int n = std::snprintf(line, sizeof(line), "%d %d %d\nx y z type flowId temp\n", d.nx, d.ny, d.nz);
        t.add(line, static_cast<size_t>(n));
        size_t i = 0;
        for (int z = 0; z < d.nz; ++z)
            for (int y = 0; y < d.ny; ++y)
                for (int x = 0; x < d.nx; ++x, ++i) {
                    // temp: the characters of printf("%.1f", (double)v) (ResultWriter.cpp:105-150)
                    n = std::snprintf(line, sizeof(line), "%d %d %d %d %d %.1f\n", x, y, z, static_cast<int>(s.type[i]),
                                      s.flow[i], static_cast<double>(s.temp[i]));
                    t.add(line, static_cast<size_t>(n));
                }
    }
    out.close();
}

void write_crystal_bin(const std::string& path, const Dims& d, const std::vector<int32_t>& crystal) {
    ResultOut out(path, true);
    const char magic[4] = {'V', 'M', 'C', '1'};
    const int32_t dims[3] = {d.nx, d.ny, d.nz};
    out.write(magic, 4);
    out.write(reinterpret_cast<const char*>(dims), 12);
    out.put(crystal);
    out.close();
}

void write_crystal_txt(const std::string& path, const Dims& d, const std::vector<int32_t>& crystal) {
    ResultOut out(path, true);
    {
        Text t(out);
        char line[96];
        int n = std::snprintf(line, sizeof(line), "%d %d %d\nx y z crystal\n", d.nx, d.ny, d.nz);
        t.add(line, static_cast<size_t>(n));
        size_t i = 0;
        for (int z = 0; z < d.nz; ++z)
            for (int y = 0; y < d.ny; ++y)
                for (int x = 0; x < d.nx; ++x, ++i) {
                    n = std::snprintf(line, sizeof(line), "%d %d %d %d\n", x, y, z, crystal[i]);
                    t.add(line, static_cast<size_t>(n));
                }
    }
    out.close();
}

void write_mc_stats_csv(const std::string& path, const std::vector<McRow>& rows) {
    std::ofstream out = open_out(path, false);
    out << "mcs,grains,meanGrainSize,energy,acceptRate\n";
    char line[160];
    for (const McRow& r : rows) {
        std::snprintf(line, sizeof(line), "%d,%llu,%.4f,%.3f,%.6f\n", r.mcs, static_cast<unsigned long long>(r.grains),
                      r.mean_grain_size, r.energy, r.accept_rate);
        out << line;
    }
    finish(out, path);
}

void write_json(const std::string& path, const Json& j) {
    std::ofstream out = open_out(path, false);
    out << j.dump(2) << '\n';
    finish(out, path);
}

CaState read_ca_folder(const std::string& folder, CaSource* src) {
    namespace fs = std::filesystem;
    const fs::path bin = fs::path(folder) / "ca_result.bin";
    const fs::path txt = fs::path(folder) / "ca_result.txt";
    CaState s;
    CaSource info;
    if (fs::exists(bin)) {
        std::ifstream in(bin, std::ios::binary);
        char magic[4];
        uint32_t version = 0;
        int32_t d[3];
        in.read(magic, 4);
        in.read(reinterpret_cast<char*>(&version), 4);
        in.read(reinterpret_cast<char*>(d), 12);
        if (!in || std::memcmp(magic, "VOLC", 4) != 0 || (version != 1 && version != 2))
            throw std::runtime_error("Nieprawidłowy plik wyniku CA: " + bin.string());
        if (d[0] < 1 || d[1] < 1 || d[2] < 1) throw std::runtime_error("Nieprawidłowe wymiary w " + bin.string());
        s.resize(Dims{d[0], d[1], d[2]});
        const size_t n = s.type.size();
        in.read(reinterpret_cast<char*>(s.type.data()), static_cast<std::streamsize>(n));
        in.read(reinterpret_cast<char*>(s.flow.data()), static_cast<std::streamsize>(n * 4));
        in.read(reinterpret_cast<char*>(s.temp.data()), static_cast<std::streamsize>(n * 4));
        in.read(reinterpret_cast<char*>(s.solid.data()), static_cast<std::streamsize>(n * 2));
        if (version >= 2) in.read(reinterpret_cast<char*>(s.arrival.data()), static_cast<std::streamsize>(n * 2));
        if (!in) throw std::runtime_error("Plik wyniku CA jest obcięty: " + bin.string());
        info = CaSource{bin.string(), "bin", static_cast<int>(version), true, version >= 2};
    } else if (fs::exists(txt)) {
        std::ifstream in(txt);
        int nx, ny, nz;
        std::string header;
        if (!(in >> nx >> ny >> nz)) throw std::runtime_error("Nieprawidłowy plik wyniku CA: " + txt.string());
        std::getline(in, header);
        std::getline(in, header);
        s.resize(Dims{nx, ny, nz});
        int x, y, z, t, f;
        double temp;
        while (in >> x >> y >> z >> t >> f >> temp) {
            if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) throw std::runtime_error("Komórka poza siatką w " + txt.string());
            const size_t i = static_cast<size_t>(x) + static_cast<size_t>(nx) * (static_cast<size_t>(y) + static_cast<size_t>(ny) * z);
            s.type[i] = static_cast<uint8_t>(t);
            s.flow[i] = f;
            s.temp[i] = static_cast<float>(temp);
        }
        info = CaSource{txt.string(), "txt", 0, false, false};
    } else {
        throw std::runtime_error("W folderze " + folder + " nie ma ca_result.bin ani ca_result.txt");
    }
    if (src) *src = info;
    return s;
}

}  // namespace eng
