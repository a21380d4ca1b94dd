// SHA-256 (FIPS 180-4) of a byte stream: the engine's --digests mode records the digest of each result file instead
// of writing it, so that runs whose outputs are far larger than a disk holds can still be compared byte for byte.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace eng {

class Sha256 {
public:
    Sha256();
    void update(const void* data, size_t n);
    // Hex digest of everything given to update (the object is finished afterwards).
    std::string hex();
    uint64_t bytes() const { return total_; }

private:
    void block(const uint8_t* p);
    uint32_t h_[8];
    uint8_t buf_[64];
    size_t used_ = 0;
    uint64_t total_ = 0;
};

}  // namespace eng
