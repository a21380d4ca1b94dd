"""Host-side engine logic on the CPU: the engine's spill levels (engine/terrain.cpp, a bucket-queue priority flood)
against the oracle's restatement of the reference's flood on random heightmaps, every boundary combination."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE_SRC = os.path.join(ROOT, "volcano_ca_mc_amd", "csrc", "engine")

HARNESS = r"""
#include <cstdio>
#include "terrain.hpp"
// stdin: nx ny px py, then nx*ny heights (row-major, y outer); stdout: the spill levels
int main() {
    int nx, ny, px, py;
    while (std::scanf("%d %d %d %d", &nx, &ny, &px, &py) == 4) {
        eng::Heightmap h;
        h.nx = nx;
        h.ny = ny;
        h.height.resize(static_cast<size_t>(nx) * ny);
        for (auto& v : h.height) std::scanf("%d", &v);
        const eng::Heightmap s = eng::fill_depressions(h, px != 0, py != 0);
        for (int v : s.height) std::printf("%d ", v);
        std::printf("\n");
    }
}
"""


@pytest.fixture(scope="module")
def fill_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fill")
    src = d / "fill.cpp"
    src.write_text(HARNESS)
    exe = d / "fill"
    srcs = [os.path.join(ENGINE_SRC, f) for f in ("terrain.cpp", "config.cpp", "json.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}", f"-I{ENGINE_SRC}", str(src),
                    *srcs, "-o", str(exe)], check=True)
    return str(exe)


def test_engine_spill_levels_match_oracle(fill_exe):
    rng = np.random.default_rng(11)
    maps, lines = [], []
    for _ in range(300):
        nx, ny = int(rng.integers(1, 33)), int(rng.integers(1, 33))
        h = rng.integers(0, int(rng.integers(1, 40)), size=(ny, nx), dtype=np.int32)
        for px in (0, 1):
            for py in (0, 1):
                maps.append((h, px, py))
                lines.append(f"{nx} {ny} {px} {py} " + " ".join(map(str, h.ravel().tolist())))
    out = subprocess.run([fill_exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    rows = out.stdout.strip().split("\n")
    assert len(rows) == len(maps)
    for (h, px, py), row in zip(maps, rows):
        got = np.array(row.split(), dtype=np.int32).reshape(h.shape)
        assert np.array_equal(got, O.fill_depressions(h, bool(px), bool(py))), (h.shape, px, py)


SHA_HARNESS = r"""
#include <cstdio>
#include <vector>
#include "sha256.hpp"
// stdin: records "<n> <chunk>" then n raw bytes; stdout: one hex digest per record (the bytes fed in chunks)
int main() {
    unsigned long n, chunk;
    while (std::scanf("%lu %lu", &n, &chunk) == 2) {
        std::getchar();
        std::vector<unsigned char> b(n);
        if (n) std::fread(b.data(), 1, n, stdin);
        eng::Sha256 h;
        for (unsigned long at = 0; at < n; at += chunk) h.update(b.data() + at, n - at < chunk ? n - at : chunk);
        std::printf("%s %llu\n", h.hex().c_str(), (unsigned long long)h.bytes());
    }
}
"""


def test_engine_sha256_matches_hashlib(tmp_path):
    """The digest of --digests mode (engine/sha256.cpp, FIPS 180-4) equals Python's hashlib on lengths around every
    padding boundary, with the bytes fed in chunks of different sizes."""
    import hashlib
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "sha.cpp"
    src.write_text(SHA_HARNESS)
    exe = tmp_path / "sha"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ENGINE_SRC}", str(src), os.path.join(ENGINE_SRC, "sha256.cpp"),
                    "-o", str(exe)], check=True)
    rng = np.random.default_rng(5)
    blobs, payload = [], b""
    for n in list(range(0, 130)) + [447, 448, 511, 512, 1000, 65536, 300001]:
        for chunk in (1, 7, 64, 1 << 20):
            if chunk == 1 and n > 1000:
                continue
            b = rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()
            blobs.append(b)
            payload += f"{n} {chunk}\n".encode() + b
    out = subprocess.run([str(exe)], input=payload, capture_output=True, check=True).stdout.decode().split("\n")
    for b, line in zip(blobs, out):
        assert line == f"{hashlib.sha256(b).hexdigest()} {len(b)}"
    assert hashlib.sha256(b"abc").hexdigest().startswith("ba7816bf")  # FIPS 180-4 example B.1
