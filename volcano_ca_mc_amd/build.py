"""Builds the native parts in-tree:

  volcano_ca_mc_amd/_lib/libvca.so   HIP kernels + C ABI (include/vca.h), hipcc for gfx950
  volcano_ca_mc_amd/_lib/engine      the drop-in `engine` CLI (C++17 host code linked to libvca.so)

Every HIP translation unit is compiled with -ffp-contract=off (temperatures are bit-exact float
subtractions) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libvca.so")
ENGINE = os.path.join(LIB_DIR, "engine")
CALIB = os.path.join(LIB_DIR, "calib_copy")
INCLUDE = os.path.join(ROOT, "include")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
]
LIB_SOURCES = ["vca_device.hip", "vca_api.hip"]
ENGINE_SOURCES = ["engine/sha256.cpp", "engine/json.cpp", "engine/config.cpp", "engine/terrain.cpp", "engine/writer.cpp",
                  "engine/simulation.cpp", "engine/batch.cpp", "engine/pipeline.cpp", "engine/main.cpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(jobs: int = 8) -> str:
    """Compiles every HIP source (objects in parallel) and links libvca.so."""
    os.makedirs(os.path.join(LIB_DIR, "obj"), exist_ok=True)
    procs = []
    objs = []
    for src in LIB_SOURCES:
        obj = os.path.join(LIB_DIR, "obj", os.path.basename(src) + ".o")
        objs.append(obj)
        procs.append((src, subprocess.Popen([hipcc(), *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj],
                                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(f"{src}:\n{out}")
    if errors:
        raise RuntimeError("HIP compile failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_engine() -> str:
    """Compiles the C++17 host engine and links it against libvca.so (rpath to its own directory)."""
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return ""
    tmp = ENGINE + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", f"-I{INCLUDE}",
          f"-I{os.path.join(CSRC, 'engine')}", *srcs, "-o", tmp, f"-L{LIB_DIR}", "-lvca",
          "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(tmp, ENGINE)
    return ENGINE


def build_calib() -> str:
    """Byte-counter calibration program (a plain copy of known size) for rocprofv3 --pmc runs."""
    tmp = CALIB + ".tmp"
    _run([hipcc(), *HIP_FLAGS, os.path.join(CSRC, "calib_copy.hip"), "-o", tmp])
    os.replace(tmp, CALIB)
    return CALIB


def build(engine: bool = True) -> None:
    """The product: libvca.so and the engine (the counter-calibration program is measurement tooling, built by
    build_calib on its own)."""
    build_lib()
    if engine:
        build_engine()


if __name__ == "__main__":
    build(engine="--no-engine" not in sys.argv)
    if "--calib" in sys.argv:
        build_calib()
    print(LIB)
