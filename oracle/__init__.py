"""Python binding of the CPU oracle (oracle/oracle.c) and a whole-config replay.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg are the only users; the product package never imports this module.

``run_config`` replays the reference engine's plain run (``engine/src/app/Simulation.cpp:294-429``
and ``engine/src/app/McStage.cpp:93-301``) on top of the C restatement and returns the bytes of
``ca_result.bin`` / ``mc_result.bin``, the text of ``mc_stats.csv`` and the ``stats.json``
values, so every output file can be compared with the reference's and with the GPU engine's.
"""
from __future__ import annotations

import copy
import ctypes as C
import json
import math
import os
import struct
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_ENGINE = os.path.join(HERE, "_ref", "engine")

AIR, ROCK, LAVA, BASALT = 0, 1, 2, 3
NB_KINDS = {"von_neumann": 0, "moore": 1, "pentagonal": 2, "hexagonal": 3}
STOP_TEXT = {0: "trwa", 1: "brak płynnej lawy", 2: "brak wolnych komórek (AIR)", 3: "brak zmian w kroku",
             4: "osiągnięto ca.maxSteps"}  # CaRules.hpp:110-119

_lib = None


def build() -> str:
    """Compiles liboracle.so (gcc) if it is missing or older than its sources."""
    src = [os.path.join(HERE, "oracle.c"), os.path.join(HERE, "oracle.h")]
    if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(os.path.getmtime(s) for s in src):
        import subprocess
        subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return LIB_PATH


class _NB(C.Structure):
    _fields_ = [("count", C.c_int), ("size", C.c_int * 6), ("d", C.c_int8 * (6 * 26 * 3))]


class CaParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("classic", C.c_int),
                ("neighborhood", C.c_int), ("periodic", C.c_int), ("seed", C.c_uint64),
                ("max_steps", C.c_int), ("eruption_steps", C.c_int), ("erupt_temp", C.c_float),
                ("solidus_temp", C.c_float), ("cooling_ambient", C.c_float), ("cooling_lateral", C.c_float),
                ("cooling_fall", C.c_float), ("ponding", C.c_int)]


class McParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("neighborhood", C.c_int),
                ("periodic", C.c_int * 3), ("seed", C.c_uint64), ("Q", C.c_int), ("J", C.c_double),
                ("kT", C.c_double), ("checkerboard", C.c_int)]


class McStats(C.Structure):
    _fields_ = [("mcs", C.c_int), ("grains", C.c_uint64), ("cells", C.c_uint64), ("largest", C.c_uint64),
                ("mean_grain_size", C.c_double), ("energy", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(build())
        u64, i32p = C.c_uint64, C.POINTER(C.c_int32)
        for name, res, args in [
            ("orc_splitmix64", u64, [u64]), ("orc_hash_rand", u64, [u64, u64, u64, u64]),
            ("orc_hash_prefix", u64, [u64, u64]), ("orc_hash_index", u64, [u64, u64]),
            ("orc_hash_salt", u64, [u64, u64]), ("orc_salt", u64, [C.c_int]),
            ("orc_nb_build", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_NB)]),
            ("orc_fill_depressions", None, [i32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p]),
            ("orc_ca_create", C.c_void_p, [C.POINTER(CaParams), i32p, i32p, i32p, C.c_int]),
            ("orc_ca_destroy", None, [C.c_void_p]), ("orc_ca_step", C.c_int, [C.c_void_p]),
            ("orc_ca_run", C.c_int, [C.c_void_p]), ("orc_ca_steps", C.c_int, [C.c_void_p]),
            ("orc_ca_stop_reason", C.c_int, [C.c_void_p]),
            ("orc_ca_get", None, [C.c_void_p] + [C.c_void_p] * 5),
            ("orc_ca_histories", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
            ("orc_mc_create", C.c_void_p, [C.POINTER(McParams), C.c_void_p, C.c_void_p]),
            ("orc_mc_destroy", None, [C.c_void_p]), ("orc_mc_sweep", u64, [C.c_void_p]),
            ("orc_mc_counters", None, [C.c_void_p] + [C.POINTER(u64)] * 3),
            ("orc_mc_colors", C.c_int, [C.c_void_p]), ("orc_mc_domain_cells", u64, [C.c_void_p]),
            ("orc_mc_spins", None, [C.c_void_p, C.c_void_p]),
            ("orc_mc_measure", None, [C.c_void_p, C.POINTER(McStats), C.c_void_p]),
            ("orc_mc_total_energy", C.c_double, [C.c_void_p]),
            ("orc_mc_set_spin", None, [C.c_void_p, C.c_int64, C.c_int32]),
            ("orc_mc_color_list", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
            ("orc_grains", u64, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int * 3),
                                 C.c_void_p, C.POINTER(u64), C.POINTER(u64)]),
            ("orc_mobility_from_cooling", None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double,
                                                 C.c_void_p, C.c_void_p]),
            ("orc_cone", None, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, u64, i32p]),
            ("orc_place_vents", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, i32p, u64, i32p]),
            ("orc_place_seeds", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, u64, i32p]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray, ctype=C.c_int32):
    return a.ctypes.data_as(C.POINTER(ctype))


# ------------------------------------------------------------------------------------------
# small wrappers
# ------------------------------------------------------------------------------------------

def neighborhood(kind: int, nx: int, ny: int, nz: int):
    """Variant offset lists [(dx, dy, dz), ...] per variant, or None when unsupported."""
    nb = _NB()
    if lib().orc_nb_build(kind, nx, ny, nz, C.byref(nb)) != 0:
        return None
    d = np.frombuffer(bytes(nb.d), dtype=np.int8).reshape(6, 26, 3)
    return [[tuple(int(v) for v in d[var, k]) for k in range(nb.size[var])] for var in range(nb.count)]


def fill_depressions(height: np.ndarray, periodic_x: bool, periodic_y: bool) -> np.ndarray:
    h = np.ascontiguousarray(height, dtype=np.int32)
    ny, nx = h.shape
    out = np.empty_like(h)
    lib().orc_fill_depressions(_p(h), nx, ny, int(periodic_x), int(periodic_y), _p(out))
    return out


def grains(crystal: np.ndarray, dims, moore: bool, periodic):
    """(grains, cells, largest, root) of a crystal map (global-index order)."""
    nx, ny, nz = dims
    c = np.ascontiguousarray(crystal, dtype=np.int32)
    root = np.empty(c.size, dtype=np.uint32)
    cells, largest = C.c_uint64(), C.c_uint64()
    per = (C.c_int * 3)(*[int(bool(v)) for v in periodic])
    g = lib().orc_grains(c.ctypes.data, nx, ny, nz, int(moore), C.byref(per), root.ctypes.data,
                         C.byref(cells), C.byref(largest))
    return int(g), int(cells.value), int(largest.value), root


class OracleCA:
    def __init__(self, params: CaParams, height: Optional[np.ndarray], spill: Optional[np.ndarray], vents):
        self.params = params
        v = np.ascontiguousarray(np.asarray(vents, dtype=np.int32).reshape(-1, 4))
        h = None if height is None else np.ascontiguousarray(height, dtype=np.int32)
        s = None if spill is None else np.ascontiguousarray(spill, dtype=np.int32)
        self.h = lib().orc_ca_create(C.byref(params), None if h is None else _p(h), None if s is None else _p(s),
                                     _p(v), len(v))
        if not self.h:
            raise ValueError("unsupported CA configuration")
        self.n = params.nx * params.ny * params.nz

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ca_destroy(self.h)
            self.h = None

    def step(self) -> bool:
        return bool(lib().orc_ca_step(self.h))

    def run(self) -> int:
        return lib().orc_ca_run(self.h)

    @property
    def steps(self) -> int:
        return lib().orc_ca_steps(self.h)

    @property
    def stop_reason(self) -> int:
        return lib().orc_ca_stop_reason(self.h)

    def state(self):
        n = self.n
        t = np.empty(n, np.uint8); f = np.empty(n, np.int32); tp = np.empty(n, np.float32)
        s = np.empty(n, np.uint16); a = np.empty(n, np.uint16)
        lib().orc_ca_get(self.h, t.ctypes.data, f.ctypes.data, tp.ctypes.data, s.ctypes.data, a.ctypes.data)
        return {"type": t, "flow": f, "temp": tp, "solid": s, "arrival": a}

    def histories(self):
        cap = max(1, self.steps)
        act = np.zeros(cap, np.uint64); chg = np.zeros(cap, np.uint64)
        k = lib().orc_ca_histories(self.h, act.ctypes.data, chg.ctypes.data, cap)
        return act[:k].tolist(), chg[:k].tolist()


class OracleMC:
    def __init__(self, params: McParams, domain: Optional[np.ndarray] = None, mobility: Optional[np.ndarray] = None):
        self.params = params
        self.dom = None if domain is None else np.ascontiguousarray(domain, dtype=np.uint8)
        self.mob = None if mobility is None else np.ascontiguousarray(mobility, dtype=np.float32)
        self.h = lib().orc_mc_create(C.byref(params), None if self.dom is None else self.dom.ctypes.data,
                                     None if self.mob is None else self.mob.ctypes.data)
        if not self.h:
            raise ValueError("unsupported MC configuration")
        self.n = params.nx * params.ny * params.nz

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mc_destroy(self.h)
            self.h = None

    def sweep(self) -> int:
        return int(lib().orc_mc_sweep(self.h))

    def counters(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib().orc_mc_counters(self.h, C.byref(a), C.byref(b), C.byref(c))
        return int(a.value), int(b.value), int(c.value)

    @property
    def colors(self) -> int:
        return lib().orc_mc_colors(self.h)

    @property
    def domain_cells(self) -> int:
        return int(lib().orc_mc_domain_cells(self.h))

    def spins(self) -> np.ndarray:
        out = np.empty(self.n, np.int32)
        lib().orc_mc_spins(self.h, out.ctypes.data)
        return out

    def set_spin(self, g: int, s: int) -> None:
        lib().orc_mc_set_spin(self.h, g, s)

    def total_energy(self) -> float:
        return float(lib().orc_mc_total_energy(self.h))

    def color_lists(self):
        starts = np.zeros(29, np.int64)
        lst = np.zeros(max(1, self.domain_cells), np.int64)
        c = lib().orc_mc_color_list(self.h, starts.ctypes.data, lst.ctypes.data)
        return [lst[starts[k]:starts[k + 1]] for k in range(c)]

    def measure(self, want_root: bool = False):
        st = McStats()
        root = np.empty(self.n, np.uint32) if want_root else None
        lib().orc_mc_measure(self.h, C.byref(st), None if root is None else root.ctypes.data)
        return st, root


# ------------------------------------------------------------------------------------------
# whole-config replay
# ------------------------------------------------------------------------------------------

# Config.cpp:324-346 (keys and defaults the reference accepts)
DEFAULTS = {
    "space": {"nx": 200, "ny": 1, "nz": 100}, "grid": "vector", "boundary": "absorbing", "neighborhood": "moore",
    "rngSeed": 42,
    "terrain": {"type": "cone", "peakHeight": 70, "baseRadius": 90, "craterRadius": 6, "craterDepth": 4,
                "noise": 0.05, "file": None},
    "vents": {"count": 1, "placement": "crater", "positions": [], "eruptionSteps": 300, "eruptTemp": 1200},
    "ca": {"mode": "lava", "maxSteps": 5000, "coolingAmbient": 0.3, "coolingLateral": 1.5, "coolingFall": 0.1,
           "solidusTemp": 800, "ponding": True, "sweep": "front"},
    "mc": {"enabled": False, "domain": "basalt", "Q": 50, "mcs": 100, "J": 1.0, "kT": 0.0,
           "mobilityFromCooling": True, "mobilityMin": 0.05, "statsEvery": 10, "snapshotsEvery": 0,
           "update": "random"},
    "output": {"format": "txt", "snapshotsEvery": 0},
    "parallel": {"threads": 1, "schedule": "static", "chunk": 0, "affinity": "none", "priority": "normal"},
    "mpi": {"exchange": "always", "pack": "manual", "bcast": "json", "io": "gather"},
}


def merged(cfg: dict) -> dict:
    out = copy.deepcopy(DEFAULTS)

    def rec(base, patch):
        for k, v in patch.items():
            if k.startswith("_"):
                continue
            if isinstance(base.get(k), dict):
                rec(base[k], v)
            else:
                base[k] = v
    rec(out, cfg)
    return out


def load_heightmap_csv(path: str, nx: int, ny: int, nz: int) -> np.ndarray:
    """HeightmapLoader.cpp:37-85: ceil(h) clamped to nz; rows = y, columns = x."""
    rows = []
    with open(path) as f:
        for line in f:
            fields = [t for t in line.replace(";", " ").replace(",", " ").replace("\t", " ").replace("\r", " ").split(" ") if t]
            if not fields or fields[0].startswith("#"):
                continue
            rows.append([min(int(math.ceil(float(v))), nz) for v in fields])
    h = np.asarray(rows, dtype=np.int32)
    assert h.shape == (ny, nx), (h.shape, ny, nx)
    return h


_libm = C.CDLL("libm.so.6")
_libm.hypot.restype = C.c_double
_libm.hypot.argtypes = [C.c_double, C.c_double]
_libm.round.restype = C.c_double
_libm.round.argtypes = [C.c_double]
_libm.pow.restype = C.c_double
_libm.pow.argtypes = [C.c_double, C.c_double]


def _rounded(v: float, decimals: int) -> float:
    f = _libm.pow(10.0, float(decimals))
    return _libm.round(v * f) / f


def terrain_for(cfg: dict, config_dir: str = ".") -> np.ndarray:
    nx, ny, nz = cfg["space"]["nx"], cfg["space"]["ny"], cfg["space"]["nz"]
    t = cfg["terrain"]
    if t["type"] == "file":
        path = t["file"]
        if not os.path.isabs(path) and not os.path.exists(path):
            path = os.path.join(config_dir, path)
        return load_heightmap_csv(path, nx, ny, nz)
    out = np.empty((ny, nx), np.int32)
    lib().orc_cone(nx, ny, nz, float(t["peakHeight"]), float(t["baseRadius"]), float(t["craterRadius"]),
                   float(t["craterDepth"]), float(t["noise"]), int(cfg["rngSeed"]), _p(out))
    return out


def vents_for(cfg: dict, hm: Optional[np.ndarray]) -> np.ndarray:
    nx, ny, nz = cfg["space"]["nx"], cfg["space"]["ny"], cfg["space"]["nz"]
    v = cfg["vents"]
    count = int(v["count"])
    out = np.empty((count, 4), np.int32)
    if cfg["ca"]["mode"] == "classic":
        lib().orc_place_seeds(count, nx, ny, nz, int(cfg["rngSeed"]), _p(out))
        return out
    if v["placement"] == "positions":
        for i, (x, y) in enumerate(v["positions"]):
            out[i] = (x, y, 0 if hm is None else hm[y, x], i + 1)
        return out
    t = cfg["terrain"]
    k = lib().orc_place_vents(count, 0 if v["placement"] == "crater" else 1, int(t["type"] == "cone"),
                              float(t["craterRadius"]), float(t["baseRadius"]), nx, ny, nz,
                              _p(np.ascontiguousarray(hm)), int(cfg["rngSeed"]), _p(out))
    if k < 0:
        raise ValueError("not enough vent columns")
    return out


def ca_params(cfg: dict) -> CaParams:
    """Simulation.cpp:122-136 (double config values narrowed to float)."""
    s, ca, v = cfg["space"], cfg["ca"], cfg["vents"]
    f32 = lambda x: float(np.float32(x))
    return CaParams(s["nx"], s["ny"], s["nz"], int(ca["mode"] == "classic"), NB_KINDS[cfg["neighborhood"]],
                    int(cfg["boundary"] == "periodic"), int(cfg["rngSeed"]), int(ca["maxSteps"]),
                    int(v["eruptionSteps"]), f32(v["eruptTemp"]), f32(ca["solidusTemp"]),
                    f32(ca["coolingAmbient"]), f32(ca["coolingLateral"]), f32(ca["coolingFall"]),
                    int(bool(ca["ponding"])))


def ca_result_bin(dims, st) -> bytes:
    """ResultWriter.cpp:155-167, 204-213."""
    nx, ny, nz = dims
    return (b"VOLC" + struct.pack("<Iiii", 2, nx, ny, nz) + st["type"].tobytes() + st["flow"].astype("<i4").tobytes()
            + st["temp"].astype("<f4").tobytes() + st["solid"].astype("<u2").tobytes()
            + st["arrival"].astype("<u2").tobytes())


def mc_result_bin(dims, crystal: np.ndarray) -> bytes:
    """ResultWriter.cpp:194-202, 215-222."""
    nx, ny, nz = dims
    return b"VMC1" + struct.pack("<iii", nx, ny, nz) + crystal.astype("<i4").tobytes()


def mc_stats_csv(rows) -> str:
    """ResultWriter.cpp:224-236."""
    out = "mcs,grains,meanGrainSize,energy,acceptRate\n"
    for r in rows:
        out += "%d,%d,%.4f,%.3f,%.6f\n" % (r["mcs"], r["grains"], r["meanGrainSize"], r["energy"], r["acceptRate"])
    return out


def ca_stats(cfg: dict, steps: int, stop: int, vents: np.ndarray, st, act, chg) -> dict:
    """Simulation.cpp:408-524 (flowSums + statsFromSums)."""
    nx, ny, nz = cfg["space"]["nx"], cfg["space"]["ny"], cfg["space"]["nz"]
    flows = len(vents)
    t = st["type"].reshape(nz, ny, nx)
    f = st["flow"].reshape(nz, ny, nx)
    lava = [0] * (flows + 1)
    basalt = [0] * (flows + 1)
    reach = [0.0] * (flows + 1)
    sel = ((t == LAVA) | (t == BASALT)) & (f > 0) & (f <= flows)
    zz, yy, xx = np.nonzero(sel)
    for z, y, x in zip(zz.tolist(), yy.tolist(), xx.tolist()):
        fi = int(f[z, y, x])
        if t[z, y, x] == LAVA:
            lava[fi] += 1
        else:
            basalt[fi] += 1
        v = vents[fi - 1]
        d = _libm.hypot(float(x - int(v[0])), float(y - int(v[1])))
        if reach[fi] < d:
            reach[fi] = d
    lava_mode = cfg["ca"]["mode"] == "lava"
    j = {"mode": cfg["ca"]["mode"], "steps": steps, "stopReason": STOP_TEXT[stop]}
    total = total_lava = 0
    max_reach = 0.0
    mins, maxs = None, 0
    lst = []
    for fi in range(1, flows + 1):
        cells = lava[fi] + basalt[fi]
        total += cells
        total_lava += lava[fi]
        max_reach = max(max_reach, reach[fi])
        mins = cells if mins is None else min(mins, cells)
        maxs = max(maxs, cells)
        if fi <= 10000:
            v = vents[fi - 1]
            e = {"flowId": fi, "vent": [int(v[0]), int(v[1]), int(v[2])], "cells": cells}
            if lava_mode:
                e.update({"lava": lava[fi], "basalt": basalt[fi], "maxReach": _rounded(reach[fi], 2)})
            lst.append(e)
    if lava_mode:
        j.update({"totalLavaVolume": total, "liquidLava": total_lava, "maxReach": _rounded(max_reach, 2)})
    else:
        j.update({"grains": flows, "meanGrainSize": (total / flows) if flows else 0.0,
                  "minGrainSize": mins if flows else 0, "maxGrainSize": maxs})
    j["flows"] = lst
    if flows > 10000:
        j["flowsTruncated"] = True
    j["activeCells"] = act
    j["changedCells"] = chg
    return j


def _classify(root: np.ndarray, mobility: np.ndarray, threshold: float = 0.5):
    """McStage.cpp:131-150: a grain is 'high' when at least half of its cells have mobility >= threshold."""
    inside = root != 0xFFFFFFFF
    r = root[inside].astype(np.int64)
    hi = (mobility[inside].astype(np.float64) >= threshold).astype(np.int64)
    n = root.size
    size = np.bincount(r, minlength=n)
    high_cells = np.bincount(r, weights=hi, minlength=n).astype(np.int64)
    roots = np.nonzero(size > 0)[0]
    out = {}
    for name, mask in (("high", 2 * high_cells[roots] >= size[roots]), ("low", 2 * high_cells[roots] < size[roots])):
        s = size[roots][mask]
        out[name] = {"grains": int(s.size), "grainCells": int(s.sum()), "largest": int(s.max()) if s.size else 0}
    return out


def run_mc(cfg: dict, st=None):
    """McStage::run (McStage.cpp:93-301) on a CA result ``st`` (basalt domain) or the whole space."""
    nx, ny, nz = cfg["space"]["nx"], cfg["space"]["ny"], cfg["space"]["nz"]
    mc = cfg["mc"]
    full = st is None
    periodic = cfg["boundary"] == "periodic"
    per = [periodic and nx > 1, periodic and ny > 1, periodic and full and nz > 1]  # McStage.cpp:102-106
    params = McParams(nx, ny, nz, NB_KINDS[cfg["neighborhood"]], (C.c_int * 3)(*map(int, per)),
                      int(cfg["rngSeed"]), int(mc["Q"]), float(mc["J"]), float(mc["kT"]),
                      int(mc["update"] == "checkerboard"))
    domain = mobility = cell_mobility = None
    if not full:
        n = nx * ny * nz
        domain = np.empty(n, np.uint8)
        cell_mobility = np.empty(n, np.float32)
        lib().orc_mobility_from_cooling(st["type"].ctypes.data, st["solid"].ctypes.data, st["arrival"].ctypes.data,
                                        n, float(mc["mobilityMin"]), domain.ctypes.data, cell_mobility.ctypes.data)
        if mc["mobilityFromCooling"]:
            mobility = cell_mobility
    m = OracleMC(params, domain, mobility)
    moore = cfg["neighborhood"] != "von_neumann"
    rows = []

    def measure(keep: bool):
        s, root = m.measure(want_root=keep)
        return {"mcs": s.mcs, "grains": int(s.grains), "meanGrainSize": s.mean_grain_size, "energy": s.energy,
                "acceptRate": 0.0, "cells": int(s.cells), "largest": int(s.largest)}, root

    r0, start_root = measure(True)
    rows.append(r0)
    att_before = acc_before = 0
    end_root = start_root
    end_row = r0
    mcs = int(mc["mcs"])
    every = int(mc["statsEvery"])
    for k in range(1, mcs + 1):
        m.sweep()
        if k % every == 0 or k == mcs:
            row, root = measure(k == mcs)
            att, acc, _ = m.counters()
            row["acceptRate"] = (acc - acc_before) / (att - att_before) if att - att_before > 0 else 0.0
            att_before, acc_before = att, acc
            rows.append(row)
            if k == mcs:
                end_root, end_row = root, row
    attempts, accepted, uphill = m.counters()
    first, last = rows[0], rows[-1]
    j = {"domain": "full" if full else "basalt", "domainCells": m.domain_cells, "space": [nx, ny, nz],
         "Q": int(mc["Q"]), "mcs": mcs, "J": float(mc["J"]), "kT": float(mc["kT"]), "update": mc["update"]}
    if mc["update"] == "checkerboard":
        j["colors"] = m.colors
    j.update({"neighborhood": cfg["neighborhood"], "boundary": cfg["boundary"],
              "grainConnectivity": "moore" if moore else "von_neumann",
              "mobility": "cooling" if mobility is not None else "off"})
    if mobility is not None:
        j["mobilityMin"] = float(mc["mobilityMin"])
    j.update({"grainsStart": first["grains"], "grainsEnd": last["grains"],
              "meanGrainSizeStart": _rounded(first["meanGrainSize"], 3),
              "meanGrainSizeEnd": _rounded(last["meanGrainSize"], 3),
              "largestGrainEnd": end_row["largest"], "energyStart": first["energy"], "energyEnd": last["energy"],
              "attempts": attempts, "accepted": accepted, "acceptedUphill": uphill})
    if cell_mobility is not None:
        cs, ce = _classify(start_root, cell_mobility), _classify(end_root, cell_mobility)
        hi_cells = int(((start_root != 0xFFFFFFFF) & (cell_mobility.astype(np.float64) >= 0.5)).sum())

        def cls(cells, a, b):
            mean = lambda c: c["grainCells"] / c["grains"] if c["grains"] > 0 else 0.0
            return {"cells": cells, "grainsStart": a["grains"], "grainsEnd": b["grains"],
                    "meanGrainSizeStart": _rounded(mean(a), 3), "meanGrainSizeEnd": _rounded(mean(b), 3),
                    "largestGrainEnd": b["largest"]}
        j["byMobility"] = {"threshold": 0.5, "mobilityMin": float(mc["mobilityMin"]),
                           "rule": "ziarno należy do klasy, w której jest większość jego komórek",
                           "high": cls(hi_cells, cs["high"], ce["high"]),
                           "low": cls(m.domain_cells - hi_cells, cs["low"], ce["low"])}
    return {"crystal": m.spins(), "rows": rows, "stats": j}


def run_config(cfg_in: dict, config_dir: str = ".") -> dict:
    """Replays one plain engine run.  Returns file bytes / texts and the stats dict."""
    cfg = merged(cfg_in)
    nx, ny, nz = cfg["space"]["nx"], cfg["space"]["ny"], cfg["space"]["nz"]
    dims = (nx, ny, nz)
    out = {"dims": dims}
    mc_full = cfg["mc"]["enabled"] and cfg["mc"]["domain"] == "full"
    stats = {}
    st = None
    if not mc_full:
        classic = cfg["ca"]["mode"] == "classic"
        hm = spill = None
        if not classic:
            hm = terrain_for(cfg, config_dir)
            periodic = cfg["boundary"] == "periodic"
            spill = fill_depressions(hm, periodic, periodic)
        vents = vents_for(cfg, hm)
        ca = OracleCA(ca_params(cfg), hm, spill, vents)
        ca.run()
        st = ca.state()
        act, chg = ca.histories()
        out.update(state=st, steps=ca.steps, stop=ca.stop_reason, active=act, changed=chg, vents=vents)
        out["ca_result_bin"] = ca_result_bin(dims, st)
        stats = ca_stats(cfg, ca.steps, ca.stop_reason, vents, st, act, chg)
    if cfg["mc"]["enabled"]:
        res = run_mc(cfg, st)
        out["crystal"] = res["crystal"]
        out["mc_result_bin"] = mc_result_bin(dims, res["crystal"])
        out["mc_stats_csv"] = mc_stats_csv(res["rows"])
        stats["mc"] = res["stats"]
    out["stats"] = stats
    return out


def load_stats_json(path: str) -> dict:
    with open(path, encoding="utf-8") as f:
        j = json.load(f)
    j.pop("parallel", None)
    return j
