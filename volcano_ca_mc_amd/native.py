"""ctypes binding of libvca.so (include/vca.h) with a Python mirror of the reference's classes.

``LavaCA`` mirrors ``LavaCA<Storage>`` (engine/src/ca/LavaCA.hpp:80-296): construct with dims, params,
terrain and seeds, then ``run()``/``step()``, ``steps``, ``stop_reason``, histories and ``result()``.
``PottsMC`` mirrors ``PottsMC`` (engine/src/mc/PottsMC.hpp:52-268): ``sweep()``, ``measure()``,
``crystals()``.  All state lives on the GPU; the library raises ``RuntimeError`` with the library's
message on any failure and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libvca.so")

NEIGHBORHOODS = {"von_neumann": 0, "moore": 1, "pentagonal": 2, "hexagonal": 3}
STOP_REASONS = {0: "running", 1: "no_lava", 2: "no_air", 3: "no_change", 4: "max_steps"}


class CaDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("classic", C.c_int32),
                ("neighborhood", C.c_int32), ("periodic", C.c_int32), ("seed", C.c_uint64),
                ("max_steps", C.c_int32), ("eruption_steps", C.c_int32), ("erupt_temp", C.c_float),
                ("solidus_temp", C.c_float), ("cooling_ambient", C.c_float), ("cooling_lateral", C.c_float),
                ("cooling_fall", C.c_float), ("ponding", C.c_int32), ("dense", C.c_int32),
                ("tile_shape", C.c_int32)]


class Vent(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32), ("flow_id", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("tile_shape", C.c_int32), ("dense", C.c_int32), ("candidates", C.c_int32),
                ("probe_steps", C.c_int32), ("ms_per_step", (C.c_float * 5) * 2)]


class McDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("neighborhood", C.c_int32),
                ("periodic", C.c_int32 * 3), ("seed", C.c_uint64), ("Q", C.c_int32), ("J", C.c_double),
                ("kT", C.c_double), ("checkerboard", C.c_int32)]


class McCounts(C.Structure):
    _fields_ = [("attempts", C.c_uint64), ("accepted", C.c_uint64), ("uphill", C.c_uint64), ("sweeps", C.c_int32)]


class McStats(C.Structure):
    _fields_ = [("mcs", C.c_int32), ("grains", C.c_uint64), ("cells", C.c_uint64), ("largest", C.c_uint64),
                ("mean_grain_size", C.c_double), ("energy", C.c_double)]


class McClass(C.Structure):
    _fields_ = [("grains", C.c_uint64), ("grain_cells", C.c_uint64), ("largest", C.c_uint64)]


class KernelTime(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    """Loads libvca.so; raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m volcano_ca_mc_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, u64 = C.c_void_p, C.c_int32, C.c_uint64
    sig = {
        "vca_version": [], "vca_last_error": [C.c_char_p, C.c_size_t],
        "vca_ctx_create": [C.c_int, vp, C.POINTER(vp)], "vca_ctx_destroy": [vp], "vca_ctx_synchronize": [vp],
        "vca_device_memory": [vp, C.POINTER(u64), C.POINTER(u64)],
        "vca_ctx_set_timing": [vp, i32], "vca_ctx_kernel_time": [vp, C.c_char_p, C.POINTER(KernelTime)],
        "vca_ca_create": [vp, C.POINTER(CaDesc), vp, vp, vp, i32, C.POINTER(vp)], "vca_ca_destroy": [vp],
        "vca_ca_run": [vp, i32, C.POINTER(i32), C.POINTER(i32)], "vca_ca_enqueue": [vp, i32],
        "vca_ca_run_noalloc": [vp, i32, C.POINTER(i32), C.POINTER(i32)],
        "vca_ca_status": [vp, C.POINTER(i32), C.POINTER(i32)],
        "vca_ca_set_fast_forward": [vp, i32], "vca_ca_fast_forwarded": [vp, C.POINTER(i32), C.POINTER(i32)],
        "vca_ca_histories": [vp, vp, vp, i32, C.POINTER(i32)], "vca_ca_download": [vp, vp, vp, vp, vp, vp],
        "vca_ca_work": [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32)],
        "vca_ca_flow_stats": [vp, vp, i32, vp, vp, vp], "vca_ca_flow_points": [vp, vp, i32, u64, vp, i32],
        "vca_ca_autotune": [vp, C.POINTER(CaDesc), vp, vp, vp, i32, i32, C.POINTER(TuneResult)],
        "vca_mc_create": [vp, C.POINTER(McDesc), vp, vp, i32, C.POINTER(vp)],
        "vca_mc_create_from_ca": [vp, C.POINTER(McDesc), vp, i32, C.c_double, C.POINTER(vp)],
        "vca_mc_destroy": [vp], "vca_mc_sweep": [vp, i32, C.POINTER(McCounts)], "vca_mc_set_sweep_mode": [vp, i32],
        "vca_mc_counts_get": [vp, C.POINTER(McCounts)],
        "vca_mc_measure": [vp, C.POINTER(McStats), C.POINTER(McClass), C.POINTER(u64)],
        "vca_mc_download": [vp, vp], "vca_mc_download_labels": [vp, vp],
        "vca_mc_info": [vp, C.POINTER(i32), C.POINTER(u64)], "vca_mc_download_mobility": [vp, vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


EXPORTED = ("vca_version", "vca_last_error", "vca_ctx_create", "vca_ctx_destroy", "vca_ctx_synchronize",
            "vca_device_memory", "vca_ctx_set_timing", "vca_ctx_kernel_time", "vca_ca_create", "vca_ca_destroy",
            "vca_ca_run", "vca_ca_run_noalloc", "vca_ca_enqueue", "vca_ca_status", "vca_ca_histories", "vca_ca_download", "vca_ca_work",
            "vca_ca_set_fast_forward", "vca_ca_fast_forwarded",
            "vca_ca_footprint", "vca_ca_tiles", "vca_batch_create", "vca_batch_destroy", "vca_batch_add",
            "vca_batch_remove", "vca_batch_set_graph", "vca_batch_enqueue", "vca_batch_run", "vca_batch_info",
            "vca_batch_status", "vca_ca_collect", "vca_ca_flow_stats", "vca_ca_flow_points", "vca_ca_autotune",
            "vca_mc_create", "vca_mc_create_from_ca", "vca_mc_destroy", "vca_mc_sweep", "vca_mc_set_sweep_mode",
            "vca_mc_counts_get",
            "vca_mc_measure", "vca_mc_download", "vca_mc_download_labels", "vca_mc_info",
            "vca_mc_download_mobility")


def _check(rc: int) -> None:
    if rc != 0:
        buf = C.create_string_buffer(1024)
        lib().vca_last_error(buf, 1024)
        raise RuntimeError(f"libvca error {rc}: {buf.value.decode('utf-8', 'replace')}")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Context:
    """One device and one HIP stream (``stream``: optional hipStream_t handle as int)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self.h = C.c_void_p()
        _check(lib().vca_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vca_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().vca_ctx_synchronize(self.h))

    def memory(self):
        f, t = C.c_uint64(), C.c_uint64()
        _check(lib().vca_device_memory(self.h, C.byref(f), C.byref(t)))
        return int(f.value), int(t.value)

    def set_timing(self, on: bool):
        _check(lib().vca_ctx_set_timing(self.h, int(on)))

    def kernel_time(self, name: str):
        k = KernelTime()
        _check(lib().vca_ctx_kernel_time(self.h, name.encode(), C.byref(k)))
        return float(k.ms), int(k.launches)


@dataclass
class CaParams:
    """CaParams (engine/src/ca/CaRules.hpp:93-105), temperatures already narrowed to float32."""
    seed: int = 42
    classic: bool = False
    max_steps: int = 5000
    eruption_steps: int = 300
    erupt_temp: float = 1200.0
    solidus_temp: float = 800.0
    cooling_ambient: float = 0.3
    cooling_lateral: float = 1.5
    cooling_fall: float = 0.1
    ponding: bool = True
    dense: bool = False
    tile_shape: int = 0  # 0 default; 1..5: 64x4x4, 64x2x8, 64x8x2, 64x1x16, 64x16x1


def _ca_desc(dims, params: CaParams, neighborhood: str, periodic: bool) -> CaDesc:
    nx, ny, nz = dims
    return CaDesc(nx, ny, nz, int(params.classic), NEIGHBORHOODS[neighborhood], int(periodic), params.seed,
                  params.max_steps, params.eruption_steps, params.erupt_temp, params.solidus_temp,
                  params.cooling_ambient, params.cooling_lateral, params.cooling_fall, int(params.ponding),
                  int(params.dense), int(params.tile_shape))


def ca_footprint(dims, params: CaParams, neighborhood: str = "moore", periodic: bool = False) -> int:
    """Device bytes a simulation of these dims takes (vca_ca_footprint)."""
    d = _ca_desc(dims, params, neighborhood, periodic)
    b = C.c_uint64()
    _check(lib().vca_ca_footprint(C.byref(d), C.byref(b)))
    return int(b.value)


def ca_autotune(ctx: "Context", dims, params: CaParams, neighborhood: str, periodic: bool,
                height: Optional[np.ndarray], spill: Optional[np.ndarray], vents: Sequence[Sequence[int]],
                probe_steps: int = 64) -> dict:
    """vca_ca_autotune: the fastest (tile_shape, dense) over probe_steps steps of this simulation, and the
    measured ms per step of every candidate, keyed (dense, tile_shape)."""
    d = _ca_desc(dims, params, neighborhood, periodic)
    v = (Vent * max(1, len(vents)))(*[Vent(*map(int, x)) for x in vents])
    h = None if height is None else np.ascontiguousarray(height, dtype=np.int32)
    s = None if spill is None else np.ascontiguousarray(spill, dtype=np.int32)
    r = TuneResult()
    _check(lib().vca_ca_autotune(ctx.h, C.byref(d), _ptr(h), _ptr(s), C.cast(v, C.c_void_p), len(vents), probe_steps,
                                 C.byref(r)))
    ms = {(dn, sh + 1): float(r.ms_per_step[dn][sh]) for dn in range(2) for sh in range(5) if r.ms_per_step[dn][sh] > 0}
    return {"tile_shape": int(r.tile_shape), "dense": int(r.dense), "candidates": int(r.candidates), "ms": ms}


class LavaCA:
    def __init__(self, ctx: Context, dims: Sequence[int], params: CaParams, neighborhood: str, periodic: bool,
                 height: Optional[np.ndarray], spill: Optional[np.ndarray], vents: Sequence[Sequence[int]]):
        nx, ny, nz = dims
        self.ctx, self.dims, self.n = ctx, (nx, ny, nz), nx * ny * nz
        d = _ca_desc(dims, params, neighborhood, periodic)
        v = (Vent * max(1, len(vents)))(*[Vent(*map(int, x)) for x in vents])
        h = None if height is None else np.ascontiguousarray(height, dtype=np.int32)
        s = None if spill is None else np.ascontiguousarray(spill, dtype=np.int32)
        self.h = C.c_void_p()
        _check(lib().vca_ca_create(ctx.h, C.byref(d), _ptr(h), _ptr(s), C.cast(v, C.c_void_p), len(vents),
                                   C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vca_ca_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_new_steps: int = 1 << 30, noalloc: bool = False):
        """LavaCA::run: steps until a stop rule fires or max_new_steps are done.  noalloc: through vca_ca_run_noalloc
        (no device allocation; its lists in the library's static device array)."""
        steps, stop = C.c_int32(), C.c_int32()
        fn = lib().vca_ca_run_noalloc if noalloc else lib().vca_ca_run
        _check(fn(self.h, min(max_new_steps, 2 ** 31 - 1), C.byref(steps), C.byref(stop)))
        return int(steps.value), int(stop.value)

    def step(self) -> bool:
        return self.run(1)[1] == 0

    def set_fast_forward(self, on: bool) -> None:
        """Quiescent fast-forward of run() on (the default) or off (every step launched); results are the same."""
        _check(lib().vca_ca_set_fast_forward(self.h, 1 if on else 0))

    def fast_forwarded(self):
        """(step the fast-forward started from or -1, steps it completed)."""
        f, n = C.c_int32(), C.c_int32()
        _check(lib().vca_ca_fast_forwarded(self.h, C.byref(f), C.byref(n)))
        return int(f.value), int(n.value)

    def enqueue(self, n: int):
        _check(lib().vca_ca_enqueue(self.h, n))

    def status(self):
        steps, stop = C.c_int32(), C.c_int32()
        _check(lib().vca_ca_status(self.h, C.byref(steps), C.byref(stop)))
        return int(steps.value), int(stop.value)

    @property
    def steps(self) -> int:
        return self.status()[0]

    @property
    def stop_reason(self) -> int:
        return self.status()[1]

    def histories(self):
        steps = self.steps
        a = np.zeros(max(1, steps), np.uint64)
        c = np.zeros(max(1, steps), np.uint64)
        n = C.c_int32()
        _check(lib().vca_ca_histories(self.h, _ptr(a), _ptr(c), steps, C.byref(n)))
        return a[:n.value].tolist(), c[:n.value].tolist()

    def result(self):
        n = self.n
        t = np.empty(n, np.uint8); f = np.empty(n, np.int32); tp = np.empty(n, np.float32)
        s = np.empty(n, np.uint16); a = np.empty(n, np.uint16)
        _check(lib().vca_ca_download(self.h, _ptr(t), _ptr(f), _ptr(tp), _ptr(s), _ptr(a)))
        return {"type": t, "flow": f, "temp": tp, "solid": s, "arrival": a}

    def tiles(self) -> int:
        t = C.c_uint64()
        _check(lib().vca_ca_tiles(self.h, C.byref(t)))
        return int(t.value)

    def work(self):
        tv, ch, tc = C.c_uint64(), C.c_uint64(), C.c_int32()
        _check(lib().vca_ca_work(self.h, C.byref(tv), C.byref(ch), C.byref(tc)))
        return int(tv.value), int(ch.value), int(tc.value)

    def flow_stats(self, vents: Sequence[Sequence[int]]):
        """Per flowId 1..len(vents): (lava cells, basalt cells, largest squared vent distance), on the device."""
        m = len(vents) + 1
        v = (Vent * max(1, len(vents)))(*[Vent(*map(int, x)) for x in vents])
        la, ba, d2 = (np.zeros(m, np.uint64) for _ in range(3))
        _check(lib().vca_ca_flow_stats(self.h, C.cast(v, C.c_void_p), len(vents), _ptr(la), _ptr(ba), _ptr(d2)))
        return la, ba, d2

    def flow_points(self, vent: Sequence[int], flow: int, d2: int, cap: int) -> np.ndarray:
        v = Vent(*map(int, vent))
        seen = np.zeros(cap, np.uint8)
        _check(lib().vca_ca_flow_points(self.h, C.byref(v), flow, d2, _ptr(seen), cap))
        return seen


class Batch:
    """Simulations advanced together, one launch per kernel and step for all members (vca_batch_*)."""

    def __init__(self, ctx: Context, max_sims: int, max_tiles: int, steps_per_graph: int = 0):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().vca_batch_create(ctx.h, max_sims, max_tiles, C.byref(self.h)))
        if steps_per_graph:
            self.set_graph(steps_per_graph)

    def close(self):
        if self.h:
            lib().vca_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, ca: LavaCA) -> int:
        slot = C.c_int32()
        _check(lib().vca_batch_add(self.h, ca.h, C.byref(slot)))
        return int(slot.value)

    def status(self, cap: int):
        """[(steps, stop)] per slot; stop -1 = empty slot."""
        out = (C.c_int32 * (2 * cap))()
        _check(lib().vca_batch_status(self.h, out, cap))
        return [(out[2 * k], out[2 * k + 1]) for k in range(cap)]

    def remove(self, ca: LavaCA):
        _check(lib().vca_batch_remove(self.h, ca.h))

    def set_graph(self, steps_per_graph: int):
        _check(lib().vca_batch_set_graph(self.h, steps_per_graph))

    def enqueue(self, steps: int):
        _check(lib().vca_batch_enqueue(self.h, steps))

    def run(self, max_steps: int = 1 << 30) -> int:
        running = C.c_int32()
        _check(lib().vca_batch_run(self.h, min(max_steps, 2 ** 31 - 1), C.byref(running)))
        return int(running.value)

    def info(self):
        m, t, r = C.c_int32(), C.c_uint64(), C.c_int32()
        _check(lib().vca_batch_info(self.h, C.byref(m), C.byref(t), C.byref(r)))
        return int(m.value), int(t.value), int(r.value)


class PottsMC:
    def __init__(self, ctx: Context, dims: Sequence[int], neighborhood: str, periodic: Sequence[bool], seed: int,
                 Q: int, J: float, kT: float, checkerboard: bool = True, domain: Optional[np.ndarray] = None,
                 mobility: Optional[np.ndarray] = None, mobility_dynamics: bool = True, _handle=None):
        nx, ny, nz = dims
        self.ctx, self.dims, self.n = ctx, (nx, ny, nz), nx * ny * nz
        self.J = J
        self.desc = McDesc(nx, ny, nz, NEIGHBORHOODS[neighborhood], (C.c_int32 * 3)(*[int(bool(p)) for p in periodic]),
                           seed, Q, J, kT, int(checkerboard))
        self.has_mobility = mobility is not None
        if _handle is not None:
            self.h = _handle
            return
        dom = None if domain is None else np.ascontiguousarray(domain, dtype=np.uint8)
        mob = None if mobility is None else np.ascontiguousarray(mobility, dtype=np.float32)
        self.h = C.c_void_p()
        _check(lib().vca_mc_create(ctx.h, C.byref(self.desc), _ptr(dom), _ptr(mob), int(mobility_dynamics),
                                   C.byref(self.h)))

    @classmethod
    def from_ca(cls, ctx: Context, ca: LavaCA, neighborhood: str, periodic, seed, Q, J, kT, checkerboard=True,
                use_mobility=True, mobility_min=0.05):
        obj = cls(ctx, ca.dims, neighborhood, periodic, seed, Q, J, kT, checkerboard, _handle=C.c_void_p())
        _check(lib().vca_mc_create_from_ca(ctx.h, C.byref(obj.desc), ca.h, int(use_mobility), mobility_min,
                                           C.byref(obj.h)))
        obj.has_mobility = True
        return obj

    def close(self):
        if self.h:
            lib().vca_mc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # vca_mc_set_sweep_mode values
    SWEEP_AUTO, SWEEP_PER_COLOUR, SWEEP_PERSISTENT = 0, 1, 2

    def set_sweep_mode(self, mode: int) -> None:
        """How checkerboard sweeps are launched (same results): auto, one launch per colour pass, or one persistent
        launch per call."""
        _check(lib().vca_mc_set_sweep_mode(self.h, mode))

    def sweep(self, n: int = 1):
        c = McCounts()
        _check(lib().vca_mc_sweep(self.h, n, C.byref(c)))
        return int(c.attempts), int(c.accepted), int(c.uphill)

    def counters(self):
        c = McCounts()
        _check(lib().vca_mc_counts_get(self.h, C.byref(c)))
        return int(c.attempts), int(c.accepted), int(c.uphill)

    def measure(self):
        st = McStats()
        cls = (McClass * 2)()
        hi = C.c_uint64()
        _check(lib().vca_mc_measure(self.h, C.byref(st), cls, C.byref(hi)))
        out = {"mcs": st.mcs, "grains": int(st.grains), "cells": int(st.cells), "largest": int(st.largest),
               "mean_grain_size": st.mean_grain_size, "energy": st.energy}
        if self.has_mobility:
            out["classes"] = [{"grains": int(c.grains), "grain_cells": int(c.grain_cells), "largest": int(c.largest)}
                              for c in cls]
            out["high_cells"] = int(hi.value)
        return out

    def crystals(self) -> np.ndarray:
        out = np.empty(self.n, np.int32)
        _check(lib().vca_mc_download(self.h, _ptr(out)))
        return out

    def labels(self) -> np.ndarray:
        out = np.empty(self.n, np.uint32)
        _check(lib().vca_mc_download_labels(self.h, _ptr(out)))
        return out

    def info(self):
        c, d = C.c_int32(), C.c_uint64()
        _check(lib().vca_mc_info(self.h, C.byref(c), C.byref(d)))
        return int(c.value), int(d.value)
