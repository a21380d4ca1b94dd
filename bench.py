#!/usr/bin/env python3
"""Benchmark of the GPU engine on the 3D cone workload (one MI355X).

Workload: 512 x 512 x 256 voxels, radial cone 220 -> 20 at r = 250 plus 4-octave
value noise (amplitude 12), three craters within r = 30 of the summit, Moore neighbourhood, absorbing
boundaries, up to 1000 lava-CA steps, then 200 checkerboard Potts sweeps (Q 64, kT 0.5) on the basalt.
Synthetic, seeded terrain (volcano_ca_mc_amd/workloads.py).

One bench "step" = one complete job of that workload through the drop-in `engine` binary (C++ host +
HIP kernels): the CA runs to its stop rule, then the MC sweeps.  The timed regions are the engine's
caSim (all CA steps, state resident in HBM) and mcSim (all sweeps); terrain, allocation, MC set-up,
statistics and file output are outside them (output.format = none).

Printed: one JSON line.  value = CA cell-updates/s = cells x executed steps / caSim (the reference's
"ns per cell and step" unit); mc_value = MC spin-attempts/s = domain cells x sweeps / mcSim.
roofline: the dominant kernel's algorithmic bytes per launch (26 B per processed cell-step for the CA
kernels) / its mean launch time from HIP events recorded on the engine's stream during the timed runs (on every
10th CA step: events on every launch would add their own cost to caSim);
traffic = measured HBM bytes per launch of that kernel from the committed rocprofv3 counter summary;
valu_issue_frac / salu_issue_frac = its VALU / SALU wave-instructions per launch (committed rocprofv3 counter
summary) / the live mean launch time / the issue peak (256 CUs x 2.4 GHz).  mc_roofline: the same for mc_color,
at 2 * colours + 4 B per spin attempt.
cpu_baseline: the reference engine (oracle/_ref/engine, built from the reference sources) on the host
cores, on a bounded sample (the full CA phase, the first 20 sweeps) of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ENGINE = os.path.join(ROOT, "volcano_ca_mc_amd", "_lib", "engine")
REF_ENGINE = os.path.join(ROOT, "oracle", "_ref", "engine")
HBM_PEAK_GBS = 8000.0
TIMING_SAMPLE = 10  # kernel events on every 10th CA step and MC sweep
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
PROFILE_ISSUE = os.path.join(ROOT, "profiles", "valu.json")
# wave-instruction issue peak: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU instruction
ISSUE_PEAK_PER_S = 256 * 2.4e9


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_engine(exe: str, cfg_path: str, out: str, extra, timeout: int = 1800) -> dict:
    t0 = time.time()
    r = subprocess.run([exe, "--config", cfg_path, "--out", out, *extra], capture_output=True, text=True,
                       timeout=timeout, cwd=os.path.dirname(cfg_path))
    wall = time.time() - t0
    if r.returncode != 0:
        raise RuntimeError(f"{exe} failed ({r.returncode}): {r.stderr[-2000:]}")
    with open(os.path.join(out, "times.json")) as f:
        times = json.load(f)
    times["_wall_s"] = wall
    return times


# Algorithmic bytes of one launch of each kernel, from the engine's work counters.  The CA figure is 26 B per
# processed cell-step (13 B of state read + 13 B written: u8 type, i32 flowId, f32 temperature, u16 solid and
# arrival steps), processed = visited tiles x 1024 cells; neighbour reads count once.  The MC figure is
# 2 * colours + 4 B per trial (spin read once per colour pass, class read, spin written).
CA_BYTES_PER_CELL = 26


def kernel_bytes(times: dict) -> dict:
    d = times.get("detail", {})
    w = d.get("caWork", {})
    k = d.get("kernels", {})
    out = {}
    steps = max(1, times.get("steps", 1))
    visited_cells = w.get("tilesVisited", 0) * w.get("tileCells", 0)
    if "ca_rules" in k:
        out["ca_rules"] = CA_BYTES_PER_CELL * visited_cells / steps
    if "mc_color" in k:
        colours = max(1, k["mc_color"]["launches"] // max(1, times.get("mcs", 1)))
        dom = times.get("domainCells", 0)
        out["mc_color"] = dom * (2 * colours + 4) * times.get("mcs", 0) / max(1, k["mc_color"]["launches"])
    return out


def minimal_bytes(times: dict) -> float:
    """Lower bound of what one ca_rules launch must move in this in-place design: the type byte of every
    visited cell, flowId + temperature of every lava cell, one 16 B record per change."""
    w = times.get("detail", {}).get("caWork", {})
    steps = max(1, times.get("steps", 1))
    return (w.get("tilesVisited", 0) * w.get("tileCells", 0) + 8 * w.get("lavaCellSteps", 0)
            + 16 * w.get("changes", 0)) / steps


def cpu_baseline(cfg_path: str, cells: int, tmp: str) -> dict | None:
    if not os.path.exists(REF_ENGINE):
        return None
    threads = min(16, os.cpu_count() or 1)
    out = os.path.join(tmp, "cpu")
    log(f"cpu baseline: reference engine, {threads} threads, full CA phase + first 20 sweeps")
    t = run_engine(REF_ENGINE, cfg_path, out, ["--threads", str(threads), "--set", "mc.mcs=20",
                                               "--set", "output.format=none"])
    ca_rate = cells * t["steps"] / (t["caSim"] / 1000.0)
    mc_rate = t["domainCells"] * t["mcs"] / (t["mcSim"] / 1000.0) if t.get("mcSim") else None
    return {"value": ca_rate, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
            "mc_value": mc_rate, "mc_unit": "spin-attempts/s",
            "sample": f"reference engine (oracle/_ref, OpenMP) on the same workload: all {t['steps']} CA steps and the "
                      f"first {t['mcs']} Potts sweeps ({t['domainCells']} domain cells); caSim {t['caSim']:.0f} ms, "
                      f"mcSim {t['mcSim']:.0f} ms"}


def _write_jobs(workloads, folder: str) -> None:
    os.makedirs(folder, exist_ok=True)
    for wl in workloads:
        wl.write(folder)


def _pipeline(args, out: str) -> dict:
    t0 = time.time()
    # the engine's per-job progress lines go to our stderr (the run's log)
    r = subprocess.run([ENGINE, *args, "--out", out, "--set", "output.format=none"], stdout=sys.stderr,
                       stderr=sys.stderr, timeout=3600)
    if r.returncode != 0:
        raise RuntimeError(f"engine {' '.join(args)} failed ({r.returncode})")
    name = "serve_summary.json" if args[0] == "--serve" else "pipeline_summary.json"
    with open(os.path.join(out, name)) as f:
        s = json.load(f)
    s["_wall_s"] = time.time() - t0
    return s


def _same_stats(a: str, b: str, names) -> bool:
    for n in names:
        with open(os.path.join(a, n, "stats.json"), encoding="utf-8") as f:
            x = json.load(f)
        with open(os.path.join(b, n, "stats.json"), encoding="utf-8") as f:
            y = json.load(f)
        if x != y:
            return False
    return True


def bench_batched(a) -> dict:
    """256 uneven jobs: the batched pipeline against the same jobs one at a time (--max-jobs 1), same process
    and engine; stats.json of every job must be equal between the two."""
    from volcano_ca_mc_amd import workloads as W
    jobs = W.batched_mix(jobs=a.jobs)
    with tempfile.TemporaryDirectory(prefix="vca_batched_") as tmp:
        folder = os.path.join(tmp, "jobs")
        _write_jobs(jobs, folder)
        log(f"{len(jobs)} jobs written")
        one = _pipeline(["--batch", folder, "--gpu-batch", "--max-jobs", "1"], os.path.join(tmp, "one"))
        log(f"one at a time: {one['makespanS']:.2f} s, {one['jobsPerS']:.2f} jobs/s")
        many = _pipeline(["--batch", folder, "--gpu-batch"], os.path.join(tmp, "many"))
        log(f"batched: {many['makespanS']:.2f} s, {many['jobsPerS']:.2f} jobs/s")
        same = _same_stats(os.path.join(tmp, "one"), os.path.join(tmp, "many"), [w.name for w in jobs])
    return {"metric": "jobs/s (batched uneven simulations, bit-exact)", "value": many["jobsPerS"], "unit": "jobs/s",
            "n_gpus": 1, "higher_is_better": True, "vs_one_at_a_time": many["jobsPerS"] / max(1e-9, one["jobsPerS"]),
            "one_at_a_time_jobs_per_s": one["jobsPerS"], "makespan_s": many["makespanS"],
            "one_at_a_time_makespan_s": one["makespanS"], "cell_updates_per_s": many["cellUpdatesPerS"],
            "p50_job_ms": many["p50LatencyMs"], "p99_job_ms": many["p99LatencyMs"], "identical_stats": same,
            "gpu_thread_ms": many.get("gpuThread"), "one_at_a_time_gpu_thread_ms": one.get("gpuThread"),
            "dtype": "u8+i32+f32", "data": "synthetic (seeded terrains)",
            "config": {"workload": f"{len(jobs)} jobs from one seed: 2D edges 128-2048, 3D edges 64-256, 200-800 steps, "
                                   "50 Potts sweeps Q64", "parallelism": "single GPU"}}


def bench_serving(a) -> dict:
    from volcano_ca_mc_amd import workloads as W
    arr = W.serving_stream(rate=a.rate, duration=a.duration, scale=a.scale)
    with tempfile.TemporaryDirectory(prefix="vca_serve_") as tmp:
        folder = os.path.join(tmp, "jobs")
        _write_jobs([x.workload for x in arr], folder)
        with open(os.path.join(tmp, "jobs.jsonl"), "w") as f:
            for x in arr:
                f.write(json.dumps({"at": x.t, "config": os.path.join(folder, x.workload.name + ".json"),
                                    "name": x.workload.name}) + "\n")
        log(f"{len(arr)} requests over {a.duration} s written")
        s = _pipeline(["--serve", os.path.join(tmp, "jobs.jsonl")], os.path.join(tmp, "out"))
    return {"metric": "served jobs/s with p50/p99 latency (bit-exact)", "value": s["jobsPerS"], "unit": "jobs/s",
            "n_gpus": 1, "higher_is_better": True, "p50_job_ms": s["p50LatencyMs"], "p99_job_ms": s["p99LatencyMs"],
            "offered_rate": a.rate, "makespan_s": s["makespanS"], "cell_updates_per_s": s["cellUpdatesPerS"],
            "spin_attempts_per_s": s["spinAttemptsPerS"], "jobs": s["jobs"], "ok": s["ok"], "dtype": "u8+i32+f32",
            "gpu_thread_ms": s.get("gpuThread"),
            "data": "synthetic (seeded terrains)",
            "config": {"workload": f"Poisson {a.rate}/s for {a.duration} s: 70% 2D 512x256 x300 steps, 25% 3D 256^3 x600, "
                                   f"5% 768x768x256 x1200, 50-200 sweeps; grid scale {a.scale}",
                       "parallelism": "single GPU"}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["cone", "batched", "serving"], default="cone",
                    help="cone: the 3D cone job (default line); batched: 256 uneven jobs; serving: Poisson stream")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed jobs")
    ap.add_argument("--warmup", type=int, default=1, help="untimed jobs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="grid scale (1.0 = the benchmark workload)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": a.gpus}))
        return 2
    if not os.path.exists(ENGINE):
        raise SystemExit(f"{ENGINE} missing: run `python -m volcano_ca_mc_amd.build` first")
    if a.workload == "batched":
        print(json.dumps(bench_batched(a)))
        return 0
    if a.workload == "serving":
        print(json.dumps(bench_serving(a)))
        return 0
    from volcano_ca_mc_amd import workloads as W
    n = max(16, int(round(512 * a.scale)))
    wl = W.cone_3d(n=n, nz=max(16, int(round(256 * a.scale))))
    nx, ny, nz = wl.dims
    cells = nx * ny * nz
    with tempfile.TemporaryDirectory(prefix="vca_bench_") as tmp:
        cfg_path = wl.write(tmp)
        # kernel times from HIP events on every TIMING_SAMPLE-th CA step / MC sweep of the timed jobs: timing every
        # launch would add its own events' cost to the measured caSim
        extra = ["--timing-sample", str(TIMING_SAMPLE), "--set", "output.format=none"]
        runs = []
        for i in range(a.warmup + a.steps):
            t = run_engine(ENGINE, cfg_path, os.path.join(tmp, f"run{i}"), extra)
            tag = "warmup" if i < a.warmup else "timed"
            log(f"{tag} job {i}: steps {t['steps']} caSim {t['caSim']:.1f} ms mcSim {t['mcSim']:.1f} ms "
                f"total {t['total']:.1f} ms")
            if i >= a.warmup:
                runs.append(t)
        ca_ms = [t["caSim"] for t in runs]
        mc_ms = [t["mcSim"] for t in runs]
        steps = runs[0]["steps"]
        dom = runs[0]["domainCells"]
        mcs = runs[0]["mcs"]
        ca_rate = cells * steps / (statistics.median(ca_ms) / 1000.0)
        mc_rate = dom * mcs / (statistics.median(mc_ms) / 1000.0)
        lat = sorted(t["total"] for t in runs)
        p50 = statistics.median(lat)
        p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]
        # roofline of the dominant kernel (largest summed device time across the timed jobs)
        ktot = {}
        for t in runs:
            for name, v in t.get("detail", {}).get("kernels", {}).items():
                if name in ("ca_step", "mc_sweep"):
                    continue
                e = ktot.setdefault(name, [0.0, 0])
                e[0] += v["ms"]
                e[1] += v["launches"]
        bytes_per_launch = kernel_bytes(runs[0])
        dom_k = max((k for k in ktot if k in bytes_per_launch), key=lambda k: ktot[k][0])
        mean_ms = ktot[dom_k][0] / ktot[dom_k][1]
        achieved = bytes_per_launch[dom_k] / (mean_ms / 1000.0) / 1e9
        traffic = None
        if os.path.exists(PROFILE_TRAFFIC):
            with open(PROFILE_TRAFFIC) as f:
                traffic = json.load(f).get(dom_k)
        issue = {}
        if os.path.exists(PROFILE_ISSUE):
            with open(PROFILE_ISSUE) as f:
                issue = json.load(f)

        def issue_fracs(k: str, ms: float) -> dict:
            e = issue.get(k)
            if not e:
                return {"valu_issue_frac": None, "salu_issue_frac": None}
            return {"valu_issue_frac": round(e["valu"] / (ms / 1000.0) / ISSUE_PEAK_PER_S, 4),
                    "salu_issue_frac": round(e["salu"] / (ms / 1000.0) / ISSUE_PEAK_PER_S, 4)}

        mc_roofline = None
        if "mc_color" in ktot and "mc_color" in bytes_per_launch:
            mc_launch_ms = ktot["mc_color"][0] / ktot["mc_color"][1]
            mc_ach = bytes_per_launch["mc_color"] / (mc_launch_ms / 1000.0) / 1e9
            mc_traffic = None
            if os.path.exists(PROFILE_TRAFFIC):
                with open(PROFILE_TRAFFIC) as f:
                    mc_traffic = json.load(f).get("mc_color")
            mc_roofline = {"bound": "latency", "kernel": "mc_color", "kernel_mean_ms": round(mc_launch_ms, 5),
                           "achieved": round(mc_ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(mc_ach / HBM_PEAK_GBS, 4), "traffic": mc_traffic,
                           "bytes_per_unit": "2 * colours + 4 B per spin attempt",
                           "algorithmic_bytes_per_launch": round(bytes_per_launch["mc_color"]),
                           **issue_fracs("mc_color", mc_launch_ms)}
        w = runs[0].get("detail", {}).get("caWork", {})
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_k,
                    "kernel_mean_ms": round(mean_ms, 5), "timed_every_nth_step": TIMING_SAMPLE, "algorithmic_bytes_per_launch": round(bytes_per_launch[dom_k]),
                    "bytes_per_unit": f"{CA_BYTES_PER_CELL} B per processed cell-step" if dom_k == "ca_rules" else None,
                    "processed_cells_per_launch": round(w.get("tilesVisited", 0) * w.get("tileCells", 0) /
                                                        max(1, steps)),
                    "staged_cells_per_launch": round(w.get("tilesStaged", 0) * w.get("tileCells", 0) /
                                                     max(1, steps)),
                    "minimal_bytes_per_launch": round(minimal_bytes(runs[0])),
                    **issue_fracs(dom_k, mean_ms),
                    "kernel_share": {k: round(v[0] / sum(x[0] for x in ktot.values()), 3) for k, v in ktot.items()}}
        cpu = None if a.no_cpu else cpu_baseline(cfg_path, cells, tmp)
    result = {
        "metric": "CA cell-updates/s + MC spin-attempts/s (bit-exact); p50/p99 job latency",
        "value": ca_rate,
        "unit": "cell-updates/s",
        "mc_value": mc_rate,
        "mc_unit": "spin-attempts/s",
        "p50_job_ms": p50,
        "p99_job_ms": p99,
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": statistics.median(ca_ms) + statistics.median(mc_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8+i32+f32",
        "data": "synthetic (seeded terrain)",
        "config": {"workload": f"cone_3d {nx}x{ny}x{nz} moore absorbing, 3 craters, maxSteps 1000, "
                               f"Potts Q64 kT0.5 200 sweeps (checkerboard)", "cells": cells, "ca_steps": steps,
                   "mc_domain_cells": dom, "mc_sweeps": mcs, "parallelism": "single GPU"},
        "roofline": roofline,
        "mc_roofline": mc_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
