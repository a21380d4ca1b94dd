#!/usr/bin/env python3
"""Benchmark of the GPU engine (one MI355X).

Default workload: the 3D caldera, the largest single-GPU configuration -- 1024 x 1024 x 384 voxels, a double
cone with an inner ring depression (depth 60, radius 150) plus 6-octave value noise (amplitude 20), eight
craters on the ring, random pentagonal neighbourhood drawn by the hash per cell, periodic x/y and absorbing z,
up to 2000 lava-CA steps, then 100 checkerboard Potts sweeps (Q 128, kT 0.5) on the basalt.  Synthetic,
seeded terrain (volcano_ca_mc_amd/workloads.py).

One bench "step" = one complete job of that workload through the drop-in `engine` binary (C++ host + HIP
kernels): the CA runs to its stop rule, then the MC sweeps.  The timed regions are the engine's caSim (all CA
steps, state resident in HBM) and mcSim (all sweeps); terrain, allocation, MC set-up, statistics and file
output are outside them (output.format = none).

Printed: one JSON line.
  value      CA cell-updates/s = cells x executed steps / caSim (the reference's "ns per cell and step" unit)
  mc_value   MC spin-attempts/s = domain cells x sweeps / mcSim
  roofline   the dominant CA kernel (ca_rules): 26 B per cell it stages (13 B of state read + 13 B written) x
             the cells of the tiles it stages per launch, over its mean launch time from HIP events recorded by
             the kernel's own dispatch (every 10th step); traffic = HBM bytes per launch measured by rocprofv3
             counters (profiles/traffic.json).  step_roofline: 26 B x the cells of every visited tile over the
             whole step time (caSim / steps).
  mc_roofline the MC sweep kernel: per sweep, every domain cell's list entry (4 B), mobility threshold (8 B, when
             cooling mobility drives the trial skip) and spin written (spin bytes), and the whole domain's spins read
             once per colour pass (colours x spin bytes) -- dom x (4 + 8 + spin_bytes x (colours + 1)) B per sweep;
             over the sweep time of the persistent kernel (mc_sweeps: all sweeps of a statistics interval in one
             launch) or of the per-colour launches (mc_color x colours).
  cpu_baseline the reference engine (oracle/_ref/engine, built from the reference sources) on every host core
             the process may use, on a bounded sample of the same workload.
  extra      the cone workload line, the 256-job batched mix (against the same jobs one at a time) and the
             serving stream (Poisson arrivals; and the loop's capacity with every request offered at once).
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
CA_BYTES_PER_CELL = 26
# CPU sample of each workload: the first CA steps and Potts sweeps the reference engine runs
CPU_SAMPLE = {"caldera": (300, 5), "cone": (1000, 20), "cross": (400, 100)}


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
    with open(os.path.join(out, "stats.json"), encoding="utf-8") as f:
        times["_stats"] = json.load(f)
    times["_wall_s"] = wall
    return times


def physical_cores() -> int | None:
    """Physical cores of the host (distinct (package, core) pairs in /proc/cpuinfo)."""
    try:
        pairs, phys, core = set(), None, None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    phys = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                elif not line.strip():
                    if core is not None:
                        pairs.add((phys, core))
                    phys = core = None
        return len(pairs) or None
    except OSError:
        return None


def cpu_baseline(kind: str, cfg_path: str, cells: int, tmp: str) -> dict | None:
    if not os.path.exists(REF_ENGINE):
        return None
    threads = len(os.sched_getaffinity(0))
    steps, sweeps = CPU_SAMPLE[kind]
    out = os.path.join(tmp, "cpu")
    log(f"cpu baseline: reference engine, {threads} threads, first {steps} CA steps + {sweeps} sweeps")
    t = run_engine(REF_ENGINE, cfg_path, out, ["--threads", str(threads), "--set", f"ca.maxSteps={steps}",
                                               "--set", f"mc.mcs={sweeps}", "--set", "output.format=none",
                                               "--set", "parallel.affinity=none"], timeout=3600)
    ca_rate = cells * t["steps"] / (t["caSim"] / 1000.0)
    mc_rate = t["domainCells"] * t["mcs"] / (t["mcSim"] / 1000.0) if t.get("mcSim") else None
    return {"value": ca_rate, "unit": "cell-updates/s", "cores": threads, "physical_cores": physical_cores(),
            "kind": "reference", "mc_value": mc_rate, "mc_unit": "spin-attempts/s",
            "sample": f"reference engine (oracle/_ref, OpenMP, {threads} threads = every core this process may use) "
                      f"on the same workload cut to its first {t['steps']} CA steps and {t['mcs']} Potts sweeps "
                      f"({t['domainCells']} domain cells at that point); caSim {t['caSim']:.0f} ms, mcSim "
                      f"{t['mcSim']:.0f} ms.  Early steps have a smaller lava front than the whole run, so the "
                      f"sampled rate overstates the CPU's full-run rate"}


def _issue(workload: str, k: str, ms: float) -> dict:
    if not os.path.exists(PROFILE_ISSUE):
        return {"valu_issue_frac": None, "salu_issue_frac": None}
    with open(PROFILE_ISSUE) as f:
        e = json.load(f).get(workload, {}).get(k)
    if not e:
        return {"valu_issue_frac": None, "salu_issue_frac": None}
    return {"valu_issue_frac": round(e["valu"] / (ms / 1000.0) / ISSUE_PEAK_PER_S, 4),
            "salu_issue_frac": round(e["salu"] / (ms / 1000.0) / ISSUE_PEAK_PER_S, 4)}


def _traffic(workload: str, k: str):
    if not os.path.exists(PROFILE_TRAFFIC):
        return None
    with open(PROFILE_TRAFFIC) as f:
        t = json.load(f)
    return t.get(workload, {}).get(k)


def workload_line(kind: str, a, with_cpu: bool, engine_args=(), profiled: bool = True) -> dict:
    """K timed jobs (after W warm-up jobs) of one scenario through the engine.  profiled: the committed counter
    summaries (profiles/*.json) describe this run's launches, so its issue fractions and traffic come from them."""
    from volcano_ca_mc_amd import workloads as W
    wl = {"caldera": W.caldera_3d, "cone": W.cone_3d, "cross": W.cross_section_2d}[kind]()
    nx, ny, nz = wl.dims
    cells = nx * ny * nz
    with tempfile.TemporaryDirectory(prefix="vca_bench_") as tmp:
        cfg_path = wl.write(tmp)
        # kernel times from HIP events on every TIMING_SAMPLE-th CA step / MC sweep of the timed jobs: timing every
        # launch would add its own events' cost to the measured caSim
        extra = ["--timing-sample", str(TIMING_SAMPLE), "--set", "output.format=none", *engine_args]
        runs = []
        for i in range(a.warmup + a.steps):
            t = run_engine(ENGINE, cfg_path, os.path.join(tmp, f"run{i}"), extra)
            tag = "warmup" if i < a.warmup else "timed"
            log(f"{kind} {tag} job {i}: steps {t['steps']} caSim {t['caSim']:.1f} ms mcSim {t['mcSim']:.1f} ms "
                f"total {t['total']:.1f} ms")
            if i >= a.warmup:
                runs.append(t)
        cpu = cpu_baseline(kind, cfg_path, cells, tmp) if with_cpu else None
    ca_ms = statistics.median(t["caSim"] for t in runs)
    mc_ms = statistics.median(t["mcSim"] for t in runs)
    r0 = runs[0]
    steps, dom, mcs = r0["steps"], r0["domainCells"], r0["mcs"]
    colours = r0["_stats"].get("mc", {}).get("colors", 1)
    lat = sorted(t["total"] for t in runs)
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]
    # kernel means over the sampled launches of every timed job
    ktot = {}
    for t in runs:
        for name, v in t.get("detail", {}).get("kernels", {}).items():
            e = ktot.setdefault(name, [0.0, 0])
            e[0] += v["ms"]
            e[1] += v["launches"]
    mean = {k: v[0] / v[1] for k, v in ktot.items() if v[1]}
    w = r0.get("detail", {}).get("caWork", {})
    tile_cells = w.get("tileCells", 1024)
    # steps the step kernels ran; the rest (a quiescent tail: lava only cooling) completed by the fast-forward
    launched = w.get("stepsLaunched", steps)
    # ca_rules launches per job: the launches after the step that held a quiet simulation (until the host's status
    # read) are no-ops and are timed like the others, so the staged cells are spread over every launch; one launch in
    # TIMING_SAMPLE is timed (the first included), which fixes the count to within TIMING_SAMPLE / 2
    rules_timed = ktot.get("ca_rules", [0.0, 0])[1] / max(1, len(runs))
    rules_launches = max(launched, round(rules_timed * TIMING_SAMPLE - TIMING_SAMPLE / 2))
    staged = w.get("tilesStaged", 0) * tile_cells / max(1, rules_launches)  # cells ca_rules stages per launch
    visited = w.get("tilesVisited", 0) * tile_cells / max(1, launched)    # cells of every visited tile per step
    step_ms = ca_ms / max(1, steps)
    rules_bytes = CA_BYTES_PER_CELL * staged
    rules_ach = rules_bytes / (mean["ca_rules"] / 1000.0) / 1e9
    # the whole CA phase: 26 B per cell of every visited tile over caSim
    step_ach = CA_BYTES_PER_CELL * w.get("tilesVisited", 0) * tile_cells / (ca_ms / 1000.0) / 1e9
    step_kernels = [k for k in mean if k.startswith("ca_") and k != "ca_step"]
    roofline = {"bound": "hbm", "achieved": round(rules_ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(rules_ach / HBM_PEAK_GBS, 4), "traffic": _traffic(kind, "ca_rules") if profiled else None,
                "kernel": "ca_rules", "kernel_mean_ms": round(mean["ca_rules"], 5),
                "timed_every_nth_step": TIMING_SAMPLE, "bytes_per_unit": f"{CA_BYTES_PER_CELL} B per staged cell",
                "staged_cells_per_launch": round(staged), "algorithmic_bytes_per_launch": round(rules_bytes),
                "launches": rules_launches,
                **(_issue(kind, "ca_rules", mean["ca_rules"]) if profiled
                   else {"valu_issue_frac": None, "salu_issue_frac": None}),
                "kernel_mean_ms_all": {k: round(mean[k], 5) for k in step_kernels},
                "launches_per_step": round(sum(ktot[k][1] for k in step_kernels) / max(1, ktot["ca_rules"][1]), 3)}
    step_roofline = {"achieved": round(step_ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(step_ach / HBM_PEAK_GBS, 4), "step_ms": round(step_ms, 5),
                     "visited_cells_per_launched_step": round(visited),
                     "bytes_per_unit": f"{CA_BYTES_PER_CELL} B per cell of every visited tile, over the whole caSim"}
    mc_roofline = None
    mst = r0["_stats"].get("mc", {})
    if dom and ("mc_sweeps" in mean or "mc_color" in mean):
        q = int(mst.get("Q", 2))
        spin_bytes = 1 if q <= 255 else (2 if q <= 65535 else 4)
        dyn = mst.get("mobility") == "cooling"
        sweep_bytes = dom * (4 + (8 if dyn else 0) + spin_bytes * (colours + 1))
        if "mc_sweeps" in mean:
            kname = "mc_sweeps"
            launches_per_job = ktot[kname][1] / len(runs)
            sweeps_per_launch = mcs / launches_per_job
            sweep_ms = mean[kname] / sweeps_per_launch
        else:
            kname = "mc_color"
            sweeps_per_launch = 1.0 / colours
            sweep_ms = mean[kname] * colours
        mc_bytes = sweep_bytes * sweeps_per_launch
        mc_ach = mc_bytes / (mean[kname] / 1000.0) / 1e9
        mc_roofline = {"bound": "latency", "kernel": kname, "kernel_mean_ms": round(mean[kname], 5),
                       "sweeps_per_launch": round(sweeps_per_launch, 4), "sweep_ms": round(sweep_ms, 5),
                       "achieved": round(mc_ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(mc_ach / HBM_PEAK_GBS, 4), "traffic": _traffic(kind, kname),
                       "colours": colours, "spin_bytes": spin_bytes, "mobility_skip": dyn,
                       "bytes_per_unit": f"{4 + (8 if dyn else 0) + spin_bytes * (colours + 1)} B per domain cell "
                                         f"and sweep",
                       "algorithmic_bytes_per_launch": round(mc_bytes), **_issue(kind, kname, mean[kname])}
    desc = {"caldera": f"caldera_3d {nx}x{ny}x{nz} random pentagonal, periodic x/y, 8 craters, maxSteps 2000, "
                       f"Potts Q128 kT0.5 100 sweeps (checkerboard)",
            "cone": f"cone_3d {nx}x{ny}x{nz} moore absorbing, 3 craters, maxSteps 1000, Potts Q64 kT0.5 200 sweeps",
            "cross": f"cross_section_2d {nx}x{ny}x{nz} von Neumann absorbing, 1 crater, maxSteps 400, Potts Q64 kT0.5 "
                     f"100 sweeps"}[kind]
    return {"value": cells * steps / (ca_ms / 1000.0), "mc_value": dom * mcs / (mc_ms / 1000.0),
            "p50_job_ms": statistics.median(lat), "p99_job_ms": p99, "ca_ms": ca_ms, "mc_ms": mc_ms,
            "config": {"workload": desc, "cells": cells, "ca_steps": steps, "ca_steps_launched": launched,
                       "ca_steps_fast_forwarded": w.get("fastForwardSteps", 0),
                       "ca_fast_forward_from": w.get("fastForwardFrom", -1), "mc_domain_cells": dom, "mc_sweeps": mcs,
                       "mc_colours": colours, "parallelism": "single GPU"},
            "roofline": roofline, "step_roofline": step_roofline, "mc_roofline": mc_roofline, "cpu_baseline": cpu}


def _write_jobs(workloads, folder: str) -> None:
    os.makedirs(folder, exist_ok=True)
    for wl in workloads:
        wl.write(folder)


def _pipeline(args, out: str, engine_args=()) -> dict:
    t0 = time.time()
    # the engine's per-job progress lines go to our stderr (the run's log)
    r = subprocess.run([ENGINE, *args, "--out", out, "--set", "output.format=none", *engine_args], stdout=sys.stderr,
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


def bench_batched(jobs_n: int, engine_args=()) -> dict:
    """The uneven mix: the batched pipeline against the same jobs one at a time (--max-jobs 1), same engine;
    stats.json of every job must be equal between the two."""
    from volcano_ca_mc_amd import workloads as W
    jobs = W.batched_mix(jobs=jobs_n)
    with tempfile.TemporaryDirectory(prefix="vca_batched_") as tmp:
        folder = os.path.join(tmp, "jobs")
        _write_jobs(jobs, folder)
        log(f"{len(jobs)} jobs written")
        one = _pipeline(["--batch", folder, "--gpu-batch", "--max-jobs", "1"], os.path.join(tmp, "one"), engine_args)
        log(f"one at a time: {one['makespanS']:.2f} s, {one['jobsPerS']:.2f} jobs/s")
        many = _pipeline(["--batch", folder, "--gpu-batch"], os.path.join(tmp, "many"), engine_args)
        log(f"batched: {many['makespanS']:.2f} s, {many['jobsPerS']:.2f} jobs/s")
        same = _same_stats(os.path.join(tmp, "one"), os.path.join(tmp, "many"), [w.name for w in jobs])
    return {"metric": "jobs/s (batched uneven simulations, bit-exact)", "value": many["jobsPerS"], "unit": "jobs/s",
            "vs_one_at_a_time": many["jobsPerS"] / max(1e-9, one["jobsPerS"]),
            "one_at_a_time_jobs_per_s": one["jobsPerS"], "makespan_s": many["makespanS"],
            "one_at_a_time_makespan_s": one["makespanS"], "cell_updates_per_s": many["cellUpdatesPerS"],
            "p50_job_ms": many["p50LatencyMs"], "p99_job_ms": many["p99LatencyMs"], "identical_stats": same,
            "gpu_thread_ms": many.get("gpuThread"),
            "workload": f"{len(jobs)} jobs from one seed: 2D edges 128-2048, 3D edges 64-256, 200-800 steps, "
                        "50 Potts sweeps Q64"}


def bench_serving(rate: float, duration: float, scale: float, capacity_jobs: int) -> dict:
    """The serving mix as a Poisson stream (latency at the offered rate), then the first capacity_jobs requests
    of the same stream all offered at t = 0 (the loop's capacity)."""
    from volcano_ca_mc_amd import workloads as W
    arr = W.serving_stream(rate=rate, duration=duration, scale=scale)
    out = {"workload": f"Poisson {rate}/s for {duration} s: 70% 2D 512x256 x300 steps, 25% 3D 256^3 x600, "
                       f"5% 768x768x256 x1200, 50-200 sweeps; grid scale {scale}", "offered_rate": rate}
    with tempfile.TemporaryDirectory(prefix="vca_serve_") as tmp:
        folder = os.path.join(tmp, "jobs")
        _write_jobs([x.workload for x in arr], folder)

        def stream(path, items, at0):
            with open(path, "w") as f:
                for x in items:
                    f.write(json.dumps({"at": 0.0 if at0 else x.t,
                                        "config": os.path.join(folder, x.workload.name + ".json"),
                                        "name": x.workload.name}) + "\n")
            return path

        log(f"{len(arr)} requests over {duration} s written")
        s = _pipeline(["--serve", stream(os.path.join(tmp, "jobs.jsonl"), arr, False)], os.path.join(tmp, "out"))
        out.update({"jobs_per_s": s["jobsPerS"], "p50_job_ms": s["p50LatencyMs"], "p99_job_ms": s["p99LatencyMs"],
                    "jobs": s["jobs"], "ok": s["ok"], "makespan_s": s["makespanS"],
                    "cell_updates_per_s": s["cellUpdatesPerS"], "spin_attempts_per_s": s["spinAttemptsPerS"]})
        cap = arr[:capacity_jobs]
        c = _pipeline(["--serve", stream(os.path.join(tmp, "cap.jsonl"), cap, True)], os.path.join(tmp, "cap"))
        out["capacity"] = {"jobs": c["jobs"], "ok": c["ok"], "jobs_per_s": c["jobsPerS"],
                           "makespan_s": c["makespanS"], "cell_updates_per_s": c["cellUpdatesPerS"],
                           "note": "the same requests all offered at t = 0"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["caldera", "cone", "cross", "batched", "serving"], default="caldera",
                    help="caldera (default line), cone, cross: one scenario; batched: 256 uneven jobs; "
                         "serving: Poisson stream")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=30.0)
    ap.add_argument("--capacity-jobs", type=int, default=300)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed jobs")
    ap.add_argument("--warmup", type=int, default=1, help="untimed jobs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="only the main workload line")
    ap.add_argument("--scale", type=float, default=1.0, help="grid scale of the serving mix (1.0 = the benchmark)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; runs on more than one GPU are not part of it",
                          "n_gpus": a.gpus}))
        return 2
    if not os.path.exists(ENGINE):
        raise SystemExit(f"{ENGINE} missing: run `python -m volcano_ca_mc_amd.build` first")
    if a.workload == "batched":
        print(json.dumps({"n_gpus": 1, **bench_batched(a.jobs)}))
        return 0
    if a.workload == "serving":
        s = bench_serving(a.rate, a.duration, a.scale, a.capacity_jobs)
        print(json.dumps({"metric": "served jobs/s with p50/p99 latency (bit-exact)", "value": s["jobs_per_s"],
                          "unit": "jobs/s", "n_gpus": 1, **s}))
        return 0
    line = workload_line(a.workload, a, not a.no_cpu)
    extra = {}
    if not a.no_extra:
        # the same workload with every CA step launched (no quiescent fast-forward): what the step kernels alone give
        log(f"extra: {a.workload} with every step launched")
        # (no counter summary describes its launches: no issue fractions or measured traffic)
        e = workload_line(a.workload, argparse.Namespace(steps=2, warmup=0), False, ["--no-fast-forward"],
                          profiled=False)
        extra["every_step_launched"] = {k: e[k] for k in ("value", "ca_ms", "config", "roofline", "step_roofline")}
        if a.workload != "cone":
            log("extra: cone workload")
            sub = argparse.Namespace(steps=3, warmup=1)
            c = workload_line("cone", sub, False)
            extra["cone"] = {k: c[k] for k in ("value", "mc_value", "p50_job_ms", "p99_job_ms", "config", "roofline",
                                               "step_roofline", "mc_roofline")}
        log("extra: batched mix")
        extra["batched"] = bench_batched(a.jobs)
        log("extra: serving stream")
        extra["serving"] = bench_serving(a.rate, a.duration, 1.0, a.capacity_jobs)
    result = {
        "metric": "CA cell-updates/s + MC spin-attempts/s (bit-exact); p50/p99 job latency",
        "value": line["value"],
        "unit": "cell-updates/s",
        "mc_value": line["mc_value"],
        "mc_unit": "spin-attempts/s",
        "p50_job_ms": line["p50_job_ms"],
        "p99_job_ms": line["p99_job_ms"],
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": line["ca_ms"] + line["mc_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8+i32+f32",
        "data": "synthetic (seeded terrain)",
        "config": line["config"],
        "roofline": line["roofline"],
        "step_roofline": line["step_roofline"],
        "mc_roofline": line["mc_roofline"],
        "cpu_baseline": line["cpu_baseline"],
        "extra": extra,
    }
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
