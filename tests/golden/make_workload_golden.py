"""Reference digests for members of the batched mix (256 uneven jobs) and of the serving stream, at their own sizes.

The jobs come from volcano_ca_mc_amd/workloads.py (seeded); the reference engine ``oracle/_ref/engine`` (compiled from
the reference sources by ``make -C oracle ref``, never committed) runs each selected job alone, and its outputs are
stored as SHA-256 digests (ca_result.bin, mc_result.bin, mc_stats.csv) plus stats.json without its machine section,
together with digests of the job's config and heightmap, so that a change of the generator is detected instead of
compared against stale digests.  Output: tests/golden/workloads.json.

Usage:  python tests/golden/make_workload_golden.py [--threads 8]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from volcano_ca_mc_amd import workloads as W  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "engine")
OUT = os.path.join(HERE, "workloads.json")

# the serving stream members the GPU test runs (its first SERVE_JOBS requests; four of them are 768 x 768 x 256)
SERVE_DURATION = 30.0
SERVE_JOBS = 40


def sha_file(p: str) -> str:
    h = hashlib.sha256()
    with open(p, "rb") as f:
        for b in iter(lambda: f.read(1 << 22), b""):
            h.update(b)
    return h.hexdigest()


def job_inputs_digest(wl, folder: str) -> dict:
    cfg = wl.write(folder)
    return {"config_sha256": sha_file(cfg), "heights_sha256": sha_file(os.path.join(folder, wl.name + ".csv"))}


def batched_selection(jobs):
    """The largest 2D and 3D jobs, and for every neighbourhood its smallest 2D and 3D jobs."""
    d2 = [w for w in jobs if w.dims[1] == 1]
    d3 = [w for w in jobs if w.dims[1] > 1]
    pick = {max(d2, key=lambda w: w.cells).name, max(d3, key=lambda w: w.cells).name}
    for nb in ("von_neumann", "moore", "pentagonal", "hexagonal"):
        for group in (d2, d3):
            c = sorted((w for w in group if w.config["neighborhood"] == nb), key=lambda w: w.cells)
            if c:
                pick.add(c[0].name)
    return sorted(pick)


def serving_selection(arr):
    """Two requests of every kind among the first SERVE_JOBS."""
    picked, seen = [], {}
    for x in arr[:SERVE_JOBS]:
        k = x.workload.dims
        if seen.get(k, 0) < 2:
            seen[k] = seen.get(k, 0) + 1
            picked.append(x.workload.name)
    return picked


def run_reference(wl, threads: int) -> dict:
    with tempfile.TemporaryDirectory() as tmp:
        entry = job_inputs_digest(wl, tmp)
        out = os.path.join(tmp, "out")
        r = subprocess.run([REF, "--config", os.path.join(tmp, wl.name + ".json"), "--out", out, "--threads",
                            str(threads)], cwd=tmp, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        for fn in ("ca_result.bin", "mc_result.bin", "mc_stats.csv"):
            p = os.path.join(out, fn)
            if os.path.exists(p):
                entry[fn] = sha_file(p)
        stats = json.load(open(os.path.join(out, "stats.json"), encoding="utf-8"))
        stats.pop("parallel", None)
        entry["stats"] = stats
        entry["dims"] = list(wl.dims)
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/engine is missing: run `make -C oracle ref` where /root/reference exists")
    out = {"batched_mix": {}, "serving": {"duration": SERVE_DURATION, "jobs": SERVE_JOBS, "members": {}}}
    jobs = {w.name: w for w in W.batched_mix()}
    for name in batched_selection(list(jobs.values())):
        out["batched_mix"][name] = run_reference(jobs[name], a.threads)
        print("batched", name, jobs[name].dims, flush=True)
    arr = W.serving_stream(duration=SERVE_DURATION)
    reqs = {x.workload.name: x.workload for x in arr}
    for name in serving_selection(arr):
        out["serving"]["members"][name] = run_reference(reqs[name], a.threads)
        print("serving", name, reqs[name].dims, flush=True)
    with open(OUT, "w", encoding="utf-8") as f:
        json.dump(out, f, indent=1, ensure_ascii=False, sort_keys=True)


if __name__ == "__main__":
    main()
