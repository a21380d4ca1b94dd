"""The batched and serving workloads at their own sizes, bit-exact.

Batched mix: 256 uneven jobs from one seed (2D x-z grids with edges up to 2048, 3D grids with edges up to 256, every
neighbourhood, 200-800 CA steps, 50 Potts sweeps).  The whole series runs through the batched pipeline (every step
launch serves all running jobs) and again one job at a time (--max-jobs 1); every job's ca_result.bin, mc_result.bin
(SHA-256, --digests), mc_stats.csv and stats.json must be the same in both runs, and equal to the reference engine's
for the jobs the reference ran (tests/golden/workloads.json: the largest 2D and 3D jobs, the smallest 2D and 3D job of
every neighbourhood).  The series semantics follow engine/src/app/Batch.cpp:186-246: each job its own output folder,
a failing job does not stop the others.

Serving stream: the first 40 requests of the seeded Poisson stream (70 % 2D 512 x 256 with 300 steps, 25 % 256^3
with 600, 5 % 768 x 768 x 256 with 1200 -- four of these; 50-200 sweeps), at their arrival times (the first 2 s of
the 600 s stream).  Each result must equal the same job run alone (`engine --config`), and the reference engine's
for two requests of each kind.
"""
import json
import os
import subprocess

import pytest

from volcano_ca_mc_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = os.path.join(ROOT, "volcano_ca_mc_amd", "_lib", "engine")
GOLD = os.path.join(ROOT, "tests", "golden", "workloads.json")
pytestmark = pytest.mark.gpu


def _sha(path):
    import hashlib
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for b in iter(lambda: f.read(1 << 22), b""):
            h.update(b)
    return h.hexdigest()


def _outputs(folder):
    """Digests and statistics of one job's output folder (result files as written by --digests)."""
    out = {}
    for fn in ("ca_result.bin", "mc_result.bin"):
        p = os.path.join(folder, fn + ".sha256")
        out[fn] = open(p).read().split()[0] if os.path.exists(p) else None
        assert not os.path.exists(os.path.join(folder, fn)), fn
    p = os.path.join(folder, "mc_stats.csv")
    out["mc_stats.csv"] = _sha(p) if os.path.exists(p) else None
    with open(os.path.join(folder, "stats.json"), encoding="utf-8") as f:
        stats = json.load(f)
    stats.pop("parallel", None)
    out["stats"] = stats
    return out


def _reference(entry):
    return {k: entry.get(k) for k in ("ca_result.bin", "mc_result.bin", "mc_stats.csv", "stats")}


def _run(args, timeout):
    print("engine " + " ".join(args[:3]), flush=True)
    r = subprocess.run([ENGINE, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


def test_batched_mix_256_jobs(tmp_path):
    jobs = W.batched_mix(jobs=256)
    folder = tmp_path / "jobs"
    for wl in jobs:
        wl.write(str(folder))
    gold = json.load(open(GOLD, encoding="utf-8"))["batched_mix"]
    for name, e in gold.items():
        wl = next(w for w in jobs if w.name == name)
        assert _sha(os.path.join(folder, name + ".json")) == e["config_sha256"], name
        assert _sha(os.path.join(folder, name + ".csv")) == e["heights_sha256"], name
    _run(["--batch", str(folder), "--gpu-batch", "--digests", "--out", str(tmp_path / "many")], 900)
    _run(["--batch", str(folder), "--gpu-batch", "--max-jobs", "1", "--digests", "--out", str(tmp_path / "one")], 900)
    checked = 0
    for wl in jobs:
        many = _outputs(tmp_path / "many" / wl.name)
        one = _outputs(tmp_path / "one" / wl.name)
        assert many == one, wl.name
        assert many["ca_result.bin"] and many["mc_result.bin"], wl.name
        if wl.name in gold:
            assert many == _reference(gold[wl.name]), wl.name
            checked += 1
    assert checked == len(gold) >= 8
    dims = [tuple(gold[n]["dims"]) for n in gold]
    assert max(d[0] for d in dims if d[1] == 1) > 2000 and max(d[0] * d[1] * d[2] for d in dims if d[1] > 1) > 1.3e7


def test_serving_stream_results_equal_solo_runs(tmp_path):
    g = json.load(open(GOLD, encoding="utf-8"))["serving"]
    arr = W.serving_stream(duration=g["duration"])[: g["jobs"]]
    big = [x for x in arr if x.workload.dims == (768, 768, 256)]
    assert len(big) >= 2
    folder = tmp_path / "jobs"
    lines = []
    for x in arr:
        cfg = x.workload.write(str(folder))
        lines.append(json.dumps({"at": x.t, "config": cfg, "name": x.workload.name}))
    for name, e in g["members"].items():
        assert _sha(os.path.join(folder, name + ".json")) == e["config_sha256"], name
    (tmp_path / "jobs.jsonl").write_text("\n".join(lines) + "\n")
    _run(["--serve", str(tmp_path / "jobs.jsonl"), "--digests", "--out", str(tmp_path / "served")], 900)
    summary = json.load(open(tmp_path / "served" / "serve_summary.json"))
    assert summary["ok"] == len(arr)
    for x in arr:
        name = x.workload.name
        _run(["--config", str(folder / (name + ".json")), "--digests", "--out", str(tmp_path / "solo" / name)], 300)
        served = _outputs(tmp_path / "served" / name)
        assert served == _outputs(tmp_path / "solo" / name), name
        if name in g["members"]:
            assert served == _reference(g["members"][name]), name
