"""The batched pipeline of the engine (`--batch --gpu-batch`, `--serve`): every job's files identical to the
reference engine's, whether the job runs alone or in a mix that shares GPU steps.

All golden configs go through one in-process series whose CA phases share batched launches (with a broken
config among them: the series goes on, exit code 5, an error row in batch_summary.csv), and through a
serving stream where jobs arrive while others are running.
"""
import csv
import gzip
import json
import os
import subprocess

import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = os.path.join(ROOT, "volcano_ca_mc_amd", "_lib", "engine")
pytestmark = pytest.mark.gpu


def _cases():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return sorted(json.load(f))


def _write_config(name: str, dst: str) -> None:
    """The golden config with its terrain file made absolute, so configs of many cases share one folder."""
    src = os.path.join(GOLD, "cases", name)
    cfg = json.load(open(os.path.join(src, "config.json")))
    t = cfg.get("terrain", {})
    if t.get("type") == "file" and not os.path.isabs(t["file"]):
        t["file"] = os.path.join(src, t["file"])
    with open(dst, "w") as f:
        json.dump(cfg, f)


def _check_outputs(out: str, name: str) -> None:
    d = os.path.join(GOLD, "cases", name)
    for fn in ("ca_result.bin", "mc_result.bin"):
        p = os.path.join(d, fn + ".gz")
        got = os.path.join(out, fn)
        if os.path.exists(p):
            assert open(got, "rb").read() == gzip.decompress(open(p, "rb").read()), (name, fn)
        else:
            assert not os.path.exists(got), (name, fn)
    p = os.path.join(d, "mc_stats.csv")
    if os.path.exists(p):
        assert open(os.path.join(out, "mc_stats.csv")).read() == open(p).read(), name
    got = json.load(open(os.path.join(out, "stats.json"), encoding="utf-8"))
    got.pop("parallel", None)
    assert got == json.load(open(os.path.join(d, "stats.json"), encoding="utf-8")), name
    times = json.load(open(os.path.join(out, "times.json")))
    assert times["pipeline"] == "gpu-batch" and times["total"] > 0


def test_gpu_batch_series_matches_reference(tmp_path):
    series = tmp_path / "series"
    series.mkdir()
    names = _cases()
    for n in names:
        _write_config(n, str(series / f"{n}.json"))
    (series / "zz_broken.json").write_text(json.dumps({"space": {"nx": 0}}))
    out = tmp_path / "out"
    print(f"--gpu-batch series of {len(names)} configs + 1 broken", flush=True)
    r = subprocess.run([ENGINE, "--batch", str(series), "--gpu-batch", "--out", str(out)], capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 5, r.stdout[-3000:] + r.stderr[-3000:]
    for n in names:
        _check_outputs(str(out / n), n)
    rows = list(csv.DictReader(open(out / "batch_summary.csv", encoding="utf-8")))
    assert len(rows) == len(names) + 1
    status = {row["name"]: row["status"] for row in rows}
    assert status["zz_broken"] == "error" and all(status[n] == "ok" for n in names)


def test_serve_stream_jobs_match_reference_and_report_latency(tmp_path):
    cfgs = tmp_path / "cfg"
    cfgs.mkdir()
    names = _cases()
    lines = []
    # 48 requests over ~1.2 s: every golden case once, some twice, arriving while others run
    picks = names + names[::4]
    for k, n in enumerate(picks):
        p = cfgs / f"{n}.json"
        if not p.exists():
            _write_config(n, str(p))
        lines.append(json.dumps({"at": 0.025 * k, "config": str(p), "name": f"r{k:03d}_{n}"}))
    jobs = tmp_path / "jobs.jsonl"
    jobs.write_text("\n".join(lines) + "\n")
    out = tmp_path / "out"
    print(f"--serve stream of {len(lines)} requests", flush=True)
    r = subprocess.run([ENGINE, "--serve", str(jobs), "--out", str(out)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    for k, n in enumerate(picks):
        _check_outputs(str(out / f"r{k:03d}_{n}"), n)
    s = json.load(open(out / "serve_summary.json"))
    assert s["jobs"] == len(picks) and s["ok"] == len(picks)
    assert 0 < s["p50LatencyMs"] <= s["p99LatencyMs"]
    assert s["jobsPerS"] > 0 and s["cellUpdatesPerS"] > 0
    print(f"serve: {s['jobsPerS']:.1f} jobs/s, p50 {s['p50LatencyMs']:.1f} ms, p99 {s['p99LatencyMs']:.1f} ms",
          flush=True)
