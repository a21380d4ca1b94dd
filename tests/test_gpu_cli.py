"""The engine's process contract on the GPU: text and binary result files, CA and MC snapshots, --progress lines,
--mc-from and the --batch summary, each against what the reference engine wrote for the same command
(fixtures made by tests/golden/make_cli_golden.py).

- ca_result.txt / mc_result.txt: engine/src/io/ResultWriter.cpp:105-150 (byte for byte)
- snapshots/step_<k>.bin (app/Simulation.cpp:212-243) and mc_snapshots/mcs_<k>.bin (app/McStage.cpp:186-203)
- --progress: app/Progress.hpp -- the lines printed regardless of the clock (start and end of each stage) must be the
  reference's; lines the 500 ms clock lets through in between must carry the run's own step counts and histories
- --mc-from <folder>: app/Simulation.cpp:334-402, from ca_result.bin and from ca_result.txt
- --batch: app/Batch.cpp:186-330, the batch_summary.csv columns that depend neither on the machine nor on the clock
"""
import csv
import gzip
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tests", "golden", "cli")
ENGINE = os.path.join(ROOT, "volcano_ca_mc_amd", "_lib", "engine")
pytestmark = pytest.mark.gpu

RUNS = sorted(os.listdir(os.path.join(CLI, "runs")))


def _unzip(p):
    with open(p, "rb") as f:
        return gzip.decompress(f.read())


def _check_outputs(expected, got, ca_folder=None):
    for fn in sorted(os.listdir(expected)):
        p = os.path.join(expected, fn)
        if fn.endswith(".gz"):
            assert (got / fn[:-3]).read_bytes() == _unzip(p), fn
        elif fn == "mc_stats.csv":
            assert (got / fn).read_text() == open(p).read(), fn
        elif fn == "stats.json":
            s = json.loads((got / fn).read_text(encoding="utf-8"))
            s.pop("parallel", None)
            want = json.load(open(p, encoding="utf-8"))
            if "source" in want:
                # --mc-from records the absolute path of the CA folder it read: the run's own folder here
                assert s["source"]["folder"] == os.path.abspath(ca_folder), s["source"]
                s["source"]["folder"] = want["source"]["folder"]
            assert s == want, fn
        elif fn in ("snapshots", "mc_snapshots"):
            want = sorted(os.listdir(p))
            assert sorted(x + ".gz" for x in os.listdir(got / fn)) == want, fn
            for w in want:
                assert (got / fn / w[:-3]).read_bytes() == _unzip(os.path.join(p, w)), f"{fn}/{w}"
    for fn in ("ca_result.bin", "ca_result.txt", "mc_result.bin", "mc_result.txt"):
        if not os.path.exists(os.path.join(expected, fn + ".gz")):
            assert not (got / fn).exists(), fn


def _progress_ok(ref_lines, got_lines, stats):
    """Lines printed regardless of the clock equal the reference's; any line in between carries the run's own
    numbers (CA: step k and the lava count after step k; MC: a sweep count between 0 and mcs)."""
    hist = stats.get("activeCells")
    kept = []
    for ln in got_lines:
        if ln in ref_lines:
            kept.append(ln)
            continue
        f = ln.split()
        assert f[0] == "PROGRESS" and f[1] in ("ca", "mc"), ln
        done, total = map(int, f[2].split("/"))
        assert 0 < done < total, ln
        if f[1] == "ca" and hist is not None:
            assert f[3] == f"active={hist[done - 1]}", ln
    assert kept == ref_lines


@pytest.mark.parametrize("name", RUNS)
def test_run_matches_reference(tmp_path, name):
    d = os.path.join(CLI, "runs", name)
    args = json.load(open(os.path.join(d, "args.json")))
    work = tmp_path / "case"
    shutil.copytree(d, work, ignore=shutil.ignore_patterns("expected"))
    # a CA result folder given to --mc-from holds the reference's ca_result file
    ca = work / "ca"
    if ca.is_dir():
        for fn in os.listdir(ca):
            if fn.endswith(".gz"):
                (ca / fn[:-3]).write_bytes(_unzip(str(ca / fn)))
                (ca / fn).unlink()
    out = tmp_path / "out"
    print(f"engine run {name}: {args}", flush=True)
    r = subprocess.run([ENGINE, "--config", "config.json", "--out", str(out), *args], cwd=work, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    _check_outputs(os.path.join(d, "expected"), out, str(ca))
    p = os.path.join(d, "expected", "progress.txt")
    if os.path.exists(p):
        ref = [ln for ln in open(p).read().splitlines() if ln]
        got = [ln for ln in r.stdout.splitlines() if ln.startswith("PROGRESS")]
        _progress_ok(ref, got, json.loads((out / "stats.json").read_text(encoding="utf-8")))


def _summary(path, columns):
    with open(path, newline="", encoding="utf-8") as f:
        return [{k: row[k] for k in columns} for row in csv.DictReader(f)]


@pytest.mark.parametrize("gpu_batch", [False, True])
def test_batch_summary_matches_reference(tmp_path, gpu_batch):
    """--batch over a folder with good and broken configs: exit code 5 and the reference's summary rows (status, exit
    code, error text, dimensions, steps, stop reason, MC sizes, result file sizes); --gpu-batch runs the same series in
    one process and must report the same."""
    d = os.path.join(CLI, "batch")
    exp = json.load(open(os.path.join(d, "expected.json"), encoding="utf-8"))
    work = tmp_path / "b"
    shutil.copytree(d, work)
    out = tmp_path / "out"
    print(f"engine --batch (gpu_batch={gpu_batch})", flush=True)
    extra = ["--gpu-batch"] if gpu_batch else []
    r = subprocess.run([ENGINE, "--batch", "series", "--out", str(out), *extra], cwd=work, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == exp["rc"], r.stdout + r.stderr
    assert _summary(out / "batch_summary.csv", exp["columns"]) == exp["rows"]
