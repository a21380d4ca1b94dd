"""Drop-in parity of the `engine` CLI on the GPU: every golden config, every output file.

The engine (C++17 host + HIP kernels) is run exactly like the reference engine was run to make the
fixtures (`engine --config <case>/config.json --out <dir>`); ca_result.bin, mc_result.bin and
mc_stats.csv must be byte-identical, stats.json equal value for value (its machine-dependent
"parallel" section excluded, as in the fixtures).
"""
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


@pytest.mark.parametrize("name", _cases())
def test_engine_outputs_match_reference(tmp_path, name):
    d = os.path.join(GOLD, "cases", name)
    print(f"engine case {name}", flush=True)
    r = subprocess.run([ENGINE, "--config", os.path.join(d, "config.json"), "--out", str(tmp_path)], cwd=d,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    for fn in ("ca_result.bin", "mc_result.bin"):
        p = os.path.join(d, fn + ".gz")
        if os.path.exists(p):
            assert (tmp_path / fn).read_bytes() == gzip.decompress(open(p, "rb").read()), fn
        else:
            assert not (tmp_path / fn).exists()
    p = os.path.join(d, "mc_stats.csv")
    if os.path.exists(p):
        assert (tmp_path / "mc_stats.csv").read_text() == open(p).read()
    got = json.loads((tmp_path / "stats.json").read_text(encoding="utf-8"))
    got.pop("parallel", None)
    ref = json.load(open(os.path.join(d, "stats.json"), encoding="utf-8"))
    assert got == ref
    times = json.loads((tmp_path / "times.json").read_text())
    assert times["unit"] == "ms" and "total" in times


def test_engine_batch_runs_children_and_reports_failures(tmp_path):
    """--batch: one good and one bad config -> exit 5, summary with one ok row and one error row."""
    src = os.path.join(GOLD, "cases", "lava2d_von_neumann_absorbing", "config.json")
    folder = tmp_path / "series"
    folder.mkdir()
    (folder / "a_good.json").write_text(open(src).read())
    (folder / "b_bad.json").write_text(json.dumps({"space": {"nx": 0}}))
    (folder / "_ignored.json").write_text("{}")
    out = tmp_path / "out"
    r = subprocess.run([ENGINE, "--batch", str(folder), "--out", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 5, r.stdout + r.stderr
    rows = (out / "batch_summary.csv").read_text().strip().splitlines()
    assert len(rows) == 3
    assert rows[1].startswith("a_good,ok,0") and rows[2].startswith("b_bad,error,1")
    assert (out / "a_good" / "ca_result.bin").exists()


@pytest.mark.parametrize("name", ["lava3d_moore_absorbing", "lava2d_von_neumann_absorbing"])
def test_engine_autotune_keeps_outputs(tmp_path, name):
    """--autotune: probes every tile shape x front mode on the run's own inputs, runs with the fastest, and the
    result files stay identical to the reference's."""
    d = os.path.join(GOLD, "cases", name)
    print(f"engine --autotune case {name}", flush=True)
    r = subprocess.run([ENGINE, "--config", os.path.join(d, "config.json"), "--out", str(tmp_path), "--autotune", "8"],
                       cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    p = os.path.join(d, "ca_result.bin.gz")
    assert (tmp_path / "ca_result.bin").read_bytes() == gzip.decompress(open(p, "rb").read())
    got = json.loads((tmp_path / "stats.json").read_text(encoding="utf-8"))
    got.pop("parallel", None)
    assert got == json.load(open(os.path.join(d, "stats.json"), encoding="utf-8"))
    tune = json.loads((tmp_path / "times.json").read_text())["autotune"]
    cfg = json.load(open(os.path.join(d, "config.json")))
    three_d = cfg["space"].get("ny", 1) > 1 and cfg["space"].get("nz", 1) > 1
    assert len(tune["msPerStep"]) == (10 if three_d else 2)
    assert tune["sweep"] in ("front", "full") and tune["probeSteps"] == 8
    assert min(tune["msPerStep"].values()) == tune["msPerStep"][f"{tune['sweep']} {tune['tileShape']}"]
