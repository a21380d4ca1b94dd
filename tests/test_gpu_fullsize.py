"""Full-size parity of the three benchmark grids on the GPU.

The 2D cross-section (1024x1x512, 400 steps), the 3D cone (512x512x256, 1000 steps) and the caldera
(1024x1024x384, random pentagonal, periodic x/y, up to 2000 steps) are written by the seeded generator,
run through the drop-in `engine`, and every output file is compared with the SHA-256 digest of the
reference engine's file on the same config (tests/golden/scenarios, made by tests/golden/make_golden.py).
stats.json must equal the reference's value for value (minus its CPU "parallel" section).
"""
import hashlib
import json
import os
import shutil
import subprocess

import pytest

from volcano_ca_mc_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden", "scenarios")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = os.path.join(ROOT, "volcano_ca_mc_amd", "_lib", "engine")
pytestmark = pytest.mark.gpu


def _sha(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        while True:
            b = f.read(1 << 24)
            if not b:
                break
            h.update(b)
    return h.hexdigest()


@pytest.mark.parametrize("name", ["cross_section_2d", "cone_3d", "caldera_3d"])
def test_fullsize_outputs_match_reference(tmp_path, name):
    ref = json.load(open(os.path.join(GOLD, f"{name}.json"), encoding="utf-8"))
    wl = W.SCENARIOS[name]()
    need = wl.cells * (13 + 4) + (1 << 30)  # ca_result + mc_result + slack
    if shutil.disk_usage(str(tmp_path)).free < need:
        pytest.skip(f"needs {need >> 30} GiB of free disk for the result files")
    print(f"full-size case {name}: dims {wl.dims}", flush=True)
    cfg = wl.write(str(tmp_path))
    assert _sha(os.path.join(str(tmp_path), f"{name}.csv")) == ref["heights_sha256"]
    out = tmp_path / "out"
    r = subprocess.run([ENGINE, "--config", cfg, "--out", str(out)], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    try:
        for fn in ("ca_result.bin", "mc_result.bin", "mc_stats.csv"):
            assert _sha(str(out / fn)) == ref[fn], fn
        got = json.loads((out / "stats.json").read_text(encoding="utf-8"))
        got.pop("parallel", None)
        assert got == ref["stats"]
    finally:
        for fn in ("ca_result.bin", "mc_result.bin"):
            p = out / fn
            if p.exists():
                p.unlink()
    times = json.loads((out / "times.json").read_text())
    print(f"{name}: steps {times['steps']} caSim {times['caSim']:.1f} ms mcSim {times['mcSim']:.1f} ms "
          f"total {times['total']:.1f} ms (reference {ref['reference_threads']} threads: "
          f"caSim {ref['reference_times_ms']['caSim']:.0f} ms)", flush=True)
