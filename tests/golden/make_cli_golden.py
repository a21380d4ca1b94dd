"""Writes the fixtures of the engine's process contract under tests/golden/cli/ by running the reference engine.

The reference engine is ``oracle/_ref/engine`` (compiled from the reference sources by ``make -C oracle ref``;
never committed).  Every fixture is data: the inputs of a command (config files, heightmaps, a CA result folder)
and what the reference printed or wrote for it.

  validate.json      ``engine --validate`` on valid and invalid configs: stdout and exit code
                     (engine/src/app/ConfigCheck.cpp:43-72, messages from engine/src/io/Config.cpp)
  print_defaults.txt ``engine --print-defaults`` stdout
  runs/<case>/       runs of small configs: text and binary result files (io/ResultWriter.cpp:105-236), CA and MC
                     snapshots (app/Simulation.cpp:212-243, app/McStage.cpp:186-203), --progress lines
                     (app/Progress.hpp), --mc-from from a CA result folder (app/Simulation.cpp:334-402)
  batch/             ``engine --batch`` over a folder with broken configs: batch_summary.csv columns that do not
                     depend on the machine or the clock (app/Batch.cpp:186-330)

Usage:  python tests/golden/make_cli_golden.py
"""
from __future__ import annotations

import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "engine")
OUT = os.path.join(HERE, "cli")

# a small 2D lava config the runs below share (reference defaults otherwise)
LAVA2D = {"space": {"nx": 70, "ny": 1, "nz": 40}, "rngSeed": 5,
          "terrain": {"peakHeight": 26, "baseRadius": 30, "craterRadius": 5, "craterDepth": 3},
          "vents": {"eruptionSteps": 40},
          "ca": {"coolingAmbient": 3, "coolingLateral": 6, "coolingFall": 2, "maxSteps": 300},
          "mc": {"enabled": True, "Q": 12, "mcs": 9, "statsEvery": 4, "kT": 0.5, "update": "checkerboard"}}
LAVA3D = {"space": {"nx": 30, "ny": 26, "nz": 18}, "neighborhood": "moore", "rngSeed": 9,
          "terrain": {"peakHeight": 14, "baseRadius": 15, "craterRadius": 3, "craterDepth": 2},
          "vents": {"count": 2, "eruptionSteps": 25},
          "ca": {"coolingAmbient": 3, "coolingLateral": 6, "coolingFall": 2, "maxSteps": 200},
          "mc": {"enabled": True, "Q": 9, "mcs": 7, "statsEvery": 3, "kT": 0.5, "update": "checkerboard"}}


def deep(base, patch):
    out = json.loads(json.dumps(base))
    for k, v in patch.items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = deep(out[k], v)
        else:
            out[k] = v
    return out


# --validate cases: name -> (files {name: text}, extra args).  The config is always config.json in the run folder.
def validate_cases():
    j = json.dumps
    c = {}
    c["ok_defaults"] = ({"config.json": "{}"}, [])
    c["ok_lava2d"] = ({"config.json": j(LAVA2D)}, [])
    c["warn_positions_ignored"] = ({"config.json": j({"vents": {"positions": [[10, 0]]}})}, [])
    c["warn_positions_ignored_random"] = ({"config.json": j({"vents": {"placement": "random",
                                                                       "positions": [[10, 0]]}})}, [])
    c["warn_classic_positions"] = ({"config.json": j({"ca": {"mode": "classic"}, "vents": {"placement": "positions"}})},
                                   [])
    c["warn_mobility_full"] = ({"config.json": j({"mc": {"enabled": True, "domain": "full"}})}, [])
    c["warn_random_threads"] = ({"config.json": j({"mc": {"enabled": True}, "parallel": {"threads": 4}})}, [])
    c["err_unknown_key"] = ({"config.json": j({"foo": 1})}, [])
    c["err_unknown_nested"] = ({"config.json": j({"ca": {"speed": 1}})}, [])
    c["err_section_scalar"] = ({"config.json": j({"space": 5})}, [])
    c["err_nx_zero"] = ({"config.json": j({"space": {"nx": 0}})}, [])
    c["err_nx_float"] = ({"config.json": j({"space": {"nx": 1.5}})}, [])
    c["err_nx_huge"] = ({"config.json": '{"space": {"nx": 18446744073709551615}}'}, [])
    c["err_nx_negative"] = ({"config.json": j({"space": {"nx": -3}})}, [])
    c["err_1d"] = ({"config.json": j({"space": {"nx": 10, "ny": 1, "nz": 1}})}, [])
    c["err_too_many_cells"] = ({"config.json": j({"space": {"nx": 100000, "ny": 100000, "nz": 1}})}, [])
    c["err_enum"] = ({"config.json": j({"boundary": "torus"})}, [])
    c["err_enum_number"] = ({"config.json": j({"neighborhood": 4})}, [])
    c["err_bool"] = ({"config.json": j({"ca": {"ponding": "yes"}})}, [])
    c["err_number_string"] = ({"config.json": j({"ca": {"coolingAmbient": "hot"}})}, [])
    c["err_number_range"] = ({"config.json": j({"mc": {"mobilityMin": 1.5}})}, [])
    c["err_number_range_neg"] = ({"config.json": j({"ca": {"coolingFall": -0.25}})}, [])
    c["err_j_zero"] = ({"config.json": j({"mc": {"J": 0}})}, [])
    c["err_seed_negative"] = ({"config.json": j({"rngSeed": -1})}, [])
    c["err_maxsteps"] = ({"config.json": j({"ca": {"maxSteps": 70000}})}, [])
    c["err_lava2d_ny"] = ({"config.json": j({"space": {"nx": 50, "ny": 40, "nz": 1}})}, [])
    c["err_hex_3d"] = ({"config.json": j({"space": {"nx": 20, "ny": 20, "nz": 20}, "neighborhood": "hexagonal"})}, [])
    c["err_terrain_file_type"] = ({"config.json": j({"terrain": {"file": 5}})}, [])
    c["err_terrain_file_missing"] = ({"config.json": j({"terrain": {"type": "file"}})}, [])
    c["err_positions_not_list"] = ({"config.json": j({"vents": {"positions": 5}})}, [])
    c["err_positions_pair"] = ({"config.json": j({"vents": {"positions": [[1]]}})}, [])
    c["err_positions_outside"] = ({"config.json": j({"vents": {"positions": [[500, 0]]}})}, [])
    c["err_positions_dup"] = ({"config.json": j({"vents": {"count": 2, "positions": [[10, 0], [10, 0]]}})}, [])
    c["err_positions_empty"] = ({"config.json": j({"vents": {"placement": "positions"}})}, [])
    c["err_positions_count"] = ({"config.json": j({"vents": {"count": 2, "placement": "positions",
                                                             "positions": [[10, 0]]}})}, [])
    c["err_erupt_temp"] = ({"config.json": j({"vents": {"eruptTemp": 700}})}, [])
    c["err_classic_count"] = ({"config.json": j({"ca": {"mode": "classic"}, "space": {"nx": 10, "ny": 1, "nz": 10},
                                                 "vents": {"count": 101}})}, [])
    c["err_basalt_classic"] = ({"config.json": j({"ca": {"mode": "classic"}, "mc": {"enabled": True}})}, [])
    c["err_output_format"] = ({"config.json": j({"output": {"format": "csv"}})}, [])
    c["err_parallel_schedule"] = ({"config.json": j({"parallel": {"schedule": "auto"}})}, [])
    c["err_mpi_io"] = ({"config.json": j({"mpi": {"io": "posix"}})}, [])
    c["err_not_object"] = ({"config.json": "[1, 2]"}, [])
    c["err_override_form"] = ({"config.json": "{}"}, ["--set", "ca.maxSteps"])
    c["err_override_empty_key"] = ({"config.json": "{}"}, ["--set", "=5"])
    c["err_override_path"] = ({"config.json": "{}"}, ["--set", "ca..maxSteps=5"])
    c["err_override_range"] = ({"config.json": "{}"}, ["--set", "ca.maxSteps=-1"])
    c["ok_override_text"] = ({"config.json": "{}"}, ["--set", "boundary=periodic"])
    c["err_override_unknown"] = ({"config.json": "{}"}, ["--set", "ca.foo=1"])
    c["err_missing_config"] = ({}, [])
    c["err_heightmap_missing"] = ({"config.json": j({"terrain": {"type": "file", "file": "none.csv"}})}, [])
    c["err_heightmap_bad"] = ({"config.json": j({"terrain": {"type": "file", "file": "h.csv"},
                                                 "space": {"nx": 3, "ny": 1, "nz": 10}}),
                               "h.csv": "1,2,x\n"}, [])
    c["err_heightmap_shape"] = ({"config.json": j({"terrain": {"type": "file", "file": "h.csv"},
                                                   "space": {"nx": 4, "ny": 1, "nz": 10}}),
                                 "h.csv": "1,2,3\n"}, [])
    c["ok_heightmap"] = ({"config.json": j({"terrain": {"type": "file", "file": "h.csv"},
                                            "space": {"nx": 3, "ny": 1, "nz": 10}}),
                          "h.csv": "1,2.5,3\n"}, [])
    # syntax errors: the reference's JSON library reports them with its own message grammar (position, context,
    # token); files given as "base64:..." hold bytes that are not UTF-8
    syntax = {
        "trailing_comma": '{"space": {"nx": 10,}}', "missing_colon": '{"a" 1}', "empty": "", "blank": "   \n  ",
        "overflow": '{"space": {"nx": 1e999}}', "bad_literal": '{"x": tru}', "unterminated": '{"x": "abc',
        "bad_escape": '{"x": "a\\qb"}', "short_u": '{"x": "\\u12"}', "lone_high": '{"x": "\\ud800"}',
        "high_then_ascii": '{"x": "\\ud800\\u0041"}', "lone_low": '{"x": "\\udc00"}', "tab_in_string": '{"x": "a\tb"}',
        "minus": '{"x": -}', "dot": '{"x": 1.}', "exp": '{"x": 1e}', "exp_sign": '{"x": 1e+}', "leading_zero": '{"x": 01}',
        "array_comma": '{"x": [1, 2,]}', "array_missing_comma": '{"x": [1 2]}', "trailing_text": "{} x",
        "stray_slash": "{}\n\n  /", "open_comment": "/* abc", "comment_then_error": '// c\n{"space": 5}',
        "multiline": '{"x":\n [1,\n 2\n  }', "crlf": '{\r\n"a":1\r\n,}', "big_unsigned": '{"x": 18446744073709551616}',
        "two_documents": '{"a":1}{', "lone_comma": "{,}", "missing_value": '{"a":}',
        "bad_utf8": "base64:eyJ4IjogIsMoIn0=", "bad_bom": "base64:77t7fQ==", "surrogate_utf8": "base64:eyJ4IjogIu2ggCJ9",
    }
    for k, text in syntax.items():
        c["syntax_" + k] = ({"config.json": text}, [])
    c["mcfrom_missing_folder"] = ({"config.json": "{}"}, ["--mc-from", "nowhere"])
    c["mcfrom_empty_folder"] = ({"config.json": "{}", "ca/.keep": ""}, ["--mc-from", "ca"])
    c["mcfrom_full_domain"] = ({"config.json": j({"mc": {"domain": "full"}}), "ca/ca_result.bin": ""},
                               ["--mc-from", "ca"])
    return c


def run(args, cwd, timeout=600):
    r = subprocess.run([REF, *args], cwd=cwd, capture_output=True, timeout=timeout)
    return r.returncode, r.stdout.decode("utf-8"), r.stderr.decode("utf-8")


def write_files(folder, files):
    import base64
    for name, text in files.items():
        p = os.path.join(folder, name)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        data = base64.b64decode(text[7:]) if text.startswith("base64:") else text.encode("utf-8")
        with open(p, "wb") as f:
            f.write(data)


def make_validate():
    out = {}
    for name, (files, extra) in validate_cases().items():
        with tempfile.TemporaryDirectory() as tmp:
            write_files(tmp, files)
            rc, so, se = run(["--validate", "--config", "config.json", *extra], tmp)
        out[name] = {"files": files, "args": extra, "rc": rc, "stdout": so,
                     "stderr_first_line": se.splitlines()[0] if se else ""}
    with open(os.path.join(OUT, "validate.json"), "w", encoding="utf-8") as f:
        json.dump(out, f, indent=1, ensure_ascii=False, sort_keys=True)
    rc, so, _ = run(["--print-defaults"], ROOT)
    assert rc == 0
    with open(os.path.join(OUT, "print_defaults.txt"), "w", encoding="utf-8") as f:
        f.write(so)


def gz(src, dst):
    with open(src, "rb") as fi, open(dst, "wb") as fo:
        fo.write(gzip.compress(fi.read(), compresslevel=9, mtime=0))


def keep_outputs(run_dir, dst):
    """Copies the result files of a run (gzip for the large ones) and stats.json without its machine section."""
    os.makedirs(dst, exist_ok=True)
    for fn in ("ca_result.bin", "ca_result.txt", "mc_result.bin", "mc_result.txt"):
        p = os.path.join(run_dir, fn)
        if os.path.exists(p):
            gz(p, os.path.join(dst, fn + ".gz"))
    p = os.path.join(run_dir, "mc_stats.csv")
    if os.path.exists(p):
        shutil.copy(p, os.path.join(dst, "mc_stats.csv"))
    stats = json.load(open(os.path.join(run_dir, "stats.json"), encoding="utf-8"))
    stats.pop("parallel", None)
    with open(os.path.join(dst, "stats.json"), "w", encoding="utf-8") as f:
        json.dump(stats, f, indent=1, ensure_ascii=False)
    for sub in ("snapshots", "mc_snapshots"):
        s = os.path.join(run_dir, sub)
        if os.path.isdir(s):
            os.makedirs(os.path.join(dst, sub), exist_ok=True)
            for fn in sorted(os.listdir(s)):
                gz(os.path.join(s, fn), os.path.join(dst, sub, fn + ".gz"))


# runs: name -> (config, extra args); the config is written as config.json in the case folder
def run_cases():
    return {
        "txt_both_2d": (deep(LAVA2D, {"output": {"format": "both"}}), []),
        "txt_3d_pentagonal": (deep(LAVA3D, {"neighborhood": "pentagonal", "boundary": "periodic",
                                            "output": {"format": "txt"}}), []),
        "txt_classic_2d": ({"space": {"nx": 30, "ny": 20, "nz": 1}, "neighborhood": "von_neumann",
                            "boundary": "periodic", "rngSeed": 4, "vents": {"count": 6}, "ca": {"mode": "classic"},
                            "mc": {"enabled": True, "domain": "full", "mobilityFromCooling": False, "Q": 5, "mcs": 4,
                                   "statsEvery": 2, "update": "checkerboard"},
                            "output": {"format": "both"}}, []),
        "snapshots_3d": (deep(LAVA3D, {"output": {"format": "bin", "snapshotsEvery": 7},
                                       "mc": {"snapshotsEvery": 3}}), []),
        "snapshots_2d_every1": (deep(LAVA2D, {"ca": {"maxSteps": 12}, "output": {"format": "bin", "snapshotsEvery": 1},
                                              "mc": {"snapshotsEvery": 1, "mcs": 3}}), []),
        "progress_2d": (deep(LAVA2D, {"output": {"format": "bin"}}), ["--progress"]),
        "progress_snapshots_3d": (deep(LAVA3D, {"output": {"format": "none", "snapshotsEvery": 50}}), ["--progress"]),
    }


def make_runs():
    base = os.path.join(OUT, "runs")
    os.makedirs(base, exist_ok=True)
    for name, (cfg, extra) in run_cases().items():
        d = os.path.join(base, name)
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "config.json"), "w") as f:
            json.dump(cfg, f, indent=2)
        with tempfile.TemporaryDirectory() as tmp:
            rc, so, se = run(["--config", "config.json", "--out", tmp, *extra], d)
            assert rc == 0, se
            keep_outputs(tmp, os.path.join(d, "expected"))
            if "--progress" in extra:
                lines = [ln for ln in so.splitlines() if ln.startswith("PROGRESS")]
                with open(os.path.join(d, "expected", "progress.txt"), "w") as f:
                    f.write("\n".join(lines) + "\n")
        with open(os.path.join(d, "args.json"), "w") as f:
            json.dump(extra, f)
        print("run", name, flush=True)
    # --mc-from: the MC stage on the CA result of a finished run, read from its folder (bin, then txt only)
    for name, fmt, patch in (("mcfrom_bin_2d", "bin", {}),
                             ("mcfrom_txt_3d", "txt", {"mc": {"kT": 0.0, "mobilityMin": 0.2}}),
                             ("mcfrom_nomobility_2d", "bin", {"mc": {"mobilityFromCooling": False, "Q": 40}})):
        cfg = deep(LAVA2D if name.endswith("2d") else LAVA3D, deep({"output": {"format": fmt}}, patch))
        d = os.path.join(base, name)
        os.makedirs(os.path.join(d, "ca"), exist_ok=True)
        with open(os.path.join(d, "config.json"), "w") as f:
            json.dump(cfg, f, indent=2)
        with tempfile.TemporaryDirectory() as tmp:
            rc, _, se = run(["--config", "config.json", "--out", tmp, "--set", "mc.enabled=false"], d)
            assert rc == 0, se
            src = os.path.join(tmp, f"ca_result.{fmt}")
            gz(src, os.path.join(d, "ca", f"ca_result.{fmt}.gz"))
            shutil.copy(src, os.path.join(d, "ca", f"ca_result.{fmt}"))
        with tempfile.TemporaryDirectory() as tmp:
            rc, _, se = run(["--mc-from", "ca", "--config", "config.json", "--out", tmp], d)
            assert rc == 0, se
            keep_outputs(tmp, os.path.join(d, "expected"))
        os.remove(os.path.join(d, "ca", f"ca_result.{fmt}"))
        with open(os.path.join(d, "args.json"), "w") as f:
            json.dump(["--mc-from", "ca"], f)
        print("run", name, flush=True)


# batch_summary.csv columns that depend neither on the machine nor on the clock
BATCH_COLUMNS = ["name", "status", "exitCode", "error", "mode", "nx", "ny", "nz", "cells", "grid", "neighborhood",
                 "boundary", "seeds", "steps", "stopReason", "mcs", "Q", "domainCells", "caResultTxtBytes",
                 "caResultBinBytes", "mcResultTxtBytes", "mcResultBinBytes"]


def read_summary(path):
    import csv
    with open(path, newline="", encoding="utf-8") as f:
        rows = list(csv.DictReader(f))
    return [{k: r[k] for k in BATCH_COLUMNS} for r in rows]


def batch_files():
    j = json.dumps
    return {
        "a_good.json": j(deep(LAVA2D, {"output": {"format": "bin"}})),
        "b_positions.json": j({"vents": {"placement": "positions"}}),
        "c_syntax.json": '{"space": {"nx": 10,}}',
        "d_hex3d.json": j({"space": {"nx": 20, "ny": 20, "nz": 20}, "neighborhood": "hexagonal"}),
        "e_good_txt.json": j(deep(LAVA3D, {"output": {"format": "txt"}, "mc": {"enabled": False}})),
        "f_erupt.json": j({"vents": {"eruptTemp": 100}}),
        "g_heightmap.json": j({"terrain": {"type": "file", "file": "missing.csv"}}),
        "_skipped.json": "{}",
    }


def make_batch():
    d = os.path.join(OUT, "batch")
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(os.path.join(d, "series"))
    write_files(os.path.join(d, "series"), batch_files())
    with tempfile.TemporaryDirectory() as tmp:
        rc, so, se = run(["--batch", "series", "--out", tmp], d, timeout=1200)
        rows = read_summary(os.path.join(tmp, "batch_summary.csv"))
    with open(os.path.join(d, "expected.json"), "w", encoding="utf-8") as f:
        json.dump({"rc": rc, "columns": BATCH_COLUMNS, "rows": rows}, f, indent=1, ensure_ascii=False)
    print("batch rc", rc, flush=True)


if __name__ == "__main__":
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/engine is missing: run `make -C oracle ref` where /root/reference exists")
    os.makedirs(OUT, exist_ok=True)
    make_validate()
    make_runs()
    make_batch()
