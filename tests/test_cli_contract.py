"""The engine's process contract on the CPU (no kernel runs): `--validate` and `--print-defaults` print exactly what
the reference engine printed for the same inputs (fixtures made by tests/golden/make_cli_golden.py).

`--validate` (engine/src/app/ConfigCheck.cpp:43-72) reports every configuration error with the reference's message
text (engine/src/io/Config.cpp); the GUI shows these texts and `--batch` copies them into batch_summary.csv.
"""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tests", "golden", "cli")
ENGINE = os.path.join(ROOT, "volcano_ca_mc_amd", "_lib", "engine")

with open(os.path.join(CLI, "validate.json"), encoding="utf-8") as _f:
    VALIDATE = json.load(_f)


def _write(folder, files):
    """Files of a case; "base64:..." holds bytes that are not UTF-8."""
    import base64
    for name, text in files.items():
        p = os.path.join(folder, name)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        data = base64.b64decode(text[7:]) if text.startswith("base64:") else text.encode("utf-8")
        with open(p, "wb") as f:
            f.write(data)


@pytest.mark.parametrize("name", sorted(VALIDATE))
def test_validate_output_is_the_references(tmp_path, name):
    case = VALIDATE[name]
    _write(str(tmp_path), case["files"])
    r = subprocess.run([ENGINE, "--validate", "--config", "config.json", *case["args"]], cwd=tmp_path,
                       capture_output=True, timeout=60)
    assert r.stdout.decode("utf-8") == case["stdout"]
    assert r.returncode == case["rc"]
    err = r.stderr.decode("utf-8").splitlines()
    assert (err[0] if err else "") == case["stderr_first_line"]


def test_validate_cases_cover_errors_and_warnings():
    """The fixture set holds valid configs, configs with warnings, and at least one error per message family."""
    oks = [c for c in VALIDATE.values() if c["rc"] == 0]
    printed = [json.loads(c["stdout"]) for c in VALIDATE.values() if c["stdout"]]
    errs = [p["errors"][0] for p in printed if p["errors"]]
    errs += [c["stderr_first_line"] for c in VALIDATE.values() if c["rc"] == 2]
    warns = [w for p in printed for w in p["warnings"]]
    assert len(oks) >= 3 and len(warns) >= 5
    for stem in ("Nieznany parametr", "jest sekcją", "musi być z zakresu", "musi być liczbą całkowitą",
                 "jest za duży", "ma nieprawidłową wartość", "musi mieć wartość true lub false", "musi być liczbą,",
                 "przekrój pionowy", "hexagonal", "leży poza siatką", "powtarza kolumnę", "wymaga listy kraterów",
                 "musi być równe liczbie kraterów", "eruptTemp", "nie może przekraczać liczby komórek",
                 "wymaga ca.mode", "wymaga postaci klucz=wartość", "Nieprawidłowa ścieżka", "Nie można otworzyć",
                 "musi zawierać obiekt JSON", "--mc-from", "nie ma ca_result.bin", "parse error at line",
                 "number overflow", "invalid literal", "invalid string", "invalid number", "invalid comment",
                 "attempting to parse an empty input", "invalid BOM", "ill-formed UTF-8 byte"):
        assert any(stem in e for e in errs), stem


def test_print_defaults_is_the_references():
    r = subprocess.run([ENGINE, "--print-defaults"], capture_output=True, timeout=60)
    assert r.returncode == 0
    with open(os.path.join(CLI, "print_defaults.txt"), encoding="utf-8") as f:
        assert r.stdout.decode("utf-8") == f.read()
