"""The `cor` command line on the CPU: flags, defaults and error behaviour of the
reference's command line (src/main.c:12-315, src/args.c:11-73) that need no GPU."""
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
COR = REPO / "cor-model_amd" / "cor"
GOLD = REPO / "tests" / "golden"

pytestmark = pytest.mark.skipif(not COR.exists(), reason="cor-model_amd/cor is not built")


def run(*args, cwd=None):
    return subprocess.run([str(COR), *map(str, args)], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_help_lists_the_reference_flags():
    out = run("--help")
    assert out.returncode == 0
    for flag in ("--par", "--output", "--steps-count", "--tu-angle", "--mode", "--particle-type",
                 "--internal-field", "--external-field", "--crustal-field"):
        assert flag in out.stdout, flag
    assert run("--version").returncode == 0


def test_errors_match_the_reference_messages(tmp_path):
    assert "Invalid internal geomagnetic field model: 'igrf99'" in run("-i", "igrf99", "x").stderr
    assert "Invalid external geomagnetic field model: 't01'" in run("-e", "t01", "x").stderr
    assert "Invalid crustal geomagnetic field model: 'x'" in run("-c", "x", "x").stderr
    r = run("-s", "12x", GOLD / "config1_igrf_nk100.in")
    assert r.returncode == 1 and "Invalid steps count '12x'" in r.stderr
    r = run("-u", "abc", GOLD / "config1_igrf_nk100.in")
    assert r.returncode == 1 and "Invalid TU angle 'abc'" in r.stderr
    r = run(tmp_path / "missing.in")
    assert r.returncode == 1 and "Cannot open input file for reading" in r.stderr
    r = run("-m", "trajectory", GOLD / "config1_igrf_nk100.in")
    assert r.returncode == 1 and "Output file name is required" in r.stderr
    r = run("-m", "trajectory", "-o", "x", GOLD / "config1_igrf_nk100.in", cwd=tmp_path)
    assert r.returncode == 1 and "starting and ending rigidity must be the same" in r.stderr
    r = run("-m", "sweep", GOLD / "config1_igrf_nk100.in")
    assert r.returncode == 1 and "Invalid calculation mode" in r.stderr
    bad = tmp_path / "bad.in"
    bad.write_text("0.01 -1 20\n1.0 10 20\n10 20\n2020 1 1 1 0 0 0\n100 0 0 0.0\n")
    r = run(bad)
    assert r.returncode == 1 and "Rigidity step size must be bigger than 0" in r.stderr
    r = run("-e", "t96", GOLD / "config1_igrf_nk100.in")
    assert r.returncode == 1 and "Invalid input file" in r.stderr


def test_batch_form_checks_its_output_directory(tmp_path):
    r = run("-o", tmp_path / "nodir", GOLD / "t96_equator_nk400.in", GOLD / "t96_mid_nk400.in", "-e", "t96")
    assert r.returncode == 1 and "existing directory" in r.stderr
    r = run("-m", "field", GOLD / "field_igrf_t96.in", GOLD / "field_igrf_chaos.in")
    assert r.returncode == 1 and "cutoff mode only" in r.stderr
