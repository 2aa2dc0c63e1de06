"""The bude-mi355x driver's command line and deck checks (no GPU needed)."""
import subprocess

import pytest

from conftest import GOLDEN, ROOT

DRIVER = ROOT / "minibude_amd" / "bin" / "bude-mi355x"


def run(*args):
    return subprocess.run([str(DRIVER), *args], capture_output=True, text=True, timeout=60)


def test_help_lists_reference_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--device", "--iter", "--poses", "--ppwi", "--wgsize", "--deck", "--out", "--rows", "--csv",
                 "--list"):
        assert flag in r.stdout


def test_unsupported_ppwi_is_rejected():
    r = run("--deck", str(GOLDEN / "bm1"), "-p", "3")
    assert r.returncode == 1 and "PPWI 3 is not a supported value" in r.stderr


def test_malformed_integer():
    r = run("-i", "abc")
    assert r.returncode == 1 and "integer required for <iter>" in r.stderr


def test_unknown_argument():
    r = run("--bogus")
    assert r.returncode == 1 and "Unrecognized argument" in r.stdout


def test_bad_deck_aborts():
    r = run("--deck", "/nonexistent")
    assert r.returncode != 0 and "Bad file" in r.stderr


def test_too_many_poses():
    r = run("--deck", str(GOLDEN / "bm1"), "-n", "70000")
    assert r.returncode != 0 and "exceeded max poses" in r.stderr


def test_deck_header_is_reported():
    r = run("--deck", str(GOLDEN / "bm1"), "-n", "1024", "-p", "8", "-w", "64")
    assert "proteins:     938" in r.stdout and "ligands:      26" in r.stdout and "poses:        1024" in r.stdout


def test_list_without_gpu():
    from minibude_amd import model
    if model.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = run("--list")
    assert r.returncode == 0 and "devices:" in r.stdout
