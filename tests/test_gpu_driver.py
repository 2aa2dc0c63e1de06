"""The drivers on the GPU: this project's bude-mi355x, and the reference's own
driver (src/main.cpp) with the mi355x model dropped in (oracle/_ref/mi355x-bude,
built where the reference sources exist).  Both validate against the deck's
reference energies with the reference's own harness."""
import re
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

DRIVER = ROOT / "minibude_amd" / "bin" / "bude-mi355x"
DROPIN = ROOT / "oracle" / "_ref" / "mi355x-bude"


def run(exe, *args):
    print(exe.name, *args, flush=True)
    return subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("exe", [DRIVER, DROPIN], ids=["bude-mi355x", "reference-driver"])
def test_bm1_sweep_yaml(gpu_device, exe):
    if not exe.exists():
        pytest.skip(f"{exe} not built")
    r = run(exe, "--deck", str(GOLDEN / "bm1"), "-p", "1,8,64", "-w", "64,256,1024", "-i", "2")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.count("valid: true") == 9 and "valid: false" not in r.stdout
    assert re.search(r"best: \{ min_ms: [0-9.]+", r.stdout)


@pytest.mark.parametrize("exe", [DRIVER, DROPIN], ids=["bude-mi355x", "reference-driver"])
def test_csv_and_energy_dump(gpu_device, exe, tmp_path):
    if not exe.exists():
        pytest.skip(f"{exe} not built")
    out = tmp_path / "e.txt"
    r = run(exe, "--deck", str(GOLDEN / "bm2"), "-n", "4096", "-p", "8", "-w", "128", "-i", "1", "--csv",
            "-o", str(out))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if re.match(r"^\d+,\d+,", ln)]
    assert len(lines) == 1 and lines[0].startswith("8,128,")
    assert len(out.read_text().split()) == 4096


def test_default_wgsize_is_incompatible_not_wrong(gpu_device):
    # the reference default wgsize=1 is not a whole wavefront on this model:
    # the driver skips it before any device buffer is made and exits with
    # failure status (not an abort), reporting no result
    r = run(DRIVER, "--deck", str(GOLDEN / "bm1"), "-n", "64", "-i", "1")
    assert r.returncode == 1, (r.returncode, r.stderr[-2000:])
    assert "not compatible" in r.stderr and "wgsize 1" in r.stderr and "skipping" in r.stderr
    assert "best: ~" in r.stdout and "valid:" not in r.stdout
    # a mix: the incompatible point is skipped, the whole-wavefront one runs and validates
    r = run(DRIVER, "--deck", str(GOLDEN / "bm1"), "-n", "256", "-i", "1", "-w", "1,64")
    assert r.returncode == 1 and r.stdout.count("valid: true") == 1, r.stdout[-2000:] + r.stderr[-2000:]
