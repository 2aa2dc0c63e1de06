"""Shared test setup.

Markers: `gpu` tests need a visible MI355X (gfx950) and call libbude_mi355x
through its C ABI; everything else runs on the CPU.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    lib = ROOT / "minibude_amd" / "lib" / "libbude_mi355x.so"
    drv = ROOT / "minibude_amd" / "bin" / "bude-mi355x"
    mkdeck = ROOT / "minibude_amd" / "bin" / "bude-makedeck"
    host_srcs = list((ROOT / "minibude_amd" / "host").glob("*.cpp"))
    mkdeck_srcs = list((ROOT / "minibude_amd" / "host" / "makedeck").glob("*.cpp"))
    if not lib.exists() or (host_srcs and not drv.exists()) or (mkdeck_srcs and not mkdeck.exists()):
        # Build only when nothing is built (a machine without the toolchain runs the shipped build).
        from minibude_amd import build
        build.build()
    if not (ROOT / "oracle" / "liboracle.so").exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True, capture_output=True)


@pytest.fixture(scope="session")
def gpu_device():
    from minibude_amd import model
    if model.device_count() == 0:
        pytest.fail("gpu test selected but no HIP device is visible")
    return 0
