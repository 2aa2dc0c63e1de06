"""Build recipe for libbude_mi355x (gfx950) and the C++ driver.

    python -m minibude_amd.build            # library + driver
    python -m minibude_amd.build --lib-only

Outputs (in-tree, git-ignored, shipped with the tree):
    minibude_amd/lib/libbude_mi355x.so      C-ABI library (include/bude_mi355x.h)
    minibude_amd/bin/bude-mi355x            driver (the reference's CLI/report)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"
OBJ = PKG / "lib" / "obj"
LIB = PKG / "lib" / "libbude_mi355x.so"
DRIVER = PKG / "bin" / "bude-mi355x"
MAKEDECK = PKG / "bin" / "bude-makedeck"

ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is switched off for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]

LIB_SOURCES = [
    "prep.cpp",
    "deck_synth.cpp",
    "deck_io.cpp",
    "fasten_plain.hip",
    "capi_plain.hip",
    "capi_noalloc.hip",
    "fasten_batch.hip",
    "capi_batch.hip",
    "mlp.hip",
    "surrogate.hip",
    "capi_screen.hip",
    "select.hip",
    "field.hip",
    "atomnet.hip",
    "capi_atomnet.hip",
    "capi_combined.hip",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libbude_mi355x needs the ROCm toolchain to build")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    obj = OBJ / (src.name + ".o")
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    if src.suffix == ".hip":
        cmd = [_hipcc(), *HIP_FLAGS, *inc, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *HOST_FLAGS, *inc, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int | None = None) -> Path:
    srcs = [CSRC / s for s in LIB_SOURCES]
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> Path | None:
    srcs = sorted(HOST.glob("*.cpp"))
    if not srcs:
        return None
    DRIVER.parent.mkdir(parents=True, exist_ok=True)
    tmp = DRIVER.with_name(DRIVER.name + ".tmp")
    _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", f"-I{HOST}", *map(str, srcs), "-o", str(tmp),
          f"-L{LIB.parent}", "-lbude_mi355x", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_makedeck() -> Path:
    """bin/bude-makedeck: decks from bhff/mol2 files (the makedeck program)."""
    srcs = sorted((HOST / "makedeck").glob("*.cpp"))
    MAKEDECK.parent.mkdir(parents=True, exist_ok=True)
    tmp = MAKEDECK.with_name(MAKEDECK.name + ".tmp")
    _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", *map(str, srcs), "-o", str(tmp),
          f"-L{LIB.parent}", "-lbude_mi355x", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, MAKEDECK)
    return MAKEDECK


def build(lib_only: bool = False) -> None:
    build_library()
    if not lib_only:
        build_driver()
        build_makedeck()


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lib-only", action="store_true")
    args = ap.parse_args(argv)
    build(lib_only=args.lib_only)
    print(f"built {LIB}" + ("" if args.lib_only else f", {DRIVER} and {MAKEDECK}"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
