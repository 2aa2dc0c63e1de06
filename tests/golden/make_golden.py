"""Write the golden fixtures under tests/golden/ (run in the build container,
where the reference sources exist; the outputs are committed).

    python tests/golden/make_golden.py [--only bm|c0|c1]

Fixtures:
  bm1/, bm2/      the reference's own decks and golden energies, copied
                  byte for byte from data/bm1 and data/bm2 (inputs and expected
                  outputs of the reference's only correctness check).
  c0/, c1/        fp64.f32: FP64-oracle energies and magnitudes (0.5 * sum of
                  |steric| + |charge| + |desolvation| over all pairs), float32
                  pairs [E, M] per pose, from oracle/liboracle.so; the
                  magnitude is the scale of cancellation in each pose's sum.
                  reference energies of the seeded synthetic workload decks:
                  ref_energies.f32 (float32, little-endian, one per pose)
                  computed by the reference's own per-pose kernel
                  (makedeck/ref-kernel.cpp:15-139, built strict by
                  oracle/Makefile into oracle/_ref/libref_kernel.so), and
                  manifest.json with the generator parameters and the SHA-256
                  of every deck file, so the deck can be regenerated and
                  checked byte for byte wherever the fixture is used.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import shutil
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from minibude_amd import deck as deckmod  # noqa: E402

REF = Path("/root/reference")
REF_LIB = ROOT / "oracle" / "_ref" / "libref_kernel.so"


def ref_energies(d: deckmod.Deck, first: int = 0, count: int | None = None) -> np.ndarray:
    lib = ctypes.CDLL(str(REF_LIB))
    count = d.nposes - first if count is None else count
    out = np.zeros(count, np.float32)
    P, S = ctypes.c_void_p, ctypes.c_size_t
    rows = [P(d.poses[k].ctypes.data) for k in range(6)]
    lib.ref_fasten(S(d.natlig), S(d.natpro), P(d.protein.ctypes.data), P(d.ligand.ctypes.data),
                   P(d.forcefield.ctypes.data), S(d.ntypes), *rows, S(first), S(count), P(out.ctypes.data))
    return out


def fp64_fixture(name: str) -> None:
    d = deckmod.workload_deck(name)
    lib = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    P, S = ctypes.c_void_p, ctypes.c_size_t
    e = np.zeros(d.nposes, np.float64)
    m = np.zeros(d.nposes, np.float64)
    rows = [P(d.poses[k].ctypes.data) for k in range(6)]
    t = time.time()
    lib.bude_oracle_fasten_f64_mag(S(d.natlig), S(d.natpro), P(d.protein.ctypes.data), P(d.ligand.ctypes.data),
                                   P(d.forcefield.ctypes.data), *rows, S(0), S(d.nposes), P(e.ctypes.data),
                                   P(m.ctypes.data))
    np.stack([e, m], axis=1).astype("<f4").tofile(HERE / name / "fp64.f32")
    print(f"{name}: FP64 energies + magnitudes in {time.time() - t:.1f}s")


def copy_bm(name: str) -> None:
    src = REF / "data" / name
    dst = HERE / name
    dst.mkdir(exist_ok=True)
    for f in ("protein.in", "ligand.in", "forcefield.in", "poses.in", "ref_energies.out"):
        shutil.copyfile(src / f, dst / f)
    print(f"copied {src} -> {dst}")


def synth_fixture(name: str) -> None:
    w = deckmod.WORKLOADS[name]
    d = deckmod.workload_deck(name)
    t = time.time()
    e = ref_energies(d)
    dt = time.time() - t
    out = HERE / name
    out.mkdir(exist_ok=True)
    e.astype("<f4").tofile(out / "ref_energies.f32")
    with tempfile.TemporaryDirectory() as tmp:
        sha = deckmod.write_deck(tmp, d)
    meta = {
        "workload": name,
        "generator": w,
        "ntypes": d.ntypes,
        "deck_sha256": sha,
        "ref_energies": "ref_energies.f32: float32 LE, makedeck/ref-kernel.cpp fasten_main per pose, g++ -O2 "
                        "-ffp-contract=off",
        "ref_seconds": round(dt, 1),
    }
    (out / "manifest.json").write_text(json.dumps(meta, indent=1) + "\n")
    print(f"{name}: {d.nposes} poses in {dt:.1f}s -> {out}")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["bm", "c0", "c1", "fp64"])
    a = ap.parse_args()
    if not REF_LIB.exists():
        print("build oracle/_ref first: make -C oracle ref", file=sys.stderr)
        return 1
    if a.only in (None, "bm"):
        copy_bm("bm1")
        copy_bm("bm2")
    for w in ("c0", "c1"):
        if a.only in (None, w):
            synth_fixture(w)
    if a.only in (None, "fp64"):
        for w in ("c0", "c1"):
            fp64_fixture(w)
    return 0


if __name__ == "__main__":
    sys.exit(main())
