"""Decks in the reference format, and the seeded synthetic decks of the
benchmark workloads.

Deck directory layout (src/bude.h:15-20, src/main.cpp:290-322):
    protein.in, ligand.in      raw packed Atom{float x,y,z; int32 type} (16 B)
    forcefield.in              raw packed FFParams{int32 hbtype; float radius, hphb, elsc}
    poses.in                   6 rows of maxPoses float32 (tilt, roll, pan, tx, ty, tz)
    ref_energies.out           one float per line

Synthetic values come from libbude_mi355x's generator (bude_synth_*), which
maps std::mt19937 words through fixed integer/IEEE mappings, so a seed gives
the same bytes on every host.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import capi
from .capi import ATOM_DTYPE, FF_DTYPE

FILE_LIGAND = "ligand.in"
FILE_PROTEIN = "protein.in"
FILE_FORCEFIELD = "forcefield.in"
FILE_POSES = "poses.in"
FILE_REF_ENERGIES = "ref_energies.out"


@dataclass
class Deck:
    protein: np.ndarray
    ligand: np.ndarray
    forcefield: np.ndarray
    poses: np.ndarray                      # (6, nposes) float32, C-contiguous
    ref_energies: np.ndarray | None = None  # (>= nposes,) float32
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def natpro(self) -> int:
        return len(self.protein)

    @property
    def natlig(self) -> int:
        return len(self.ligand)

    @property
    def ntypes(self) -> int:
        return len(self.forcefield)

    @property
    def nposes(self) -> int:
        return int(self.poses.shape[1])

    @property
    def interactions(self) -> int:
        """Pair evaluations per pass (src/main.cpp:368)."""
        return self.nposes * self.natlig * self.natpro

    def head(self, n: int) -> "Deck":
        """The first n poses of every row, like the driver's -n (src/main.cpp:305-311)."""
        if n > self.nposes:
            raise ValueError(f"Requested poses ({n}) exceeded max poses ({self.nposes}) for deck")
        ref = None if self.ref_energies is None else self.ref_energies[:n]
        return Deck(self.protein, self.ligand, self.forcefield, np.ascontiguousarray(self.poses[:, :n]), ref,
                    self.name, dict(self.meta))


def read_deck(path: str | Path, nposes: int = 0, require_ref: bool = True) -> Deck:
    """Read a deck directory as the reference driver does (src/main.cpp:290-322):
    nposes == 0 means every pose in poses.in; the reference energies must cover
    the selected poses."""
    path = Path(path)
    for f in (FILE_LIGAND, FILE_PROTEIN, FILE_FORCEFIELD, FILE_POSES):
        if not (path / f).is_file():
            raise ValueError(f"Bad file: {path / f}")
    lig = np.fromfile(path / FILE_LIGAND, dtype=ATOM_DTYPE)
    pro = np.fromfile(path / FILE_PROTEIN, dtype=ATOM_DTYPE)
    ff = np.fromfile(path / FILE_FORCEFIELD, dtype=FF_DTYPE)
    raw = np.fromfile(path / FILE_POSES, dtype="<f4")
    if raw.size % 6 != 0:
        raise ValueError(f"Pose size ({raw.size}) not divisible my 6!")
    max_poses = raw.size // 6
    n = max_poses if nposes == 0 else nposes
    if n > max_poses:
        raise ValueError(f"Requested poses ({n}) exceeded max poses ({max_poses}) for deck")
    poses = np.ascontiguousarray(raw.reshape(6, max_poses)[:, :n])
    ref = None
    if (path / FILE_REF_ENERGIES).is_file():
        ref = np.loadtxt(path / FILE_REF_ENERGIES, dtype=np.float32, ndmin=1)
    if require_ref and (ref is None or ref.size < n):
        have = 0 if ref is None else ref.size
        raise ValueError(f"Size of reference energies ({have}) is less than poses ({n})")
    return Deck(pro, lig, ff, poses, ref, name=str(path), meta={"max_poses": max_poses})


def write_deck(path: str | Path, deck: Deck, ref_fmt: str = "%.9g") -> dict[str, str]:
    """Write a deck directory in the reference format; returns the SHA-256 of each file."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    deck.protein.astype(ATOM_DTYPE).tofile(path / FILE_PROTEIN)
    deck.ligand.astype(ATOM_DTYPE).tofile(path / FILE_LIGAND)
    deck.forcefield.astype(FF_DTYPE).tofile(path / FILE_FORCEFIELD)
    np.ascontiguousarray(deck.poses, dtype="<f4").tofile(path / FILE_POSES)
    if deck.ref_energies is not None:
        np.savetxt(path / FILE_REF_ENERGIES, deck.ref_energies, fmt=ref_fmt)
    return manifest(path)


def manifest(path: str | Path) -> dict[str, str]:
    path = Path(path)
    out = {}
    for f in (FILE_PROTEIN, FILE_LIGAND, FILE_FORCEFIELD, FILE_POSES, FILE_REF_ENERGIES):
        if (path / f).is_file():
            out[f] = hashlib.sha256((path / f).read_bytes()).hexdigest()
    return out


# ---------------------------------------------------------------------------
# synthetic generation
# ---------------------------------------------------------------------------

def synth_forcefield(seed: int, ntypes: int) -> np.ndarray:
    out = np.zeros(ntypes, FF_DTYPE)
    capi.check(capi.lib().bude_synth_forcefield(seed, ntypes, capi.ptr(out)))
    return out


def synth_protein(seed: int, natpro: int, ntypes: int, half_width: float = 20.0) -> np.ndarray:
    out = np.zeros(natpro, ATOM_DTYPE)
    capi.check(capi.lib().bude_synth_protein(seed, natpro, ntypes, half_width, capi.ptr(out)))
    return out


def synth_ligand(seed: int, natlig: int, ntypes: int, radius: float = 10.0) -> np.ndarray:
    out = np.zeros(natlig, ATOM_DTYPE)
    capi.check(capi.lib().bude_synth_ligand(seed, natlig, ntypes, radius, capi.ptr(out)))
    return out


def synth_poses(seed: int, count: int, first: int = 0, max_trans: float = 5.0) -> np.ndarray:
    out = np.zeros((6, count), np.float32)
    capi.check(capi.lib().bude_synth_poses(seed, first, count, max_trans, capi.ptr(out), count))
    return out


def synth_deck(natpro: int, natlig: int, nposes: int, seed: int, ntypes: int = 34, *, protein_seed: int | None = None,
               ligand_seed: int | None = None, pose_seed: int | None = None, name: str = "") -> Deck:
    """A deck with the benchmark distributions: protein uniform in a 40 A cube,
    ligand in a 10 A ball, types U{0..ntypes-1}, angles U[-pi,pi), translations
    U[-5,5) A.  Component seeds default to `seed`."""
    ps = seed if protein_seed is None else protein_seed
    ls = seed if ligand_seed is None else ligand_seed
    qs = seed if pose_seed is None else pose_seed
    return Deck(
        protein=synth_protein(ps, natpro, ntypes),
        ligand=synth_ligand(ls, natlig, ntypes),
        forcefield=synth_forcefield(ps, ntypes),
        poses=synth_poses(qs, nposes),
        name=name or f"synth-{natpro}x{natlig}x{nposes}-s{seed}",
        meta={"seed": seed, "protein_seed": ps, "ligand_seed": ls, "pose_seed": qs},
    )


# The plain-kernel workloads (all seeds fixed).
WORKLOADS = {
    # bm1-shaped CPU case: 938 protein atoms, 26 ligand atoms, 65,536 poses.
    "c0": dict(natpro=938, natlig=26, nposes=65536, seed=0),
    # single-GPU plain case: 4,096 / 64 / 1,048,576.
    "c1": dict(natpro=4096, natlig=64, nposes=1048576, seed=1),
}


def workload_deck(name: str, nposes: int | None = None) -> Deck:
    w = dict(WORKLOADS[name])
    if nposes is not None:
        w["nposes"] = nposes
    return synth_deck(**w, name=name)
