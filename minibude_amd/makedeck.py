"""Decks from real molecules: a BUDE bhff forcefield, a protein and a ligand
mol2 file and a seeded pose grid, built as the reference's makedeck program
builds them (makedeck/main.cpp:49-234, makedeck/input-utils.cpp:72-262).

The parsers and the pose sampler are libbude_mi355x host functions
(bude_forcefield_*, bude_mol2_read, bude_poses_generate); the reference
energies of a new deck come from the library's plain kernel on the GPU
(`reference_energies`), where makedeck ran its CPU kernel.  The command-line
program with makedeck's options is minibude_amd/bin/bude-makedeck.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import capi
from . import deck as D
from .capi import ATOM_DTYPE, FF_DTYPE

# makedeck's pose grid (makedeck/main.cpp:12-23): 36 angles -170..180 deg in
# 10 deg steps for tilt, roll and pan, 15 translations -7..7 A, seed 42,
# 65,536 poses.
GRID_DEGREES = np.arange(-170, 190, 10, dtype=np.float32)
GRID_TRANSLATIONS = np.arange(-7, 8, dtype=np.float32)
DEFAULT_POSE_SEED = 42
DEFAULT_POSES = 65536
POSE_MULTIPLE = 256  # makedeck requires the pose count to divide by its work-group size


class Forcefield:
    """A parsed bhff file (bude_forcefield_read)."""

    def __init__(self, path: str | Path):
        self.path = str(path)
        self._h = ctypes.c_void_p()
        capi.check(capi.lib().bude_forcefield_read(self.path.encode(), ctypes.byref(self._h)))

    @property
    def nresidues(self) -> int:
        r, e = ctypes.c_size_t(), ctypes.c_size_t()
        capi.check(capi.lib().bude_forcefield_info(self._h, ctypes.byref(r), ctypes.byref(e)))
        return r.value

    def params(self) -> np.ndarray:
        """The deck's forcefield rows (FFParams records)."""
        n = ctypes.c_size_t()
        capi.check(capi.lib().bude_forcefield_params(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, FF_DTYPE)
        capi.check(capi.lib().bude_forcefield_params(self._h, capi.ptr(out), n.value, ctypes.byref(n)))
        return out

    def molecule(self, mol2: str | Path) -> tuple[np.ndarray, int]:
        """Heavy atoms of a mol2 file typed against this forcefield, and the
        number of @<BUDE>CONF conformations the file carries."""
        n, nconf = ctypes.c_size_t(), ctypes.c_size_t()
        path = str(mol2).encode()
        capi.check(capi.lib().bude_mol2_read(self._h, path, None, 0, ctypes.byref(n), ctypes.byref(nconf)))
        out = np.zeros(n.value, ATOM_DTYPE)
        capi.check(capi.lib().bude_mol2_read(self._h, path, capi.ptr(out), n.value, ctypes.byref(n),
                                             ctypes.byref(nconf)))
        return out, nconf.value

    def close(self) -> None:
        if self._h:
            capi.lib().bude_forcefield_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self) -> "Forcefield":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def generate_poses(nposes: int = DEFAULT_POSES, seed: int = DEFAULT_POSE_SEED,
                   grid: list[np.ndarray] | None = None) -> np.ndarray:
    """(6, nposes) float32 poses from the grid (angles in degrees in, radians
    out), as makedeck's generatePoses draws them (bude_poses_generate)."""
    if grid is None:
        grid = [GRID_DEGREES] * 3 + [GRID_TRANSLATIONS] * 3
    vals = [np.ascontiguousarray(v, dtype=np.float32) for v in grid]
    if len(vals) != 6:
        raise ValueError("a pose grid has 6 fields")
    out = np.zeros((6, nposes), np.float32)
    vp = (ctypes.c_void_p * 6)(*[v.ctypes.data for v in vals])
    cn = (ctypes.c_size_t * 6)(*[len(v) for v in vals])
    rp = (ctypes.c_void_p * 6)(*[out[k].ctypes.data for k in range(6)])
    capi.check(capi.lib().bude_poses_generate(nposes, seed, vp, cn, rp))
    return out


def build_deck(bhff: str | Path, protein_mol2: str | Path, ligand_mol2: str | Path,
               nposes: int = DEFAULT_POSES, seed: int = DEFAULT_POSE_SEED) -> D.Deck:
    """The deck makedeck writes for these inputs, without reference energies."""
    with Forcefield(bhff) as ff:
        params = ff.params()
        protein, _ = ff.molecule(protein_mol2)
        ligand, _ = ff.molecule(ligand_mol2)
    return D.Deck(protein, ligand, params, generate_poses(nposes, seed),
                  name=f"{Path(protein_mol2).stem}+{Path(ligand_mol2).stem}",
                  meta={"pose_seed": seed, "forcefield": str(bhff)})


def reference_energies(deck: D.Deck, device: int = 0) -> np.ndarray:
    """Energies of every pose of the deck from the plain kernel on `device`
    (makedeck ran its CPU reference kernel here, makedeck/main.cpp:170-194)."""
    from .model import Context
    with Context(deck, device) as c:
        return c.fasten(1, 256).energies
