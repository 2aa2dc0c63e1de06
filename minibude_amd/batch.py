"""Batched ragged screening: many ligands (own atom and pose counts) against
one protein in one persistent launch (bude_mi355x_batch_* in the C ABI).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi
from . import deck as D
from .capi import BatchConfig


@dataclass
class LigandSet:
    """One protein and its forcefield, plus ligands[i] with poses[i] (6, n_i)."""
    protein: np.ndarray
    forcefield: np.ndarray
    ligands: list[np.ndarray]
    poses: list[np.ndarray]
    name: str = ""

    @property
    def natpro(self) -> int:
        return len(self.protein)

    @property
    def interactions(self) -> int:
        return int(sum(len(l) * p.shape[1] for l, p in zip(self.ligands, self.poses))) * self.natpro

    @property
    def total_poses(self) -> int:
        return int(sum(p.shape[1] for p in self.poses))

    def deck(self, i: int) -> D.Deck:
        return D.Deck(self.protein, self.ligands[i], self.forcefield, self.poses[i], name=f"{self.name}[{i}]")


def ragged_workload(nligands: int = 256, natpro: int = 4096, seed: int = 2, natlig_range=(8, 96),
                    nposes_range=(4096, 65536), ntypes: int = 34, pose_scale: float = 1.0) -> LigandSet:
    """The ragged screen: one protein of `natpro` atoms and `nligands` ligands
    with atom counts U{natlig_range} and pose counts U{nposes_range} (both
    inclusive), all distributions as the plain workload decks.  Sizes come
    from numpy's legacy MT19937 RandomState (stable across numpy versions);
    ligand i and its poses use generator seed `seed * 100003 + i + 1`.
    pose_scale shrinks every pose count (for quick tests)."""
    rs = np.random.RandomState(seed)
    natlig = rs.randint(natlig_range[0], natlig_range[1] + 1, size=nligands)
    nposes = rs.randint(nposes_range[0], nposes_range[1] + 1, size=nligands)
    if pose_scale != 1.0:
        nposes = np.maximum(1, (nposes * pose_scale).astype(np.int64))
    protein = D.synth_protein(seed, natpro, ntypes)
    ff = D.synth_forcefield(seed, ntypes)
    ligs, poses = [], []
    for i in range(nligands):
        s = seed * 100003 + i + 1
        ligs.append(D.synth_ligand(s, int(natlig[i]), ntypes))
        poses.append(D.synth_poses(s, int(nposes[i])))
    return LigandSet(protein, ff, ligs, poses, name=f"ragged-{nligands}x{natpro}-s{seed}")


class Batch:
    """A LigandSet resident in HBM (bude_mi355x_batch_create)."""

    def __init__(self, ls: LigandSet, device: int = 0):
        self.ls = ls
        self._h = ctypes.c_void_p()
        n = len(ls.ligands)
        self._ligs = [np.ascontiguousarray(l) for l in ls.ligands]
        self._poses = [np.ascontiguousarray(p, dtype=np.float32) for p in ls.poses]
        lp = (ctypes.c_void_p * n)(*[l.ctypes.data for l in self._ligs])
        nl = (ctypes.c_size_t * n)(*[len(l) for l in self._ligs])
        pp = (ctypes.c_void_p * n)(*[p.ctypes.data for p in self._poses])
        npz = (ctypes.c_size_t * n)(*[p.shape[1] for p in self._poses])
        pro = np.ascontiguousarray(ls.protein)
        ff = np.ascontiguousarray(ls.forcefield)
        capi.check(capi.lib().bude_mi355x_batch_create(device, capi.ptr(pro), len(pro), capi.ptr(ff), len(ff), n,
                                                       lp, nl, pp, npz, ctypes.byref(self._h)))

    def run(self, cfg: BatchConfig, warmup: int = 0, iterations: int = 1) -> list[float]:
        ms = np.zeros(warmup + iterations, np.float64)
        capi.check(capi.lib().bude_mi355x_batch_run(self._h, ctypes.byref(cfg), warmup, iterations, capi.ptr(ms)))
        return ms.tolist()

    def run_separate(self, ppwi: int, wgsize: int, warmup: int = 0, iterations: int = 1) -> list[float]:
        ms = np.zeros(warmup + iterations, np.float64)
        capi.check(capi.lib().bude_mi355x_batch_run_separate(self._h, ppwi, wgsize, warmup, iterations,
                                                             capi.ptr(ms)))
        return ms.tolist()

    def autotune(self) -> tuple[BatchConfig, float]:
        cfg = BatchConfig()
        ms = ctypes.c_double()
        capi.check(capi.lib().bude_mi355x_batch_autotune(self._h, ctypes.byref(cfg), ctypes.byref(ms)))
        return cfg, ms.value

    def energies(self, i: int) -> np.ndarray:
        out = np.zeros(self._poses[i].shape[1], np.float32)
        capi.check(capi.lib().bude_mi355x_batch_energies(self._h, i, capi.ptr(out)))
        return out

    def clear(self) -> None:
        capi.check(capi.lib().bude_mi355x_batch_clear(self._h))

    def close(self) -> None:
        if self._h:
            capi.check(capi.lib().bude_mi355x_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self) -> "Batch":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass
