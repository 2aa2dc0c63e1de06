"""Combined screen (bude_mi355x_combined_* in the C ABI): a fixed ligand set,
each ligand with its own poses, screened against one protein after another.
Per protein a surrogate scores every pose of every (protein, ligand) pair,
each pair's lowest keep_fraction is rescored exactly in one ragged batched
launch, and pairs whose surrogate fails validation are scored exhaustively.

Surrogates: "mlp" (the learned per-atom field: one network per protein giving
a ligand atom's energy from its position and forcefield type -- a learned
feature grid, trilinear, and a two-layer ReLU MLP on BF16 MFMA -- summed
over a ligand's atoms, trained on exact one-atom energies), "field" (the
protein's single-ligand-atom energy maps on a grid, summed over a ligand's
atoms by trilinear interpolation) and "pose_mlp" (the learned prescreen's
pose network, one per protein, with 11 ligand-descriptor inputs, trained on
exact energies of sampled poses of every pair).  A pair whose validation
samples show too little room between the exact top fraction and the
selection boundary for the surrogate's error is scored exhaustively.
"""
from __future__ import annotations

import ctypes
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import capi
from . import deck as D

_c = ctypes
SURROGATES = {"pose_mlp": 0, "field": 1, "mlp": 2}


class CombinedConfig(_c.Structure):
    """BudeCombinedConfig (include/bude_mi355x.h)."""
    _fields_ = [("surrogate", _c.c_uint32), ("labels_per_pair", _c.c_uint32), ("keep_fraction", _c.c_float),
                ("max_val_rmse", _c.c_float), ("clip_quantile", _c.c_float), ("tail_quantile", _c.c_float),
                ("tail_weight", _c.c_float), ("batch", _c.c_uint32), ("epochs", _c.c_uint32), ("lr", _c.c_float),
                ("grid_spacing", _c.c_float), ("pipeline_chunks", _c.c_uint32), ("seed", _c.c_uint64),
                ("train_points", _c.c_uint32), ("steps", _c.c_uint32), ("grid_lr", _c.c_float), ("gate_z", _c.c_float),
                ("gate_top_fraction", _c.c_float), ("widen_max", _c.c_float)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class CombinedResult(_c.Structure):
    _fields_ = [("total_ms", _c.c_double), ("prepare_ms", _c.c_float), ("label_ms", _c.c_float),
                ("train_ms", _c.c_float), ("score_ms", _c.c_float), ("select_ms", _c.c_float),
                ("rescore_ms", _c.c_float), ("nposes", _c.c_uint64), ("nselected", _c.c_uint64),
                ("labels", _c.c_uint64), ("fallback_pairs", _c.c_uint32), ("surrogate_rmse", _c.c_float),
                ("widened_pairs", _c.c_uint32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class PairResult(_c.Structure):
    _fields_ = [("nposes", _c.c_uint32), ("nselected", _c.c_uint32), ("fallback", _c.c_int32),
                ("val_rmse", _c.c_float), ("val_rmse_calibrated", _c.c_float), ("gate_z", _c.c_float),
                ("sample_recall", _c.c_float), ("keep_fraction", _c.c_float)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def default_config(surrogate: str = "field", **kw) -> CombinedConfig:
    """Defaults per surrogate: the learned per-atom field ("mlp") trains for
    3,000 Adam steps of 32,768 on 2^22 exact one-atom energies with a 0.75 A
    feature grid and validates on 4,096 sampled poses per pair; the field's
    maps are 0.5 A and validate on 1,024; the pose MLP trains on 16,384
    sampled poses per pair (1/16 held out).  Every kind trusts a pair where
    its validation gap reaches 2.5 standard deviations of its tail error
    (top 1 %, keep 2 %: a modelled loss of 0.08 % of the top 1 %; the full-
    size screen at 2 left one pair of 512 just under 0.99), widens a pair
    short of that to the fraction that reaches it, up to 10 %, and scores a
    pair beyond that exhaustively; the pose MLP also falls back where its
    calibrated validation RMSE exceeds 0.8."""
    mlp = surrogate == "mlp"
    c = dict(surrogate=SURROGATES[surrogate], labels_per_pair={"pose_mlp": 16384, "mlp": 4096}.get(surrogate, 1024),
             keep_fraction=0.02, max_val_rmse=0.8 if surrogate == "pose_mlp" else float("inf"), clip_quantile=0.1,
             tail_quantile=0.05, tail_weight=8.0, batch=32768 if mlp else 8192, epochs=20, lr=3e-3 if mlp else 4e-3,
             grid_spacing=0.75 if mlp else 0.5, pipeline_chunks=0, seed=0, train_points=(1 << 22) if mlp else 1 << 21,
             steps=3000, grid_lr=1.5e-2, gate_z=2.5, gate_top_fraction=0.01, widen_max=0.1)
    c.update(kw)
    return CombinedConfig(**c)


def _copy_config(cfg: CombinedConfig) -> CombinedConfig:
    """A copy (the object keeps its own configuration)."""
    return CombinedConfig(**cfg.as_dict())


_P, _S = _c.c_void_p, _c.c_size_t
_DONE_FN = _c.CFUNCTYPE(None, _c.c_void_p, _c.c_size_t)
_SIG = {
    "bude_mi355x_combined_create": [_c.c_int, _P, _S, _S, _c.POINTER(_P), _c.POINTER(_S), _c.POINTER(_P),
                                    _c.POINTER(_S), _P, _c.POINTER(_P)],
    "bude_mi355x_combined_screen": [_P, _P, _S, _P],
    "bude_mi355x_combined_screen_batch": [_P, _S, _c.POINTER(_P), _c.POINTER(_S), _P, _P, _P,
                                          _c.POINTER(_c.c_double)],
    "bude_mi355x_combined_set_keep": [_P, _c.c_float],
    "bude_mi355x_combined_pair": [_P, _S, _P],
    "bude_mi355x_combined_selected": [_P, _S, _P, _P, _S, _c.POINTER(_S)],
    "bude_mi355x_combined_labels": [_P, _S, _P, _P, _S, _c.POINTER(_S)],
    "bude_mi355x_combined_scores": [_P, _S, _P],
    "bude_mi355x_combined_exhaustive": [_P, _c.POINTER(_c.c_double)],
    "bude_mi355x_combined_exact": [_P, _S, _P],
    "bude_mi355x_combined_recall": [_P, _c.c_float, _P],
    "bude_mi355x_combined_grid": [_P, _P, _c.POINTER(_c.c_float), _P],
    "bude_mi355x_combined_map": [_P, _c.c_int32, _P],
    "bude_mi355x_combined_destroy": [_P],
}


def _lib():
    lib = capi.lib()
    if not getattr(lib, "_combined_declared", False):
        for name, args in _SIG.items():
            fn = getattr(lib, name)
            fn.restype = _c.c_int
            fn.argtypes = args
        lib._combined_declared = True
    return lib


class Combined:
    """A ligand set with its poses resident in HBM, screened per protein."""

    def __init__(self, forcefield, ligands, poses, cfg: CombinedConfig | None = None, device: int = 0):
        self.cfg = _copy_config(cfg or default_config())
        self._h = _c.c_void_p()
        n = len(ligands)
        self._ff = np.ascontiguousarray(forcefield)
        self._ligs = [np.ascontiguousarray(l) for l in ligands]
        self._poses = [np.ascontiguousarray(p, dtype=np.float32) for p in poses]
        self.nposes = [p.shape[1] for p in self._poses]
        lp = (_P * n)(*[l.ctypes.data for l in self._ligs])
        nl = (_S * n)(*[len(l) for l in self._ligs])
        pp = (_P * n)(*[p.ctypes.data for p in self._poses])
        npz = (_S * n)(*self.nposes)
        capi.check(_lib().bude_mi355x_combined_create(device, capi.ptr(self._ff), len(self._ff), n, lp, nl, pp, npz,
                                                      _c.byref(self.cfg), _c.byref(self._h)))
        self.nligands = n
        self._poses = None  # resident on the device now

    def screen(self, protein) -> CombinedResult:
        pro = np.ascontiguousarray(protein)
        r = CombinedResult()
        capi.check(_lib().bude_mi355x_combined_screen(self._h, capi.ptr(pro), len(pro), _c.byref(r)))
        return r

    def screen_batch(self, proteins, done=None) -> tuple[list[CombinedResult], float]:
        """Screens every protein in order, each one's scoring beside the
        previous one's rescoring (bude_mi355x_combined_screen_batch).
        done(i), if given, runs as soon as protein i is complete, with
        protein i as the protein screened last for selected(), pair(),
        labels(), exhaustive(), exact() and recall().  Returns the results
        and the host time of the whole call in ms."""
        pros = [np.ascontiguousarray(p) for p in proteins]
        n = len(pros)
        ptrs = (_P * n)(*[p.ctypes.data for p in pros])
        counts = (_S * n)(*[len(p) for p in pros])
        res = (CombinedResult * n)()
        wall = _c.c_double(0.0)
        errors = []

        def trampoline(user, i):
            try:
                done(int(i))
            except BaseException as exc:  # (cannot propagate through C)
                errors.append(exc)

        cb = _DONE_FN(trampoline) if done is not None else None
        capi.check(_lib().bude_mi355x_combined_screen_batch(self._h, n, ptrs, counts, _c.cast(res, _P),
                                                            _c.cast(cb, _P) if cb else None, None,
                                                            _c.byref(wall)))
        if errors:
            raise errors[0]
        return list(res), wall.value

    def set_keep(self, keep_fraction: float) -> None:
        capi.check(_lib().bude_mi355x_combined_set_keep(self._h, keep_fraction))
        self.cfg.keep_fraction = keep_fraction

    def pair(self, ligand: int) -> PairResult:
        r = PairResult()
        capi.check(_lib().bude_mi355x_combined_pair(self._h, ligand, _c.byref(r)))
        return r

    def labels(self, ligand: int) -> tuple[np.ndarray, np.ndarray]:
        """The last screen's sampled poses of one pair and their exact energies."""
        n = _c.c_size_t()
        capi.check(_lib().bude_mi355x_combined_labels(self._h, ligand, None, None, 0, _c.byref(n)))
        idx = np.zeros(n.value, np.uint32)
        e = np.zeros(n.value, np.float32)
        capi.check(_lib().bude_mi355x_combined_labels(self._h, ligand, capi.ptr(idx), capi.ptr(e), n.value,
                                                      _c.byref(n)))
        return idx, e

    def selected(self, ligand: int) -> tuple[np.ndarray, np.ndarray]:
        n = _c.c_size_t()
        capi.check(_lib().bude_mi355x_combined_selected(self._h, ligand, None, None, 0, _c.byref(n)))
        idx = np.zeros(n.value, np.uint32)
        e = np.zeros(n.value, np.float32)
        capi.check(_lib().bude_mi355x_combined_selected(self._h, ligand, capi.ptr(idx), capi.ptr(e), n.value,
                                                        _c.byref(n)))
        return idx, e

    def scores(self, ligand: int) -> np.ndarray:
        out = np.zeros(self.nposes[ligand], np.float32)
        capi.check(_lib().bude_mi355x_combined_scores(self._h, ligand, capi.ptr(out)))
        return out

    def exhaustive(self) -> float:
        ms = _c.c_double()
        capi.check(_lib().bude_mi355x_combined_exhaustive(self._h, _c.byref(ms)))
        return ms.value

    def exact(self, ligand: int) -> np.ndarray:
        out = np.zeros(self.nposes[ligand], np.float32)
        capi.check(_lib().bude_mi355x_combined_exact(self._h, ligand, capi.ptr(out)))
        return out

    def recall(self, top_fraction: float = 0.01) -> np.ndarray:
        out = np.zeros(self.nligands, np.float32)
        capi.check(_lib().bude_mi355x_combined_recall(self._h, top_fraction, capi.ptr(out)))
        return out

    def grid(self) -> tuple[np.ndarray, float, tuple[int, int, int]]:
        lo = np.zeros(3, np.float32)
        n = np.zeros(3, np.int32)
        h = _c.c_float()
        capi.check(_lib().bude_mi355x_combined_grid(self._h, capi.ptr(lo), _c.byref(h), capi.ptr(n)))
        return lo, h.value, tuple(int(v) for v in n)

    def map(self, ff_type: int) -> np.ndarray:
        _, _, n = self.grid()
        out = np.zeros(n, np.float32)
        capi.check(_lib().bude_mi355x_combined_map(self._h, ff_type, capi.ptr(out)))
        return out

    def close(self) -> None:
        if self._h:
            capi.check(_lib().bude_mi355x_combined_destroy(self._h))
            self._h = _c.c_void_p()

    def __enter__(self) -> "Combined":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


@dataclass
class CombinedWorkload:
    forcefield: np.ndarray
    proteins: list[np.ndarray]
    ligands: list[np.ndarray]
    poses: list[np.ndarray]
    name: str = ""

    @property
    def pairs(self) -> int:
        return len(self.proteins) * len(self.ligands)

    @property
    def total_poses(self) -> int:
        return len(self.proteins) * int(sum(p.shape[1] for p in self.poses))

    @property
    def interactions(self) -> int:
        per_protein = sum(len(l) * p.shape[1] for l, p in zip(self.ligands, self.poses))
        return int(sum(len(p) for p in self.proteins)) * int(per_protein)


def combined_workload(nproteins: int = 16, nligands: int = 512, nposes: int = 1 << 20, seed: int = 4,
                      natpro_range=(900, 6000), natlig_range=(8, 96), ntypes: int = 34,
                      threads: int = 8) -> CombinedWorkload:
    """The combined screen: nproteins proteins of U{natpro_range} atoms and
    nligands ligands of U{natlig_range} atoms (both inclusive, from numpy's
    legacy MT19937 RandomState(seed)), each ligand with nposes poses, every
    distribution as the plain workload decks.  Protein p uses generator seed
    seed * 1000003 + p + 1, ligand i and its poses seed * 100003 + i + 1; a
    ligand's poses are the same for every protein."""
    rs = np.random.RandomState(seed)
    natpro = rs.randint(natpro_range[0], natpro_range[1] + 1, size=nproteins)
    natlig = rs.randint(natlig_range[0], natlig_range[1] + 1, size=nligands)
    ff = D.synth_forcefield(seed, ntypes)
    proteins = [D.synth_protein(seed * 1000003 + p + 1, int(natpro[p]), ntypes) for p in range(nproteins)]
    ligands = [D.synth_ligand(seed * 100003 + i + 1, int(natlig[i]), ntypes) for i in range(nligands)]
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        poses = list(ex.map(lambda i: D.synth_poses(seed * 100003 + i + 1, nposes), range(nligands)))
    return CombinedWorkload(ff, proteins, ligands, poses,
                            name=f"combined-{nproteins}x{nligands}x{nposes}-s{seed}")
