"""Learned prescreen (bude_mi355x_screen_* in the C ABI): an MLP surrogate
of the pose energy, trained on the GPU on plain-kernel energies, prescreens a
pose set; the lowest keep_fraction of its predictions is rescored exactly.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi
from . import deck as D

_c = ctypes


class ScreenConfig(_c.Structure):
    _fields_ = [("batch", _c.c_uint32), ("epochs", _c.c_uint32), ("lr", _c.c_float), ("clip_quantile", _c.c_float),
                ("keep_fraction", _c.c_float), ("max_val_rmse", _c.c_float), ("tail_quantile", _c.c_float),
                ("tail_weight", _c.c_float), ("seed", _c.c_uint64), ("features", _c.c_uint32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class TrainStats(_c.Structure):
    _fields_ = [("label_ms", _c.c_double), ("train_ms", _c.c_double), ("final_loss", _c.c_float),
                ("val_rmse", _c.c_float), ("clip", _c.c_float), ("mu", _c.c_float), ("sigma", _c.c_float),
                ("steps", _c.c_uint32), ("fallback", _c.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ScreenResult(_c.Structure):
    _fields_ = [("total_ms", _c.c_double), ("score_ms", _c.c_float), ("select_ms", _c.c_float),
                ("rescore_ms", _c.c_float), ("nposes", _c.c_uint64), ("nselected", _c.c_uint64),
                ("fallback", _c.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def default_config(**kw) -> ScreenConfig:
    c = dict(batch=8192, epochs=60, lr=4e-3, clip_quantile=0.1, keep_fraction=0.02, max_val_rmse=0.8,
             tail_quantile=0.05, tail_weight=8.0, seed=0, features=21)
    c.update(kw)
    return ScreenConfig(**c)


_SIG = {
    "bude_mi355x_screen_create": [_c.c_int, _c.c_void_p, _c.c_size_t, _c.c_void_p, _c.c_size_t, _c.c_void_p,
                                  _c.c_size_t, _c.c_void_p, _c.POINTER(_c.c_void_p)],
    "bude_mi355x_screen_train": [_c.c_void_p, _c.POINTER(_c.c_void_p), _c.c_size_t, _c.c_void_p, _c.c_void_p,
                                 _c.c_size_t],
    "bude_mi355x_screen_load": [_c.c_void_p, _c.POINTER(_c.c_void_p), _c.c_size_t],
    "bude_mi355x_screen_run": [_c.c_void_p, _c.c_void_p],
    "bude_mi355x_screen_set_keep": [_c.c_void_p, _c.c_float],
    "bude_mi355x_screen_selected": [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_size_t, _c.POINTER(_c.c_size_t)],
    "bude_mi355x_screen_predict": [_c.c_void_p, _c.c_void_p],
    "bude_mi355x_screen_exhaustive": [_c.c_void_p, _c.c_void_p, _c.POINTER(_c.c_double)],
    "bude_mi355x_screen_destroy": [_c.c_void_p],
    "bude_mi355x_mlp_predict": [_c.c_int, _c.POINTER(_c.c_void_p), _c.c_size_t, _c.POINTER(_c.c_void_p),
                                _c.POINTER(_c.c_void_p), _c.c_void_p],
    "bude_mi355x_mlp_forward_backward": [_c.c_int, _c.c_size_t, _c.c_void_p, _c.c_void_p, _c.POINTER(_c.c_void_p),
                                         _c.POINTER(_c.c_void_p), _c.c_void_p, _c.POINTER(_c.c_float),
                                         _c.POINTER(_c.c_void_p), _c.POINTER(_c.c_void_p)],
}


def _lib():
    lib = capi.lib()
    if not getattr(lib, "_screen_declared", False):
        for name, args in _SIG.items():
            fn = getattr(lib, name)
            fn.restype = _c.c_int
            fn.argtypes = args
        lib._screen_declared = True
    return lib


def _rows(poses: np.ndarray):
    p = np.ascontiguousarray(poses, dtype=np.float32)
    return p, (_c.c_void_p * 6)(*[p[k].ctypes.data for k in range(6)])


class Screen:
    def __init__(self, protein, ligand, forcefield, cfg: ScreenConfig | None = None, device: int = 0):
        self.cfg = cfg or default_config()
        self._h = _c.c_void_p()
        self._keep = [np.ascontiguousarray(a) for a in (protein, ligand, forcefield)]
        pro, lig, ff = self._keep
        capi.check(_lib().bude_mi355x_screen_create(device, capi.ptr(pro), len(pro), capi.ptr(lig), len(lig),
                                                    capi.ptr(ff), len(ff), _c.byref(self.cfg), _c.byref(self._h)))

    def train(self, poses: np.ndarray) -> tuple[TrainStats, np.ndarray]:
        p, rows = _rows(poses)
        st = TrainStats()
        curve = np.zeros(max(1, self.cfg.epochs), np.float32)
        capi.check(_lib().bude_mi355x_screen_train(self._h, rows, p.shape[1], _c.byref(st), capi.ptr(curve),
                                                   len(curve)))
        return st, curve

    def load(self, poses: np.ndarray) -> None:
        p, rows = _rows(poses)
        capi.check(_lib().bude_mi355x_screen_load(self._h, rows, p.shape[1]))
        self.n = p.shape[1]

    def run(self) -> ScreenResult:
        r = ScreenResult()
        capi.check(_lib().bude_mi355x_screen_run(self._h, _c.byref(r)))
        return r

    def set_keep(self, keep_fraction: float) -> None:
        capi.check(_lib().bude_mi355x_screen_set_keep(self._h, keep_fraction))
        self.cfg.keep_fraction = keep_fraction

    def selected(self) -> tuple[np.ndarray, np.ndarray]:
        n = _c.c_size_t()
        capi.check(_lib().bude_mi355x_screen_selected(self._h, None, None, 0, _c.byref(n)))
        idx = np.zeros(n.value, np.uint32)
        e = np.zeros(n.value, np.float32)
        capi.check(_lib().bude_mi355x_screen_selected(self._h, capi.ptr(idx), capi.ptr(e), n.value, _c.byref(n)))
        return idx, e

    def predict(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        capi.check(_lib().bude_mi355x_screen_predict(self._h, capi.ptr(out)))
        return out

    def exhaustive(self) -> tuple[np.ndarray, float]:
        out = np.zeros(self.n, np.float32)
        ms = _c.c_double()
        capi.check(_lib().bude_mi355x_screen_exhaustive(self._h, capi.ptr(out), _c.byref(ms)))
        return out, ms.value

    def close(self) -> None:
        if self._h:
            capi.check(_lib().bude_mi355x_screen_destroy(self._h))
            self._h = _c.c_void_p()

    def __enter__(self) -> "Screen":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def mlp_forward_backward(feat, target, W, b, device: int = 0):
    """GPU forward + backward of the surrogate with the given weights (see
    bude_mi355x_mlp_forward_backward)."""
    feat = np.ascontiguousarray(feat, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    Wc = [np.ascontiguousarray(w, np.float32) for w in W]
    bc = [np.ascontiguousarray(x, np.float32) for x in b]
    gW = [np.zeros_like(w) for w in Wc]
    gb = [np.zeros_like(x) for x in bc]
    pred = np.zeros(len(feat), np.float32)
    loss = _c.c_float()
    P = lambda arrs: (_c.c_void_p * 4)(*[a.ctypes.data for a in arrs])  # noqa: E731
    capi.check(_lib().bude_mi355x_mlp_forward_backward(device, len(feat), capi.ptr(feat), capi.ptr(target), P(Wc),
                                                       P(bc), capi.ptr(pred), _c.byref(loss), P(gW), P(gb)))
    return pred, loss.value, gW, gb


def mlp_predict(poses, W, b, device: int = 0) -> np.ndarray:
    """The fused inference kernel with the given weights (bude_mi355x_mlp_predict)."""
    p, rows = _rows(poses)
    Wc = [np.ascontiguousarray(w, np.float32) for w in W]
    bc = [np.ascontiguousarray(x, np.float32) for x in b]
    P = lambda arrs: (_c.c_void_p * 4)(*[a.ctypes.data for a in arrs])  # noqa: E731
    out = np.zeros(p.shape[1], np.float32)
    capi.check(_lib().bude_mi355x_mlp_predict(device, rows, p.shape[1], P(Wc), P(bc), capi.ptr(out)))
    return out


def features(poses: np.ndarray) -> np.ndarray:
    """The 21 surrogate inputs of each pose: angles / pi, translations / 5,
    sin and cos of the three angles, and the 9 rotation-matrix entries
    (R = Rz(pan) Ry(roll) Rx(tilt), row-major)."""
    a = poses[:3].astype(np.float64)
    t = poses[3:].astype(np.float64)
    sx, sy, sz = np.sin(a)
    cx, cy, cz = np.cos(a)
    R = [cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz, cy * sz, sx * sy * sz + cx * cz,
         cx * sy * sz - sx * cz, -sy, sx * cy, cx * cy]
    return np.concatenate([a / np.pi, t / 5.0, np.sin(a), np.cos(a), np.array(R)]).T


def recall(selected_idx: np.ndarray, exact: np.ndarray, top_frac: float = 0.01) -> float:
    """Fraction of the exact top_frac lowest-energy poses that were selected."""
    k = max(1, int(round(top_frac * len(exact))))
    true = np.argpartition(exact, k - 1)[:k]
    return float(np.isin(true, selected_idx).sum()) / k


@dataclass
class ScreenWorkload:
    protein: np.ndarray
    ligand: np.ndarray
    forcefield: np.ndarray
    train_poses: np.ndarray
    screen_poses: np.ndarray


def screen_workload(ntrain: int = 8 << 20, nscreen: int = 16 << 20, natlig: int = 26, seed: int = 3,
                    protein_seed: int = 1, natpro: int = 4096) -> ScreenWorkload:
    """The learned-prescreen workload: the single-GPU deck's protein (4,096
    atoms, seed 1) and forcefield, a fresh 26-atom ligand, training poses and
    a disjoint set of fresh screening poses (different pose streams)."""
    pro = D.synth_protein(protein_seed, natpro, 34)
    ff = D.synth_forcefield(protein_seed, 34)
    lig = D.synth_ligand(seed, natlig, 34)
    return ScreenWorkload(pro, lig, ff, D.synth_poses(seed * 7919 + 1, ntrain), D.synth_poses(seed * 7919 + 2, nscreen))
