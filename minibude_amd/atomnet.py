"""The learned per-atom field's parts through the C ABI (include/bude_mi355x.h):
exact one-atom energies (its training labels), one forward + backward pass
of its training kernels, and its fused scorer over one ligand's poses.

The network gives one ligand atom of forcefield slot s at position p the
energy sigma * y + mu, y = MLP([trilinear(grid, p), onehot(s)]) (two ReLU
layers of HIDDEN, BF16 weights on MFMA); a pose scores the sum over its
atoms.  The combined screen trains one per protein (surrogate "mlp" of
minibude_amd.combined)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi

FEATURES = 8
HIDDEN = 32


def input_width(nslots: int) -> int:
    return 32 * ((FEATURES + nslots + 31) // 32)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def atom_energies(protein: np.ndarray, forcefield: np.ndarray, points: np.ndarray, types: np.ndarray,
                  device: int = 0) -> np.ndarray:
    """Exact energy of one ligand atom of types[i] at points[i] (n, 3)."""
    pts = _f32(points).reshape(-1, 3)
    ty = np.ascontiguousarray(types, np.int32)
    out = np.zeros(len(pts), np.float32)
    prot = np.ascontiguousarray(protein)
    ff = np.ascontiguousarray(forcefield)
    capi.check(capi.lib().bude_mi355x_atom_energies(device, capi.ptr(prot), len(prot), capi.ptr(ff), len(ff),
                                                    capi.ptr(pts), capi.ptr(ty), len(pts), capi.ptr(out)))
    return out


def _grid_args(lo, h, n, grid):
    return _f32(lo), ctypes.c_float(h), np.ascontiguousarray(n, np.int32), _f32(grid)


def forward_backward(points, slots, target, lo, h, n, grid, nslots, W, b, device: int = 0):
    """One training forward + backward pass: (pred, loss, grad_grid, grad_W, grad_b).
    W = [W0 (input_width(nslots), HIDDEN), W1 (HIDDEN, HIDDEN), W2 (HIDDEN, 1)]."""
    pts = _f32(points).reshape(-1, 3)
    B = len(pts)
    sl = np.ascontiguousarray(slots, np.int32)
    tg = _f32(target)
    lo_, h_, n_, g_ = _grid_args(lo, h, n, grid)
    Ws = [_f32(w) for w in W]
    bs = [_f32(x) for x in b]
    pred = np.zeros(B, np.float32)
    loss = ctypes.c_float()
    gg = np.zeros_like(g_)
    gW = [np.zeros_like(w) for w in Ws]
    gb = [np.zeros_like(x) for x in bs]
    capi.check(capi.lib().bude_mi355x_atomnet_forward_backward(
        device, B, capi.ptr(pts), capi.ptr(sl), capi.ptr(tg), capi.ptr(lo_), h_, capi.ptr(n_), capi.ptr(g_), nslots,
        _ptrs(Ws), _ptrs(bs), capi.ptr(pred), ctypes.byref(loss), capi.ptr(gg), _ptrs(gW), _ptrs(gb)))
    return pred, loss.value, gg, gW, gb


def score(ligand: np.ndarray, poses: np.ndarray, lo, h, n, grid, nslots, W, b, sigma: float, mu: float,
          device: int = 0) -> np.ndarray:
    """The fused scorer over every pose (6, n) of one ligand whose type fields are slots."""
    lig = np.ascontiguousarray(ligand)
    p = _f32(poses)
    rows = [np.ascontiguousarray(p[k]) for k in range(6)]
    lo_, h_, n_, g_ = _grid_args(lo, h, n, grid)
    Ws = [_f32(w) for w in W]
    bs = [_f32(x) for x in b]
    out = np.zeros(p.shape[1], np.float32)
    capi.check(capi.lib().bude_mi355x_atomnet_score(
        device, capi.ptr(lig), len(lig), _ptrs(rows), p.shape[1], capi.ptr(lo_), h_, capi.ptr(n_), capi.ptr(g_),
        nslots, _ptrs(Ws), _ptrs(bs), ctypes.c_float(sigma), ctypes.c_float(mu), capi.ptr(out)))
    return out
