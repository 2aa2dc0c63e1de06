"""CPU restatement of the learned per-atom field (numpy, FP64) -- test
infrastructure only, the checker of the GPU kernels; the product never
imports it.

The surrogate gives one ligand atom of forcefield slot s at position p the
energy  E(p, s) = sigma * y(p, s) + mu  with

    phi(p)   = trilinear interpolation of the feature grid G [n0][n1][n2][F]
               (cell and fraction clamped to the grid)
    x0       = [phi(p), onehot(s), 0 ...]               (K0 columns)
    z1 = x0 W0 + b0,   a1 = act(z1)
    z2 = a1 W1 + b1,   a2 = act(z2)
    y  = a2 W2[:, 0] + b2[0]

and a pose the sum of E over its atoms, each at R x + t (the pose transform
of fasten_main, src/serial/fasten.hpp:28-73).  The exact quantity it
stands in for is the pair sum of one ligand atom (src/serial/fasten.hpp:
75-131 with a single ligand atom), whose sum over a ligand's atoms is the
pose energy.  act is ReLU.  Loss = mean((y - target)^2).

With ``bf16=True`` values are rounded to BF16 where the GPU stores BF16: the
interpolated features, the weights, the hidden activations and the
back-propagated dZ; the grid itself is read in FP32 in training.  Inference
(``predict_atoms``) reads the grid as FP16 and interpolates with FP16
weights (per z-pair of corners, wx wy (1 - fz) and wx wy fz, their products
formed in FP32), as the fused scorer does: each z-pair's interpolated
features are rounded to BF16 on their own and the first layer sums the four
pairs' contributions (the scorer's MFMA sums them over its K dimension),
onto the slot's FP32 bias b0 + W0[F + slot].  With ``bf16=False`` it is the
plain FP64 network, whose gradients tests pin against central finite
differences of its own loss.
"""
from __future__ import annotations

import numpy as np


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even) and return as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def act(z):
    return np.maximum(z, 0.0)


def act_grad(z):
    return (z > 0.0).astype(np.float64)


def cell(p, lo, h, n):
    """Lower corner index (clamped to [0, n - 2]) and fraction (clamped to
    [0, 1]) of positions p (B, 3) on the grid, per axis."""
    u = (np.asarray(p, np.float64) - np.asarray(lo, np.float64)) / h
    i = np.clip(np.floor(u).astype(np.int64), 0, np.asarray(n) - 2)
    f = np.clip(u - i, 0.0, 1.0)
    return i, f


def corners(i, f, n):
    """Flat node index (B, 8) and trilinear weight (B, 8) of each corner,
    corner c = (dx, dy, dz) bits (x high, z low)."""
    n = np.asarray(n)
    idx = np.zeros((len(i), 8), np.int64)
    w = np.ones((len(i), 8))
    for c in range(8):
        d = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        ii = i + d
        idx[:, c] = (ii[:, 0] * n[1] + ii[:, 1]) * n[2] + ii[:, 2]
        for a in range(3):
            w[:, c] *= f[:, a] if d[a] else 1.0 - f[:, a]
    return idx, w


def to_f16(x: np.ndarray) -> np.ndarray:
    """Round to FP16 (nearest even) and return as float64."""
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float64)


def features(points, grid, lo, h, n):
    """phi (B, F): trilinear interpolation of grid (N, F) at points."""
    i, f = cell(points, lo, h, n)
    idx, w = corners(i, f, n)
    return np.einsum("bc,bcf->bf", w, np.asarray(grid, np.float64)[idx])


def features_infer_pairs(points, grid, lo, h, n):
    """(B, 4, F): the features each z-pair of corners q = 2 dx + dy (corners
    2q at z and 2q + 1 at z + 1) contributes, as the fused scorer forms
    them: FP16 grid values and FP16 weights."""
    i, f = cell(points, lo, h, n)
    idx, _ = corners(i, f, n)
    f32 = np.asarray(f, np.float32)
    one = np.float32(1.0)
    wx = [one - f32[:, 0], f32[:, 0]]
    wy = [one - f32[:, 1], f32[:, 1]]
    wz = [one - f32[:, 2], f32[:, 2]]
    w = np.zeros((len(points), 8))
    for c in range(8):
        dx, dy, dz = (c >> 2) & 1, (c >> 1) & 1, c & 1
        w[:, c] = to_f16((wx[dx] * wy[dy]).astype(np.float32) * wz[dz])
    g = to_f16(grid)[idx]  # (B, 8, F)
    return (w[:, 0::2, None] * g[:, 0::2] + w[:, 1::2, None] * g[:, 1::2])


def features_infer(points, grid, lo, h, n):
    """phi as the fused scorer forms it (the sum of the four z-pairs)."""
    return features_infer_pairs(points, grid, lo, h, n).sum(1)


def forward_backward(points, slots, target, grid, lo, h, n, W, b, k0: int, bf16: bool = True):
    """points (B, 3), slots (B,), target (B,), grid (N, F), W[l] (in, out),
    b[l] (out,).  Returns y (B,), loss, grad_grid (N, F), grads_W, grads_b."""
    r = to_bf16 if bf16 else (lambda x: np.asarray(x, np.float64))
    grid = np.asarray(grid, np.float64)
    B, F = len(points), grid.shape[1]
    i, f = cell(points, lo, h, n)
    idx, w = corners(i, f, n)
    phi = np.einsum("bc,bcf->bf", w, grid[idx])
    x0 = np.zeros((B, k0))
    x0[:, :F] = r(phi)
    x0[np.arange(B), F + np.asarray(slots)] = 1.0
    Wr = [r(x) for x in W]
    z1 = x0 @ Wr[0] + b[0]
    a1 = r(act(z1))
    z2 = a1 @ Wr[1] + b[1]
    a2 = r(act(z2))
    y = (a2 @ Wr[2] + b[2])[:, 0]
    d = y - target
    loss = float(np.mean(d * d))
    dz3 = r((2.0 * d / B)[:, None])
    gW, gb = [None] * 3, [None] * 3
    gW[2] = a2.T @ dz3
    gb[2] = dz3.sum(0)
    dz2 = r((dz3 @ Wr[2].T) * act_grad(z2))
    gW[1] = a1.T @ dz2
    gb[1] = dz2.sum(0)
    dz1 = r((dz2 @ Wr[1].T) * act_grad(z1))
    gW[0] = x0.T @ dz1
    gb[0] = dz1.sum(0)
    dx0 = dz1 @ Wr[0].T
    gG = np.zeros_like(grid)
    np.add.at(gG, idx.reshape(-1), (w[:, :, None] * dx0[:, None, :F]).reshape(-1, F))
    return y, loss, gG, gW, gb


def predict_atoms(points, slots, grid, lo, h, n, W, b, k0: int):
    """Inference: y (B,) per atom with FP16 interpolation, BF16 features
    (per z-pair), weights and activations (the one-hot row folded into a
    per-slot first-layer bias)."""
    F = grid.shape[1]
    phi4 = to_bf16(features_infer_pairs(points, grid, lo, h, n))
    Wr = [to_bf16(x) for x in W]
    z1 = np.einsum("bqf,fh->bh", phi4, Wr[0][:F]) + b[0] + Wr[0][F + np.asarray(slots)]
    a1 = to_bf16(act(z1))
    a2 = to_bf16(act(a1 @ Wr[1] + b[1]))
    return (a2 @ Wr[2] + b[2])[:, 0]


def rotation(poses):
    """R = Rz(pan) Ry(roll) Rx(tilt) of each pose (src/serial/fasten.hpp:39-50), (P, 3, 3)."""
    sx, cx = np.sin(poses[0]), np.cos(poses[0])
    sy, cy = np.sin(poses[1]), np.cos(poses[1])
    sz, cz = np.sin(poses[2]), np.cos(poses[2])
    return np.stack([
        np.stack([cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz], -1),
        np.stack([cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz], -1),
        np.stack([-sy, sx * cy, cx * cy], -1)], -2)


def score_poses(ligand_xyz, ligand_slots, poses, grid, lo, h, n, W, b, k0: int, sigma: float, mu: float):
    """Surrogate pose energies: sum over the ligand's atoms of sigma * y + mu."""
    poses = np.asarray(poses, np.float64)
    R = rotation(poses)
    pos = np.einsum("pij,lj->pli", R, np.asarray(ligand_xyz, np.float64)) + poses[3:6].T[:, None, :]
    P, L = pos.shape[:2]
    y = predict_atoms(pos.reshape(-1, 3), np.tile(ligand_slots, P), grid, lo, h, n, W, b, k0)
    return (sigma * y + mu).reshape(P, L).sum(1)
