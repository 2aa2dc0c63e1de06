"""Pins of the learned per-atom field's CPU oracle (oracle/atomnet_oracle.py):
trilinear interpolation reproduces linear fields exactly and clamps outside
the grid, the rotation is Rz Ry Rx, and every gradient of its loss matches
central finite differences (FP64)."""
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import atomnet_oracle as A  # noqa: E402

F, H = 8, 32


def small_net(nslots=3, n=(4, 5, 3), seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    k0 = 32 * ((F + nslots + 31) // 32)
    grid = rng.normal(0, 0.5, (int(np.prod(n)), F))
    W = [np.zeros((k0, H)), rng.normal(0, scale * 0.3, (H, H)), rng.normal(0, scale * 0.3, (H, 1))]
    W[0][:F + nslots] = rng.normal(0, scale * 0.4, (F + nslots, H))
    b = [rng.normal(0, 0.1, H), rng.normal(0, 0.1, H), rng.normal(0, 0.1, 1)]
    return grid, W, b, k0


def test_trilinear_reproduces_linear_fields_and_clamps():
    n = (5, 4, 6)
    lo, h = np.array([-1.0, 2.0, 0.5]), 0.75
    idx = np.indices(n).reshape(3, -1).T  # (i, j, k) row-major, as the grid's node order
    pos = lo + h * idx
    coef = np.random.default_rng(1).normal(size=(F, 4))
    grid = pos @ coef[:, :3].T + coef[:, 3]
    pts = lo + h * np.random.default_rng(2).uniform(0, np.array(n) - 1, (200, 3))
    phi = A.features(pts, grid, lo, h, n)
    assert np.allclose(phi, pts @ coef[:, :3].T + coef[:, 3], atol=1e-12)
    # a node exactly: its own value; outside the grid: the boundary's value
    assert np.allclose(A.features(pos[[7]], grid, lo, h, n), grid[[7]])
    far = np.array([[lo[0] - 10, lo[1] + h, lo[2] + 99.0]])
    clamped = np.array([[lo[0], lo[1] + h, lo[2] + h * (n[2] - 1)]])
    assert np.allclose(A.features(far, grid, lo, h, n), A.features(clamped, grid, lo, h, n))


def test_rotation_is_rz_ry_rx():
    p = np.random.default_rng(4).uniform(-np.pi, np.pi, (6, 5))
    R = A.rotation(p)
    for j in range(5):
        tx, ty, tz = p[0, j], p[1, j], p[2, j]
        rx = np.array([[1, 0, 0], [0, np.cos(tx), -np.sin(tx)], [0, np.sin(tx), np.cos(tx)]])
        ry = np.array([[np.cos(ty), 0, np.sin(ty)], [0, 1, 0], [-np.sin(ty), 0, np.cos(ty)]])
        rz = np.array([[np.cos(tz), -np.sin(tz), 0], [np.sin(tz), np.cos(tz), 0], [0, 0, 1]])
        assert np.allclose(R[j], rz @ ry @ rx, atol=1e-12)


def test_gradients_match_finite_differences():
    n = (4, 5, 3)
    lo, h = np.array([0.0, 0.0, 0.0]), 1.0
    grid, W, b, k0 = small_net(n=n)
    rng = np.random.default_rng(5)
    B = 64
    pts = rng.uniform(0.1, np.array(n) - 1.1, (B, 3))
    slots = rng.integers(0, 3, B)
    target = rng.normal(size=B)

    def loss_of(g, Ws, bs):
        return A.forward_backward(pts, slots, target, g, lo, h, n, Ws, bs, k0, bf16=False)[1]

    _, _, gG, gW, gb = A.forward_backward(pts, slots, target, grid, lo, h, n, W, b, k0, bf16=False)
    eps = 1e-6
    checks = 0
    for (arr, grad) in [(grid, gG), (W[0], gW[0]), (W[1], gW[1]), (W[2], gW[2]), (b[0], gb[0]), (b[1], gb[1]),
                        (b[2], gb[2])]:
        flat = arr.reshape(-1)
        for i in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            if arr is W[0] and i // H >= F + 3:
                continue  # padding rows: zero input, zero gradient (checked below)
            old = flat[i]
            flat[i] = old + eps
            up = loss_of(grid, W, b)
            flat[i] = old - eps
            dn = loss_of(grid, W, b)
            flat[i] = old
            fd = (up - dn) / (2 * eps)
            assert grad.reshape(-1)[i] == pytest.approx(fd, rel=1e-4, abs=1e-8), (arr.shape, i)
            checks += 1
    assert checks > 60
    assert np.all(gW[0][F + 3:] == 0.0)


def test_inference_folds_the_one_hot_row_into_the_bias():
    n = (4, 4, 4)
    lo, h = np.zeros(3), 1.0
    grid, W, b, k0 = small_net(n=n, seed=7)
    rng = np.random.default_rng(8)
    pts = rng.uniform(0, 3, (64, 3))
    slots = rng.integers(0, 3, 64)
    y_train = A.forward_backward(pts, slots, np.zeros(64), grid, lo, h, n, W, b, k0, bf16=True)[0]
    # at grid nodes the FP16 interpolation weights are exact (0 or 1): with a
    # grid already in FP16 inference and the training forward agree
    nodes = np.floor(pts)
    g16 = A.to_f16(grid)
    y_infer = A.predict_atoms(nodes, slots, g16, lo, h, n, W, b, k0)
    assert np.allclose(y_infer, A.forward_backward(nodes, slots, np.zeros(64), g16, lo, h, n, W, b, k0)[0],
                       atol=1e-6)
    # off the nodes they differ by the FP16 rounding of grid and weights only
    y_infer = A.predict_atoms(pts, slots, grid, lo, h, n, W, b, k0)
    assert np.max(np.abs(y_infer - y_train)) < 0.05 * np.max(np.abs(y_train))


def test_inference_pairs_group_the_two_z_corners_of_each_xy_corner():
    n = (4, 4, 4)
    lo, h = np.zeros(3), 1.0
    pts = np.random.default_rng(11).uniform(0.05, 2.95, (300, 3))
    i, f = A.cell(pts, lo, h, n)
    for d in range(8):  # one corner (dx, dy, dz) of every point's cell carries the only nonzero node value
        dx, dy, dz = (d >> 2) & 1, (d >> 1) & 1, d & 1
        node = ((i[:, 0] + dx) * n[1] + i[:, 1] + dy) * n[2] + i[:, 2] + dz
        for b in range(len(pts)):
            grid = np.zeros((int(np.prod(n)), F))
            grid[node[b]] = 1.0 + np.arange(F)
            pairs = A.features_infer_pairs(pts[[b]], grid, lo, h, n)[0]
            w = (f[b, 0] if dx else 1 - f[b, 0]) * (f[b, 1] if dy else 1 - f[b, 1]) * (f[b, 2] if dz else 1 - f[b, 2])
            q = 2 * dx + dy
            assert np.all(np.delete(pairs, q, 0) == 0.0)
            # FP16 weight: 2^-11 relative
            assert np.allclose(pairs[q], w * (1.0 + np.arange(F)), rtol=1e-3, atol=1e-6)
    # and the pairs sum to the trilinear interpolation (FP16 rounding apart)
    grid = np.random.default_rng(12).normal(0, 1, (int(np.prod(n)), F))
    assert np.allclose(A.features_infer_pairs(pts, grid, lo, h, n).sum(1), A.features(pts, grid, lo, h, n),
                       atol=4e-3)


def test_pose_score_is_the_sum_over_atoms():
    n = (6, 6, 6)
    lo, h = np.full(3, -5.0), 2.0
    grid, W, b, k0 = small_net(n=n, seed=9)
    rng = np.random.default_rng(10)
    xyz = rng.uniform(-2, 2, (5, 3))
    slots = np.array([0, 1, 2, 1, 0])
    poses = np.vstack([rng.uniform(-np.pi, np.pi, (3, 16)), rng.uniform(-1, 1, (3, 16))])
    s = A.score_poses(xyz, slots, poses, grid, lo, h, n, W, b, k0, sigma=3.0, mu=-1.0)
    R = A.rotation(poses)
    for j in range(16):
        pos = xyz @ R[j].T + poses[3:6, j]
        want = np.sum(3.0 * A.predict_atoms(pos, slots, grid, lo, h, n, W, b, k0) - 1.0)
        assert s[j] == pytest.approx(want, rel=1e-12, abs=1e-9)
