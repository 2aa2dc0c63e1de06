"""CPU restatement of the pose-energy surrogate's forward and backward pass
(numpy, FP64) -- test infrastructure only, the checker of the GPU training
kernels; the product never imports it.

Network: 21 features -> 256 -> 256 -> 256 -> 1, GELU (tanh form,
0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))) on the three hidden layers, loss = mean((pred - target)^2).  With ``bf16=True`` values are
rounded to BF16 (round to nearest even) exactly where the GPU stores BF16:
inputs, weights, hidden activations, GELU'(Z) of the hidden layers (kept for
the backward pass) and the back-propagated dZ; everything else is FP64.  With ``bf16=False`` it is the plain FP64 network, whose
gradients tests/test_mlp_oracle.py pins against central finite differences.
"""
from __future__ import annotations

import math

import numpy as np

def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even) and return as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


_C = math.sqrt(2.0 / math.pi)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(_C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * z * z)


def forward_backward(feat, target, W, b, bf16: bool = True):
    """feat (B, 21), target (B,), W[l] (in, out), b[l] (out,).
    Returns pred (B,), loss, grads_W, grads_b."""
    r = to_bf16 if bf16 else (lambda x: np.asarray(x, np.float64))
    x = r(feat)
    Wr = [r(w) for w in W]
    B = x.shape[0]
    zs, acts = [], [x]
    h = x
    for l in range(3):
        z = h @ Wr[l] + b[l]
        zs.append(z)
        h = r(gelu(z))
        acts.append(h)
    pred = (h @ Wr[3] + b[3])[:, 0]
    d = pred - target
    loss = float(np.mean(d * d))
    dz = r((2.0 * d / B)[:, None])
    gW, gb = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        gW[l] = acts[l].T @ dz
        gb[l] = dz.sum(axis=0)
        if l == 0:
            break
        dz = r((dz @ Wr[l].T) * r(gelu_grad(zs[l - 1])))
    return pred, loss, gW, gb


def init(seed: int = 0, scale: float = 1.0):
    """He-normal weights (W[3] 1/sqrt(fan_in)), small random biases."""
    rng = np.random.default_rng(seed)
    shapes = [(21, 256), (256, 256), (256, 256), (256, 1)]
    W = [rng.normal(0, scale * math.sqrt((1.0 if l == 3 else 2.0) / s[0]), s) for l, s in enumerate(shapes)]
    b = [rng.normal(0, 0.05, s[1]) for s in shapes]
    return W, b
