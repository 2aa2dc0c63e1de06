"""Pin the MLP oracle (oracle/mlp_oracle.py): its analytic gradients against
central finite differences of its own loss, and its BF16 rounding."""
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding():
    assert M.to_bf16(np.float32([1.0]))[0] == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even -> 1.0
    assert M.to_bf16(np.float32([1 + 2 ** -8]))[0] == 1.0
    assert M.to_bf16(np.float32([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6
    assert M.to_bf16(np.float32([-3.0]))[0] == -3.0
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(M.to_bf16(x) - x) / np.abs(x)) <= 2 ** -8  # half an ulp of an 8-bit significand


def test_gelu_grad_is_the_derivative():
    z = np.linspace(-4, 4, 101)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
def test_gradients_match_finite_differences(layer):
    rng = np.random.default_rng(layer)
    W, b = M.init(seed=7)
    feat = rng.normal(size=(32, 21))
    target = rng.normal(size=32)
    _, _, gW, gb = M.forward_backward(feat, target, W, b, bf16=False)
    h = 1e-6
    for _ in range(6):
        i = rng.integers(W[layer].shape[0])
        j = rng.integers(W[layer].shape[1])
        Wp = [w.copy() for w in W]
        Wm = [w.copy() for w in W]
        Wp[layer][i, j] += h
        Wm[layer][i, j] -= h
        num = (M.forward_backward(feat, target, Wp, b, bf16=False)[1] -
               M.forward_backward(feat, target, Wm, b, bf16=False)[1]) / (2 * h)
        assert gW[layer][i, j] == pytest.approx(num, rel=1e-5, abs=1e-9)
    j = rng.integers(len(b[layer]))
    bp = [x.copy() for x in b]
    bm = [x.copy() for x in b]
    bp[layer][j] += h
    bm[layer][j] -= h
    num = (M.forward_backward(feat, target, W, bp, bf16=False)[1] -
           M.forward_backward(feat, target, W, bm, bf16=False)[1]) / (2 * h)
    assert gb[layer][j] == pytest.approx(num, rel=1e-5, abs=1e-9)
