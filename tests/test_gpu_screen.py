"""The learned prescreen on the GPU: training kernels against the CPU oracle
(gradients), top-k selection, exact rescoring, fallback, and what the trained
network gains over the same screen without training."""
import sys

import numpy as np
import pytest

from conftest import ROOT
from minibude_amd import capi
from minibude_amd import deck as D
from minibude_amd import screen as S
from minibude_amd.model import Context

sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("n", [256, 2560, 8192])
def test_forward_backward_matches_oracle(gpu_device, n):
    # n = 256: every weight-gradient slice is one k step (the loop's tail
    # alone); 2,560: one group of four k steps and a tail step (an odd group
    # count); 8,192: four groups, the pipelined loop of training
    rng = np.random.default_rng(3)
    W, b = M.init(seed=11)
    poses = D.synth_poses(12, n)
    feat = S.features(poses)
    target = rng.normal(size=n)
    pred, loss, gW, gb = S.mlp_forward_backward(feat, target, W, b, gpu_device)
    rp, rl, rgW, rgb = M.forward_backward(feat, target, W, b, bf16=True)
    # BF16 storage points are emulated by the oracle; the rest is FP32 (GPU)
    # vs FP64 (oracle) accumulation, plus rare one-ulp BF16 flips
    assert rel_err(pred, rp) < 2e-3
    assert loss == pytest.approx(rl, rel=5e-3)
    for l in range(4):
        assert rel_err(gW[l], rgW[l]) < 2e-2, f"dW{l}"
        assert rel_err(gb[l], rgb[l]) < 2e-2, f"db{l}"
    # and against the unrounded FP64 network the BF16 error is bounded too
    fp = M.forward_backward(feat, target, W, b, bf16=False)
    assert rel_err(pred, fp[0]) < 3e-2
    assert rel_err(gW[1], fp[2][1]) < 1.5e-1


@pytest.fixture(scope="module")
def small_workload():
    return S.screen_workload(ntrain=98304, nscreen=65536, natpro=1500, seed=9, protein_seed=9)


def test_fallback_selects_exact_lowest(gpu_device, small_workload):
    w = small_workload
    cfg = S.default_config(batch=1024, epochs=1, keep_fraction=0.02, max_val_rmse=0.0)
    with S.Screen(w.protein, w.ligand, w.forcefield, cfg, gpu_device) as s:
        st, _ = s.train(w.train_poses)
        assert st.fallback == 1
        s.load(w.screen_poses)
        exact, _ = s.exhaustive()
        # the configured fraction, then others set without retraining
        for keep in (0.02, 0.001, 0.05):
            if keep != 0.02:
                s.set_keep(keep)
            r = s.run()
            assert r.fallback == 1
            idx, e = s.selected()
            k = int(np.ceil(float(np.float32(keep)) * len(exact)))  # keep as the C float the library holds
            assert len(idx) >= k
            thr = np.sort(exact)[k - 1]
            assert set(idx.tolist()) == set(np.nonzero(exact <= thr)[0].tolist())
            assert np.array_equal(e, exact[idx])
        with pytest.raises(capi.BudeInvalidArgument):
            s.set_keep(0.0)
    assert S.recall(idx, exact) == 1.0


def test_trained_screen_beats_untrained(gpu_device, small_workload):
    w = small_workload
    results = {}
    for epochs, lr in ((0, 1e-30), (8, 2e-3)):
        cfg = S.default_config(batch=1024, epochs=max(1, epochs), lr=lr, max_val_rmse=10.0, seed=1)
        with S.Screen(w.protein, w.ligand, w.forcefield, cfg, gpu_device) as s:
            st, curve = s.train(w.train_poses)
            s.load(w.screen_poses)
            r = s.run()
            idx, e = s.selected()
            exact, _ = s.exhaustive()
            with Context(D.Deck(w.protein, w.ligand, w.forcefield, w.screen_poses[:, idx]), gpu_device) as c:
                want = c.fasten(1, 64).energies
        # the rescored energies are the plain kernel's, bit for bit
        assert np.array_equal(e.view(np.uint32), want.view(np.uint32))
        results[epochs] = (S.recall(idx, exact), st.val_rmse, curve)
        print(f"epochs={epochs}: recall@1% with 2% rescored = {results[epochs][0]:.3f}, val_rmse={st.val_rmse:.3f}",
              flush=True)
    trained, untrained = results[8][0], results[0][0]
    assert results[8][2][-1] < results[8][2][0]        # loss decreased
    # what training buys over the same screen untrained (random ranking keeps ~2% of the top 1%)
    assert trained > 10 * max(untrained, 0.02) and trained > 0.25


def test_fused_inference_matches_oracle_and_training_forward(gpu_device):
    W, b = M.init(seed=5)
    poses = D.synth_poses(13, 1000)  # not a multiple of the 64-pose tile
    fused = S.mlp_predict(poses, W, b, gpu_device)
    feat = S.features(poses)
    ref = M.forward_backward(feat, np.zeros(len(feat)), W, b, bf16=True)[0]
    assert rel_err(fused, ref) < 2e-3
    train_fwd = S.mlp_forward_backward(feat[:992], np.zeros(992), W, b, gpu_device)[0]
    assert rel_err(fused[:992], train_fwd) < 2e-3


@pytest.mark.parametrize("features", [12, 21])
def test_graph_training_equals_step_by_step(gpu_device, small_workload, features, monkeypatch):
    # the training loop replays one captured step as a hipGraph (the epoch's
    # order uploaded between replays); launched step by step the same kernels
    # give the same network, so the same predictions, bit for bit; with 12
    # inputs the rotation-matrix rows of the first layer stay zero
    w = small_workload
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("BUDE_MI355X_TRAIN_GRAPH", mode)
        cfg = S.default_config(batch=1024, epochs=3, features=features, seed=2)
        with S.Screen(w.protein, w.ligand, w.forcefield, cfg, gpu_device) as s:
            st, curve = s.train(w.train_poses[:, :32768])
            s.load(w.screen_poses[:, :4096])
            out[mode] = (s.predict(), curve)
    assert np.array_equal(out["1"][0].view(np.uint32), out["0"][0].view(np.uint32))
    assert np.array_equal(out["1"][1], out["0"][1])
    assert out["1"][1][-1] < out["1"][1][0]


def test_configured_12_inputs_at_2pct(gpu_device):
    # the learned prescreen as configured (12 pose inputs, 3 x 256 GELU) on a
    # scaled copy of its workload: recall of the exact top 1 % at the
    # configured 2 % rescored.  The pose network does not reach 0.99 there
    # (the learned per-atom field does, tests/test_gpu_atomnet.py)
    w = S.screen_workload(ntrain=1 << 20, nscreen=1 << 20, natpro=2048, seed=9, protein_seed=9)
    cfg = S.default_config(features=12, epochs=20)
    with S.Screen(w.protein, w.ligand, w.forcefield, cfg, gpu_device) as s:
        s.train(w.train_poses)
        s.load(w.screen_poses)
        exact, _ = s.exhaustive()
        s.run()
        idx, _ = s.selected()
    rec = S.recall(idx, exact, 0.01)
    print(f"12-input pose MLP: recall@1% with 2% rescored = {rec:.4f}", flush=True)
    assert rec > 0.5  # far above the 0.02 of a random ranking
    if rec < 0.99:
        pytest.xfail(f"the configured pose MLP keeps {rec:.3f} of the exact top 1 % at 2 % rescored, below 0.99")
