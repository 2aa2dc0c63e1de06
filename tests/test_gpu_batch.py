"""The batched ragged persistent kernel against the plain kernel (bitwise:
same per-pose operation sequence) and the oracle, on ragged shapes the
reference driver never runs (pose counts that divide nothing, 1-pose
ligands, ligands larger than a wavefront)."""
import numpy as np
import pytest

import oracle_lib as O
from minibude_amd import metrics
from minibude_amd.batch import Batch, ragged_workload
from minibude_amd.capi import BatchConfig
from minibude_amd.model import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_set():
    ls = ragged_workload(nligands=24, natpro=1200, seed=5, natlig_range=(1, 96), nposes_range=(1, 3000))
    ls.poses[0] = ls.poses[0][:, :1]          # a 1-pose ligand
    ls.ligands[1] = ls.ligands[1][:1]         # a 1-atom ligand
    return ls


CONFIGS = [BatchConfig(1, 64, 4, 0, 0), BatchConfig(2, 256, 8, 1, 1), BatchConfig(4, 128, 2, 0, 1),
           BatchConfig(8, 512, 1, 1, 0)]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "p{ppwi}w{wgsize}d{waves_per_simd}o{order}g{use_graph}".format(
    **c.as_dict()))
def test_batch_equals_plain_bitwise(gpu_device, small_set, cfg):
    with Batch(small_set, gpu_device) as b:
        b.clear()
        b.run(cfg, warmup=1, iterations=2)
        for i in range(len(small_set.ligands)):
            got = b.energies(i)
            assert np.all(np.isfinite(got)), f"ligand {i}: pose left unwritten"
            with Context(small_set.deck(i), gpu_device) as c:
                want = c.fasten(cfg.ppwi, 64).energies
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"ligand {i}"


def test_batch_vs_oracle(gpu_device, small_set):
    with Batch(small_set, gpu_device) as b:
        b.run(BatchConfig(2, 256, 4, 0, 1))
        for i in range(len(small_set.ligands)):
            d = small_set.deck(i)
            n = min(d.nposes, 64)
            e64, mag = O.fasten_f64_mag(d, 0, n)
            got = b.energies(i)[:n]
            # |error| relative to the pose's magnitude (sum of |terms|)
            assert np.max(np.abs(got - e64) / np.maximum(mag, 1.0)) < 2e-6, f"ligand {i}"
            ok = np.abs(e64) >= 1e-3 * mag
            if ok.any():
                assert metrics.symmetric_max_rel_pct(e64[ok], got[ok]) < metrics.DIFF_TOLERANCE_PCT


def test_separate_baseline_and_autotune(gpu_device):
    ls = ragged_workload(nligands=16, natpro=2048, seed=6, nposes_range=(512, 4096))
    with Batch(ls, gpu_device) as b:
        b.clear()
        b.run_separate(1, 256)
        sep = [b.energies(i) for i in range(16)]
        cfg, ms = b.autotune()
        assert ms > 0 and cfg.use_graph == 1
        b.clear()
        b.run(cfg)
        for i in range(16):
            assert np.array_equal(b.energies(i).view(np.uint32), sep[i].view(np.uint32))
