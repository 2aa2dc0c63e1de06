"""FP64-oracle error of the exact energies on every workload shape: samples of
the batched (256 ragged ligands), learned-prescreen (26-atom ligand vs the
4,096-atom protein) and combined (16 proteins of 900-6,000 atoms x ligands of
8-96 atoms) workloads, scored by the library and by the oracle in FP64.
Each test prints |E - E64| / M (M = 0.5 * sum |term|, in FP32 epsilons) and
the reference metric on the poses FP32 resolves (|E64| >= 0.002 M)."""
import numpy as np
import pytest

import oracle_lib as O
from minibude_amd import combined as C
from minibude_amd import deck as D
from minibude_amd import metrics
from minibude_amd import screen as S
from minibude_amd.batch import Batch, ragged_workload
from minibude_amd.capi import BatchConfig
from minibude_amd.model import Context

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -24


def errors(got, e64, mag):
    got = np.asarray(got, np.float64)
    noise = float(np.max(np.abs(got - e64) / mag) / EPS)
    ok = np.abs(e64) >= 0.002 * mag
    md, failed = metrics.validate(e64.astype(np.float32)[ok], got.astype(np.float32)[ok])
    rel = metrics.symmetric_max_rel_pct(e64[ok], got[ok])
    return noise, md, int(failed.size), rel, int(ok.sum())


def report(name, got, e64, mag):
    noise, md, failed, rel, nres = errors(got, e64, mag)
    print(f"{name}: {len(got)} poses, max |E-E64|/M = {noise:.2f} eps, reference metric on the {nres} resolvable "
          f"poses: max_diff {md:.5f} % ({failed} failed), max rel. vs FP64 {rel:.5f} %", flush=True)
    assert noise < 16.0
    assert failed == 0


def test_batched_workload_vs_fp64(gpu_device):
    ls = ragged_workload()  # the 256-ligand workload
    rng = np.random.default_rng(0)
    got, e64, mag = [], [], []
    with Batch(ls, gpu_device) as b:
        b.run(BatchConfig(1, 64, 8, 0, 1))
        for i in range(0, len(ls.ligands), 8):  # every 8th ligand, 64 random poses each
            d = ls.deck(i)
            idx = np.sort(rng.choice(d.nposes, 64, replace=False))
            sub = D.Deck(d.protein, d.ligand, d.forcefield, np.ascontiguousarray(d.poses[:, idx]))
            e, m = O.fasten_f64_mag(sub, 0, 64)
            got.append(b.energies(i)[idx])
            e64.append(e)
            mag.append(m)
    report("batched", np.concatenate(got), np.concatenate(e64), np.concatenate(mag))


def test_learned_workload_vs_fp64(gpu_device):
    w = S.screen_workload(ntrain=64, nscreen=4096)  # same protein, ligand and pose streams
    d = D.Deck(w.protein, w.ligand, w.forcefield, w.screen_poses)
    with Context(d, gpu_device) as c:
        got = c.fasten(1, 256).energies
    e64, mag = O.fasten_f64_mag(d, 0, d.nposes)
    report("learned", got, e64, mag)


def test_combined_workload_vs_fp64(gpu_device):
    w = C.combined_workload(nproteins=16, nligands=16, nposes=4096)
    rng = np.random.default_rng(1)
    got, e64, mag = [], [], []
    cfg = C.default_config("field")
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        for p in range(0, 16, 3):
            c.screen(w.proteins[p])
            c.exhaustive()
            for l in rng.choice(16, 4, replace=False):
                idx = np.sort(rng.choice(4096, 48, replace=False))
                sub = D.Deck(w.proteins[p], w.ligands[l], w.forcefield, np.ascontiguousarray(w.poses[l][:, idx]))
                e, m = O.fasten_f64_mag(sub, 0, 48)
                got.append(c.exact(int(l))[idx])
                e64.append(e)
                mag.append(m)
    report("combined", np.concatenate(got), np.concatenate(e64), np.concatenate(mag))
