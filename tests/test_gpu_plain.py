"""Parity of the plain HIP fasten_main (through the C ABI) with the
reference's golden energies and with the oracle.

Tolerance: the reference's own max_diff_% < 0.025 (src/bude.h:11) with its
signed metric, and the same 0.025 % as a symmetric bound against the FP32
and FP64 oracles (|E| >= 1), stated per deck.
"""
import json

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN
from handdecks import HB_E, HB_F, NORMAL, pair_deck
from minibude_amd import deck as D
from minibude_amd import metrics
from minibude_amd.model import PPWIS, Context, Mi355xBude, Params

pytestmark = pytest.mark.gpu

TOL_PCT = metrics.DIFF_TOLERANCE_PCT  # 0.025 %
WGSIZES = (64, 128, 256, 512, 1024)


def check(ref, got, tol=TOL_PCT, label=""):
    m, failed = metrics.validate(ref, got)
    assert metrics.is_valid(m), f"{label}: max_diff_% {m} ({failed.size} failed)"
    s = metrics.symmetric_max_rel_pct(ref, got)
    assert s < tol, f"{label}: symmetric max rel {s}%"
    return m, s


@pytest.fixture(scope="module")
def bm1():
    return D.read_deck(GOLDEN / "bm1")


@pytest.fixture(scope="module")
def bm1_ctx(gpu_device, bm1):
    with Context(bm1, gpu_device) as c:
        yield c


def test_bm1_every_ppwi_and_wgsize(bm1, bm1_ctx):
    ref32 = O.fasten_f32(bm1)
    for p in PPWIS:
        for w in WGSIZES:
            if bm1.nposes % (p * w):
                continue
            e = bm1_ctx.fasten(p, w).energies
            print(f"bm1 ppwi={p} wgsize={w}", flush=True)
            check(bm1.ref_energies, e, label=f"bm1 file p{p} w{w}")
            check(ref32, e, label=f"bm1 oracle p{p} w{w}")


def test_bm1_fp64_error(bm1, bm1_ctx):
    e = bm1_ctx.fasten(8, 64).energies
    e64 = O.fasten_f64(bm1, 0, 8192)
    s = metrics.symmetric_max_rel_pct(e64, e[:8192])
    print(f"bm1 max rel vs FP64 oracle: {s:.5f}%")
    assert s < TOL_PCT


def test_bm2(gpu_device):
    d = D.read_deck(GOLDEN / "bm2")
    with Context(d, gpu_device) as c:
        e = c.fasten(8, 256).energies
    check(d.ref_energies, e, label="bm2 file")
    e64 = O.fasten_f64(d, 0, 64)
    assert metrics.symmetric_max_rel_pct(e64, e[:64]) < TOL_PCT


EPS32 = 2.0 ** -24
# FP32 noise bound relative to a pose's magnitude M = 0.5 * sum |term| (the
# cancellation scale): the reference's own FP32 energies stay within 8.8 eps*M
# of FP64 on every pose of both synthetic decks (tests/golden/*/fp64.f32).
NOISE_EPS_M = 16.0
# Poses with |E| >= TAU * M: the reference's own FP32 is within 0.014 % of
# FP64 there, so the 0.025 % metric is meaningful (92.5 % of c1, 98.5 % of c0).
TAU = 0.002


def load_fixture(name):
    fx = GOLDEN / name
    meta = json.loads((fx / "manifest.json").read_text())
    ref = np.fromfile(fx / "ref_energies.f32", "<f4")
    em = np.fromfile(fx / "fp64.f32", "<f4").reshape(-1, 2).astype(np.float64)
    return meta, ref, em[:, 0], em[:, 1]


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_synthetic_workload_vs_reference_fixture(gpu_device, name):
    meta, ref, e64, mag = load_fixture(name)
    d = D.workload_deck(name)
    assert ref.size == d.nposes and meta["generator"]["nposes"] == d.nposes
    cond = np.abs(e64) >= TAU * mag
    ref_noise = np.max(np.abs(ref - e64) / mag) / EPS32
    with Context(d, gpu_device) as c:
        for p, w in ((1, 64), (8, 256), (16, 128), (4, 1024)):
            e = c.fasten(p, w).energies
            noise = np.max(np.abs(e - e64) / mag) / EPS32
            m_all, f_all = metrics.validate(ref, e)
            m_c, f_c = metrics.validate(ref[cond], e[cond])
            print(f"{name} ppwi={p} wgsize={w}: vs reference max_diff_%={m_all:.5f} ({f_all.size} failed of "
                  f"{e.size}); on |E|>={TAU}M ({cond.mean():.4f} of poses) max_diff_%={m_c:.5f}; "
                  f"|E-E64|/M = {noise:.2f} eps (reference FP32: {ref_noise:.2f} eps)", flush=True)
            # as close to FP64 as the reference's own FP32 is, relative to the cancellation scale
            assert noise < NOISE_EPS_M
            # and within the reference's tolerance of the reference wherever FP32 can resolve it
            check(ref[cond], e[cond], label=f"{name} p{p} w{w} conditioned")
            check(e64[cond], e[cond], label=f"{name} p{p} w{w} conditioned vs FP64")
            if name == "c0":  # well conditioned everywhere: the plain reference metric on every pose
                check(ref, e, label=f"{name} p{p} w{w}")


@pytest.mark.parametrize("nposes", [1, 3, 1000, 4097])
def test_ragged_pose_counts(gpu_device, nposes):
    # pose counts that do not divide ppwi * wgsize (the reference driver skips them)
    d = D.synth_deck(700, 19, nposes, seed=21)
    ref = O.fasten_f32(d)
    with Context(d, gpu_device) as c:
        for p, w in ((1, 64), (8, 64), (128, 256), (4, 1024)):
            check(ref, c.fasten(p, w).energies, label=f"n={nposes} p{p} w{w}")


def test_closed_form_pairs(gpu_device):
    P, L = (NORMAL, 1.5, -1.0, 0.5), (NORMAL, 1.0, -0.5, 0.4)
    dists = [1.0, 3.0, 5.0, 9.0]
    # hand-evaluated in tests/test_oracle_closed_forms.py
    want = np.array([0.5 * (45.6 + 9.0 - 1.5), 0.5 * (6.75 - 15 / 11), 0.5 * (-9 / 11), 0.0])
    for axis in range(3):
        with Context(pair_deck(P, L, dists, axis), gpu_device) as c:
            np.testing.assert_allclose(c.fasten(1, 64).energies, want, rtol=2e-6, atol=1e-6)
    # E-type flip, F-F long range, desolvation branches: compare with the oracle
    for p, l in (((HB_E, 1.5, -1.0, 0.5), L), ((HB_F, 1.5, -1.0, 0.5), (HB_F, 1.0, -0.5, 0.4)),
                 ((NORMAL, 1.5, 0.8, 0.0), (NORMAL, 1.0, -0.5, 0.0)), ((NORMAL, 1.5, 0.0, 0.3), L),
                 ((NORMAL, 1.5, 1.0, 0.3), (NORMAL, 1.0, 0.5, 0.2))):
        d = pair_deck(p, l, np.linspace(0.2, 10.0, 99))
        with Context(d, gpu_device) as c:
            np.testing.assert_allclose(c.fasten(2, 64).energies, O.fasten_f64(d), rtol=1e-5, atol=2e-5)


def test_far_field_zero(gpu_device):
    d = D.synth_deck(500, 30, 256, seed=4)
    d.poses[3] += 1000.0
    with Context(d, gpu_device) as c:
        assert np.all(c.fasten(4, 64).energies == 0.0)


def test_model_interface_like_reference_driver(gpu_device, bm1):
    # Bude<PPWI>::fasten: upload, 2 warmups + iterations, energies back
    b = Mi355xBude(16)
    assert b.enumerate_devices()[0][1].startswith("AMD Instinct MI355X") or b.enumerate_devices()
    assert b.compatible(Params(bm1), 256, gpu_device)
    s = b.fasten(Params(bm1, iterations=3), 256, gpu_device)
    assert len(s.kernel_ms) == 5 and len(s.device_ms) == 5 and s.context_ms > 0
    assert all(h >= dv * 0.5 for h, dv in zip(s.kernel_ms, s.device_ms))
    r = metrics.evaluate(bm1.ref_energies, s.energies, s.kernel_ms, 2, 3, bm1.nposes, bm1.natlig, bm1.natpro, 16)
    assert r.valid and r.interactions_per_sec > 0


def test_every_group_size(gpu_device):
    # The pair loop walks each protein-type group in pipelined blocks of four
    # atoms with a remainder of up to seven; give the groups every size from
    # 0 to 23 (and some larger ones) so each loop exit is taken, and check
    # every register-block width against the FP64 oracle.
    ntypes = 34
    sizes = [t if t < 24 else 24 + 3 * (t - 24) for t in range(ntypes)]
    pro = D.synth_protein(31, sum(sizes), ntypes)
    pro["type"] = np.repeat(np.arange(ntypes, dtype=np.int32), sizes)
    pro = pro[np.random.default_rng(5).permutation(len(pro))]  # groups are formed by the library
    d = D.Deck(pro, D.synth_ligand(32, 12, ntypes), D.synth_forcefield(31, ntypes), D.synth_poses(33, 2000))
    e64, mag = O.fasten_f64_mag(d, 0, d.nposes)
    eps = 2.0 ** -24
    with Context(d, gpu_device) as c:
        first = None
        for p, w in ((1, 64), (2, 128), (4, 256), (8, 64)):
            e = c.fasten(p, w).energies
            ok = mag > 0
            assert np.all(e[~ok] == 0.0)
            noise = float(np.max(np.abs(e.astype(np.float64) - e64)[ok] / mag[ok]) / eps)
            print(f"group sizes 0-23 and larger, ppwi {p} wgsize {w}: max |E-E64|/M = {noise:.2f} eps", flush=True)
            assert noise < 16.0
            if first is None:
                first = e
            else:  # the register block only changes which lane holds a pose, never the sums
                assert np.array_equal(e.view(np.uint32), first.view(np.uint32))
