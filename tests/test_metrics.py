"""The reference driver's validation quirks and FLOP accounting."""
import numpy as np
import pytest

import oracle_lib as O
from minibude_amd import metrics


def test_validate_matches_c_restatement():
    rng = np.random.default_rng(1)
    ref = rng.normal(0, 50, 5000).astype(np.float32)
    ref[:100] = rng.uniform(-0.9, 0.9, 100)
    sample = (ref * (1 + rng.normal(0, 2e-4, ref.size))).astype(np.float32)
    py = metrics.validate(ref, sample)
    c = O.validate_c(ref, sample)
    assert py[0] == pytest.approx(c[0], rel=1e-12)
    assert py[1].size == c[1]


def test_negative_references_never_fail():
    # src/main.cpp:329 divides by the signed reference
    ref = np.array([-10.0, -100.0], np.float32)
    sample = np.array([-20.0, 100.0], np.float32)
    m, failed = metrics.validate(ref, sample)
    assert m == 0.0 and failed.size == 0
    assert metrics.symmetric_max_rel_pct(ref, sample) == pytest.approx(200.0)


def test_small_entries_skipped_only_when_both_small():
    assert metrics.validate(np.float32([0.5]), np.float32([0.9]))[0] == 0.0
    m, failed = metrics.validate(np.float32([0.5]), np.float32([1.5]))
    assert m == pytest.approx(200.0) and failed.tolist() == [0]


def test_threshold_edge():
    ref = np.float32([1000.0, 1000.0])
    ok = np.float32([1000.2, 1000.0])      # 0.02 %
    bad = np.float32([1000.3, 1000.0])     # 0.03 %
    assert metrics.is_valid(metrics.validate(ref, ok)[0])
    assert not metrics.is_valid(metrics.validate(ref, bad)[0])


def test_flop_accounting_matches_reference_formula():
    # src/main.cpp:358-360 at the single-GPU workload (4096/64/1,048,576)
    n, L, N = 1048576, 64, 4096
    for p, want in ((1, 1.0997e13), (8, 8.591e12), (16, 8.419e12), (64, 8.290e12)):
        assert metrics.flops_per_pass(n, L, N, p) == pytest.approx(want, rel=2e-4)
    assert metrics.interactions_per_pass(n, L, N) == 274877906944
    # bm1 at ppwi=1: (27 + 26*(2+18+938*40) + 1) * 65536
    assert metrics.flops_per_pass(65536, 26, 938, 1) == (27 + 26 * (2 + 18 + 938 * 40) + 1) * 65536


def test_evaluate_invalid_reports_dbl_max():
    r = metrics.evaluate(np.float32([10.0]), np.float32([20.0]), [1.0, 1.0, 1.0], 2, 1, 1, 1, 1, 1)
    assert not r.valid and r.ms.sum == np.finfo(np.float64).max and r.gflops == 0.0
