"""Validation and performance accounting of the reference driver.

Restates src/main.cpp:329-396 (difference/validate/evaluate) so Python
callers report exactly what the reference harness reports, plus the strict
metrics the reference does not have (symmetric relative error, error vs an
FP64 oracle).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

DIFF_TOLERANCE_PCT = 0.025  # src/bude.h:11
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X, AMD published (FP32 vector)


def validate(ref: np.ndarray, sample: np.ndarray) -> tuple[float, np.ndarray]:
    """(max_diff_%, failed indices) exactly as src/main.cpp:331-342: the
    relative difference |ref - sample| / ref is SIGNED in ref (negative
    references can never fail), entries with both |ref| < 1 and |sample| < 1
    are skipped, an entry fails when 100*diff > 0.025."""
    ref = np.asarray(ref, np.float32)[: len(sample)]
    sample = np.asarray(sample, np.float32)
    keep = ~((np.abs(ref) < 1.0) & (np.abs(sample) < 1.0))
    r = ref.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        diff = np.abs(r - sample.astype(np.float64)) / r
    diff = np.where(keep, diff, 0.0)
    maxdiff = max(0.0, float(np.max(diff))) if diff.size else 0.0
    failed = np.nonzero(keep & (100.0 * diff > DIFF_TOLERANCE_PCT))[0]
    return 100.0 * maxdiff, failed


def is_valid(max_diff_pct: float) -> bool:
    return max_diff_pct < DIFF_TOLERANCE_PCT  # src/main.cpp:371


def symmetric_max_rel_pct(ref: np.ndarray, sample: np.ndarray, floor: float = 1.0) -> float:
    """max |ref - sample| / |ref| in %, over entries with |ref| >= floor (no sign quirk)."""
    ref = np.asarray(ref, np.float64)[: len(sample)]
    sample = np.asarray(sample, np.float64)
    m = np.abs(ref) >= floor
    if not m.any():
        return 0.0
    return float(np.max(np.abs(ref[m] - sample[m]) / np.abs(ref[m])) * 100.0)


def flops_per_pass(nposes: int, natlig: int, natpro: int, ppwi: int) -> float:
    """The reference's FLOP accounting (src/main.cpp:358-360): per group of
    PPWI poses PPWI*27 + L*(2 + PPWI*18 + N*(10 + PPWI*30)) + PPWI."""
    ops_per_wg = ppwi * 27 + natlig * (2 + ppwi * 18 + natpro * (10 + ppwi * 30)) + ppwi
    return float(ops_per_wg) * (float(nposes) / float(ppwi))


def finsts_per_pass(nposes: int, natlig: int, natpro: int) -> float:
    """25 instructions per interaction (src/main.cpp:364)."""
    return 25.0 * natpro * natlig * nposes


def interactions_per_pass(nposes: int, natlig: int, natpro: int) -> int:
    return nposes * natlig * natpro  # src/main.cpp:368


@dataclass
class Summary:
    """SummaryStats (src/main.cpp:96-110) over the non-warmup iterations."""
    min: float
    max: float
    sum: float
    mean: float
    stddev: float

    @classmethod
    def of(cls, xs) -> "Summary":
        a = np.asarray(xs, np.float64)
        mean = float(a.sum() / a.size)
        return cls(float(a.min()), float(a.max()), float(a.sum()), mean,
                   float(np.sqrt(np.sum((a - mean) ** 2) / a.size)))


@dataclass
class Result:
    valid: bool
    max_diff_pct: float
    failed: np.ndarray
    ms: Summary
    gflops: float
    ginsts: float
    interactions_per_sec: float


def evaluate(ref: np.ndarray, energies: np.ndarray, kernel_ms: list[float], warmup: int, iterations: int,
             nposes: int, natlig: int, natpro: int, ppwi: int) -> Result:
    """src/main.cpp:344-396: average over the non-warmup iterations (sum /
    iterations), GFLOP/s, GFInst/s, interactions/s; an invalid result reports
    ms = DBL_MAX and zero rates."""
    ms_list = list(kernel_ms[warmup:])
    runtime = (sum(ms_list) / float(iterations)) * 1e-3
    max_diff, failed = validate(ref, energies)
    if not is_valid(max_diff):
        big = float(np.finfo(np.float64).max)
        return Result(False, max_diff, failed, Summary(big, big, big, big, 0.0), 0.0, 0.0, 0.0)
    return Result(True, max_diff, failed, Summary.of(ms_list),
                  flops_per_pass(nposes, natlig, natpro, ppwi) / runtime / 1e9,
                  finsts_per_pass(nposes, natlig, natpro) / runtime / 1e9,
                  interactions_per_pass(nposes, natlig, natpro) / runtime)
