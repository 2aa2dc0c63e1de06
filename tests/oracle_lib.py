"""ctypes access to the CPU checker (oracle/liboracle.so) and, where it was
built, the reference's own kernel (oracle/_ref/libref_kernel.so).  Test
infrastructure only: the product never loads these."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle" / "liboracle.so"
REF_KERNEL = ROOT / "oracle" / "_ref" / "libref_kernel.so"

_P, _S = ctypes.c_void_p, ctypes.c_size_t
_o = None


def oracle() -> ctypes.CDLL:
    global _o
    if _o is None:
        _o = ctypes.CDLL(str(ORACLE))
        _o.bude_oracle_validate.restype = ctypes.c_double
        _o.bude_oracle_validate.argtypes = [_P, _P, _S, ctypes.c_double, ctypes.POINTER(_S)]
        _o.bude_oracle_pair_terms_f64.argtypes = [_P, _P, ctypes.c_double, _P]
        _o.bude_oracle_transform_f64.argtypes = [ctypes.c_double] * 6 + [_P]
        _o.bude_oracle_transform_f32.argtypes = [ctypes.c_float] * 6 + [_P]
    return _o


def _args(deck, first, count):
    rows = [_P(deck.poses[k].ctypes.data) for k in range(6)]
    return [_S(deck.natlig), _S(deck.natpro), _P(deck.protein.ctypes.data), _P(deck.ligand.ctypes.data),
            _P(deck.forcefield.ctypes.data), *rows, _S(first), _S(count)]


def fasten_f32(deck, first: int = 0, count: int | None = None) -> np.ndarray:
    """FP32 restatement in the reference's order (makedeck/ref-kernel.cpp)."""
    count = deck.nposes - first if count is None else count
    out = np.zeros(count, np.float32)
    oracle().bude_oracle_fasten_f32(*_args(deck, first, count), _P(out.ctypes.data))
    return out


def fasten_f64(deck, first: int = 0, count: int | None = None) -> np.ndarray:
    """The same energy evaluated in FP64."""
    count = deck.nposes - first if count is None else count
    out = np.zeros(count, np.float64)
    oracle().bude_oracle_fasten_f64(*_args(deck, first, count), _P(out.ctypes.data))
    return out


def fasten_f64_mag(deck, first: int = 0, count: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """FP64 energies and magnitudes 0.5 * sum |term| (the cancellation scale)."""
    count = deck.nposes - first if count is None else count
    e = np.zeros(count, np.float64)
    m = np.zeros(count, np.float64)
    oracle().bude_oracle_fasten_f64_mag(*_args(deck, first, count), _P(e.ctypes.data), _P(m.ctypes.data))
    return e, m


def pair_terms_f64(pff, lff, d: float) -> np.ndarray:
    out = np.zeros(3, np.float64)
    p = np.ascontiguousarray(pff)
    l = np.ascontiguousarray(lff)
    oracle().bude_oracle_pair_terms_f64(_P(p.ctypes.data), _P(l.ctypes.data), d, _P(out.ctypes.data))
    return out


def validate_c(ref: np.ndarray, sample: np.ndarray, tol_pct: float = 0.025) -> tuple[float, int]:
    r = np.ascontiguousarray(ref, np.float32)
    s = np.ascontiguousarray(sample, np.float32)
    nf = _S(0)
    m = oracle().bude_oracle_validate(_P(r.ctypes.data), _P(s.ctypes.data), _S(len(s)), tol_pct, ctypes.byref(nf))
    return m, nf.value


def have_ref_kernel() -> bool:
    return REF_KERNEL.exists()


def ref_kernel(deck, first: int = 0, count: int | None = None) -> np.ndarray:
    """The reference's own per-pose kernel (makedeck/ref-kernel.cpp:15-139)."""
    lib = ctypes.CDLL(str(REF_KERNEL))
    count = deck.nposes - first if count is None else count
    out = np.zeros(count, np.float32)
    rows = [_P(deck.poses[k].ctypes.data) for k in range(6)]
    lib.ref_fasten(_S(deck.natlig), _S(deck.natpro), _P(deck.protein.ctypes.data), _P(deck.ligand.ctypes.data),
                   _P(deck.forcefield.ctypes.data), _S(deck.ntypes), *rows, _S(first), _S(count),
                   _P(out.ctypes.data))
    return out
