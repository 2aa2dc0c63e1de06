"""ctypes binding of libbude_mi355x (include/bude_mi355x.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the built library is missing, and every non-zero status raises with the
library's own message (RuntimeError for HIP failures, ValueError for invalid
arguments, mirroring std::runtime_error / std::invalid_argument in the
reference, src/hip/fasten.hpp:161-165 and src/serial/fasten.hpp:147-150).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libbude_mi355x.so"

BUDE_OK = 0
BUDE_ERR_INVALID_ARGUMENT = 1
BUDE_ERR_RUNTIME = 2
BUDE_ERR_NO_DEVICE = 3
BUDE_ERR_UNSUPPORTED = 4
BUDE_ERR_OUT_OF_MEMORY = 5

ATOM_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("type", "<i4")])
FF_DTYPE = np.dtype([("hbtype", "<i4"), ("radius", "<f4"), ("hphb", "<f4"), ("elsc", "<f4")])

# Every C-ABI entry point declared in include/bude_mi355x.h.
EXPORTED = (
    "bude_mi355x_last_error",
    "bude_mi355x_abi_version",
    "bude_mi355x_device_count",
    "bude_mi355x_device_name",
    "bude_mi355x_ppwi_list",
    "bude_mi355x_compatible",
    "bude_mi355x_create",
    "bude_mi355x_fasten",
    "bude_mi355x_context_ms",
    "bude_mi355x_destroy",
    "bude_mi355x_fasten_noalloc",
    "bude_mi355x_batch_create",
    "bude_mi355x_batch_run",
    "bude_mi355x_batch_run_separate",
    "bude_mi355x_batch_autotune",
    "bude_mi355x_batch_energies",
    "bude_mi355x_batch_clear",
    "bude_mi355x_batch_destroy",
    "bude_mi355x_screen_create",
    "bude_mi355x_screen_train",
    "bude_mi355x_screen_load",
    "bude_mi355x_screen_run",
    "bude_mi355x_screen_set_keep",
    "bude_mi355x_combined_create",
    "bude_mi355x_combined_screen",
    "bude_mi355x_combined_screen_batch",
    "bude_mi355x_combined_set_keep",
    "bude_mi355x_combined_pair",
    "bude_mi355x_combined_selected",
    "bude_mi355x_combined_scores",
    "bude_mi355x_combined_exhaustive",
    "bude_mi355x_combined_exact",
    "bude_mi355x_combined_recall",
    "bude_mi355x_combined_grid",
    "bude_mi355x_combined_map",
    "bude_mi355x_combined_destroy",
    "bude_mi355x_screen_selected",
    "bude_mi355x_screen_predict",
    "bude_mi355x_screen_exhaustive",
    "bude_mi355x_screen_destroy",
    "bude_mi355x_mlp_forward_backward",
    "bude_mi355x_mlp_predict",
    "bude_mi355x_atom_energies",
    "bude_mi355x_atomnet_forward_backward",
    "bude_mi355x_atomnet_score",
    "bude_synth_forcefield",
    "bude_synth_protein",
    "bude_synth_ligand",
    "bude_synth_poses",
    "bude_mi355x_combined_labels",
    "bude_forcefield_read",
    "bude_forcefield_info",
    "bude_forcefield_params",
    "bude_forcefield_destroy",
    "bude_mol2_read",
    "bude_poses_generate",
)


class BatchConfig(ctypes.Structure):
    """BudeBatchConfig (include/bude_mi355x.h)."""
    _fields_ = [("ppwi", ctypes.c_int32), ("wgsize", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32),
                ("order", ctypes.c_int32), ("use_graph", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class BudeError(RuntimeError):
    """Non-zero status from libbude_mi355x."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class BudeInvalidArgument(BudeError, ValueError):
    pass


class BudeUnsupported(BudeError, ValueError):
    pass


class BudeNoDevice(BudeError):
    pass


_lib: ctypes.CDLL | None = None

_P = ctypes.c_void_p
_S = ctypes.c_size_t
_I = ctypes.c_int


def _declare(lib: ctypes.CDLL) -> None:
    sig = {
        "bude_mi355x_last_error": (ctypes.c_char_p, []),
        "bude_mi355x_abi_version": (_I, []),
        "bude_mi355x_device_count": (_I, [ctypes.POINTER(_I)]),
        "bude_mi355x_device_name": (_I, [_I, ctypes.c_char_p, _S]),
        "bude_mi355x_ppwi_list": (_I, [ctypes.POINTER(_S), _S, ctypes.POINTER(_S)]),
        "bude_mi355x_compatible": (_I, [_I, _S, _S, ctypes.POINTER(_I)]),
        "bude_mi355x_create": (_I, [_I, _P, _S, _P, _S, _P, _S, ctypes.POINTER(_P), _S, ctypes.POINTER(_P)]),
        "bude_mi355x_fasten": (_I, [_P, _S, _S, _S, _S, _P, _P, _P]),
        "bude_mi355x_context_ms": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
        "bude_mi355x_destroy": (_I, [_P]),
        "bude_mi355x_fasten_noalloc": (_I, [_I, _P, _S, _P, _S, _P, _S, _P, _S, _S, _S, _S, _P, _P]),
        "bude_mi355x_atom_energies": (_I, [_I, _P, _S, _P, _S, _P, _P, _S, _P]),
        "bude_mi355x_atomnet_forward_backward": (_I, [_I, _S, _P, _P, _P, _P, ctypes.c_float, _P, _P, _S,
                                                      ctypes.POINTER(_P), ctypes.POINTER(_P), _P, _P, _P,
                                                      ctypes.POINTER(_P), ctypes.POINTER(_P)]),
        "bude_mi355x_atomnet_score": (_I, [_I, _P, _S, ctypes.POINTER(_P), _S, _P, ctypes.c_float, _P, _P, _S,
                                           ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.c_float, ctypes.c_float,
                                           _P]),
        "bude_mi355x_batch_create": (_I, [_I, _P, _S, _P, _S, _S, ctypes.POINTER(_P), ctypes.POINTER(_S),
                                          ctypes.POINTER(_P), ctypes.POINTER(_S), ctypes.POINTER(_P)]),
        "bude_mi355x_batch_run": (_I, [_P, _P, _S, _S, _P]),
        "bude_mi355x_batch_run_separate": (_I, [_P, _S, _S, _S, _S, _P]),
        "bude_mi355x_batch_autotune": (_I, [_P, _P, ctypes.POINTER(ctypes.c_double)]),
        "bude_mi355x_batch_energies": (_I, [_P, _S, _P]),
        "bude_mi355x_batch_clear": (_I, [_P]),
        "bude_mi355x_batch_destroy": (_I, [_P]),
        "bude_synth_forcefield": (_I, [ctypes.c_uint64, _S, _P]),
        "bude_synth_protein": (_I, [ctypes.c_uint64, _S, _S, ctypes.c_float, _P]),
        "bude_synth_ligand": (_I, [ctypes.c_uint64, _S, _S, ctypes.c_float, _P]),
        "bude_synth_poses": (_I, [ctypes.c_uint64, _S, _S, ctypes.c_float, _P, _S]),
        "bude_forcefield_read": (_I, [ctypes.c_char_p, ctypes.POINTER(_P)]),
        "bude_forcefield_info": (_I, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
        "bude_forcefield_params": (_I, [_P, _P, _S, ctypes.POINTER(_S)]),
        "bude_forcefield_destroy": (_I, [_P]),
        "bude_mol2_read": (_I, [_P, ctypes.c_char_p, _P, _S, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
        "bude_poses_generate": (_I, [_S, ctypes.c_uint64, ctypes.POINTER(_P), ctypes.POINTER(_S), ctypes.POINTER(_P)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("BUDE_MI355X_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"libbude_mi355x not found at {path}: build it first (python -m minibude_amd.build); "
                "there is no CPU fallback")
        _lib = ctypes.CDLL(str(path))
        _declare(_lib)
    return _lib


def check(status: int) -> None:
    if status == BUDE_OK:
        return
    msg = lib().bude_mi355x_last_error().decode(errors="replace")
    cls = {
        BUDE_ERR_INVALID_ARGUMENT: BudeInvalidArgument,
        BUDE_ERR_UNSUPPORTED: BudeUnsupported,
        BUDE_ERR_NO_DEVICE: BudeNoDevice,
    }.get(status, BudeError)
    raise cls(status, msg)


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)
