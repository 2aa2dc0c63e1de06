"""The mi355x model behind the reference's plugin interface.

Mirrors src/bude.h:54-101 (Params, Sample, Bude<PPWI>) so host code and
tests read like the reference's own driver: ``Mi355xBude(ppwi).fasten(params,
wgsize, device)`` returns a Sample whose energies the reference's validation
accepts.  All work happens in libbude_mi355x on the GPU.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .deck import Deck

PPWIS = (1, 2, 4, 8, 16, 32, 64, 128)  # src/main.cpp:48
DEFAULT_ITERS = 8                        # src/bude.h:12
DEFAULT_WARMUP = 2                       # src/main.cpp:160


@dataclass
class Params:
    """src/bude.h:54-73 (deck + run configuration)."""
    deck: Deck
    iterations: int = DEFAULT_ITERS
    warmup_iterations: int = DEFAULT_WARMUP

    def natpro(self) -> int:
        return self.deck.natpro

    def natlig(self) -> int:
        return self.deck.natlig

    def ntypes(self) -> int:
        return self.deck.ntypes

    def nposes(self) -> int:
        return self.deck.nposes

    def total_iterations(self) -> int:
        return self.iterations + self.warmup_iterations


@dataclass
class Sample:
    """src/bude.h:84-91; times in milliseconds, warmup iterations first."""
    ppwi: int
    wgsize: int
    energies: np.ndarray
    kernel_ms: list[float] = field(default_factory=list)   # host launch+sync (the reference's kernelTimes)
    device_ms: list[float] = field(default_factory=list)   # hipEvent time of each launch
    context_ms: float | None = None


def device_count() -> int:
    n = ctypes.c_int(0)
    st = capi.lib().bude_mi355x_device_count(ctypes.byref(n))
    if st == capi.BUDE_ERR_NO_DEVICE:
        return 0
    capi.check(st)
    return n.value


def device_name(index: int) -> str:
    buf = ctypes.create_string_buffer(256)
    capi.check(capi.lib().bude_mi355x_device_name(index, buf, len(buf)))
    return buf.value.decode()


class Context:
    """A deck resident in HBM (bude_mi355x_create) that can be scored many
    times without re-uploading; the timed region of every launch excludes the
    upload."""

    def __init__(self, deck: Deck, device: int = 0):
        self._h = ctypes.c_void_p()
        self.deck = deck
        pro = np.ascontiguousarray(deck.protein)
        lig = np.ascontiguousarray(deck.ligand)
        ff = np.ascontiguousarray(deck.forcefield)
        poses = np.ascontiguousarray(deck.poses, dtype=np.float32)
        rows = (ctypes.c_void_p * 6)(*[poses[k].ctypes.data for k in range(6)])
        capi.check(capi.lib().bude_mi355x_create(device, capi.ptr(pro), len(pro), capi.ptr(lig), len(lig),
                                                 capi.ptr(ff), len(ff), rows, poses.shape[1],
                                                 ctypes.byref(self._h)))
        ms = ctypes.c_double(0.0)
        capi.check(capi.lib().bude_mi355x_context_ms(self._h, ctypes.byref(ms)))
        self.context_ms = ms.value

    def fasten(self, ppwi: int, wgsize: int, warmup: int = 0, iterations: int = 1,
               want_energies: bool = True) -> Sample:
        n = warmup + iterations
        host = np.zeros(n, np.float64)
        dev = np.zeros(n, np.float64)
        energies = np.zeros(self.deck.nposes, np.float32) if want_energies else None
        capi.check(capi.lib().bude_mi355x_fasten(self._h, ppwi, wgsize, warmup, iterations,
                                                 None if energies is None else capi.ptr(energies),
                                                 capi.ptr(host), capi.ptr(dev)))
        return Sample(ppwi, wgsize, energies, host.tolist(), dev.tolist(), self.context_ms)

    def close(self) -> None:
        if self._h:
            capi.check(capi.lib().bude_mi355x_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


class Mi355xBude:
    """Bude<PPWI> (src/bude.h:95-101) implemented by libbude_mi355x."""

    def __init__(self, ppwi: int):
        if ppwi not in PPWIS:
            raise ValueError(f"PPWI {ppwi} is not a supported value, should be one of `{','.join(map(str, PPWIS))}`")
        self.ppwi = ppwi

    def name(self) -> str:
        return "mi355x"

    def enumerate_devices(self) -> list[tuple[int, str]]:
        return [(i, device_name(i)) for i in range(device_count())]

    def compatible(self, params: Params, wgsize: int, device: int) -> bool:
        ok = ctypes.c_int(0)
        st = capi.lib().bude_mi355x_compatible(device, self.ppwi, wgsize, ctypes.byref(ok))
        return st == capi.BUDE_OK and ok.value == 1

    def fasten(self, params: Params, wgsize: int, device: int) -> Sample:
        """Upload (context time), warmup + iterations launches, energies back,
        free: the shape of HipBude::fasten (src/hip/fasten.hpp:212-267)."""
        with Context(params.deck, device) as ctx:
            return ctx.fasten(self.ppwi, wgsize, params.warmup_iterations, params.iterations)
