"""bude_mi355x_fasten_noalloc: the plain fasten_main over caller device
memory with the deck's derived tables in the library's fixed scratch array
instead of allocated buffers.  Its energies must equal those of the
allocating entry points (bude_mi355x_create + bude_mi355x_fasten) on the same
deck bit for bit, on consecutive calls whose decks shrink in every dimension
(so nothing a larger earlier call left in the shared array leaks into a
smaller later one)."""
import ctypes

import numpy as np
import pytest

from conftest import GOLDEN
from minibude_amd import capi
from minibude_amd import deck as D
from minibude_amd.model import Context


class Hip:
    """Device buffers for the test (the caller's memory, allocated here by the caller)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a: np.ndarray) -> ctypes.c_void_p:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 4))) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def download(self, p: ctypes.c_void_p, n: int) -> np.ndarray:
        out = np.empty(n, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def noalloc(h: Hip, d: D.Deck, ppwi: int, wgsize: int, nposes: int | None = None,
            sentinel: float = np.nan) -> np.ndarray:
    """Energies of the first nposes poses through the no-allocation entry
    point, the pose rows uploaded with their full stride."""
    n = d.nposes if nposes is None else nposes
    prot, lig, ff = h.upload(d.protein), h.upload(d.ligand), h.upload(d.forcefield)
    poses = h.upload(np.ascontiguousarray(d.poses, np.float32))
    energies = h.upload(np.full(n, sentinel, np.float32))
    print(f"noalloc: natpro={d.natpro} natlig={d.natlig} ntypes={d.ntypes} nposes={n} stride={d.nposes} "
          f"ppwi={ppwi} wgsize={wgsize}", flush=True)
    capi.check(capi.lib().bude_mi355x_fasten_noalloc(0, prot, d.natpro, lig, d.natlig, ff, d.ntypes, poses,
                                                     d.nposes, n, ppwi, wgsize, energies, None))
    return h.download(energies, n)


def allocating(d: D.Deck, ppwi: int, wgsize: int) -> np.ndarray:
    with Context(d, 0) as ctx:
        return ctx.fasten(ppwi, wgsize, warmup=0, iterations=1).energies


def test_rejects_decks_beyond_the_scratch_without_device_work():
    # 2^20 protein atoms need 16 MiB of grouped coordinates: refused on the
    # sizes alone, before any HIP call (this runs without a GPU)
    null = ctypes.c_void_p(8)
    st = capi.lib().bude_mi355x_fasten_noalloc(0, null, 1 << 20, null, 26, null, 34, null, 64, 64, 1, 64, null,
                                               None)
    assert st == capi.BUDE_ERR_INVALID_ARGUMENT
    assert b"fixed scratch holds 4194304" in capi.lib().bude_mi355x_last_error()
    st = capi.lib().bude_mi355x_fasten_noalloc(0, null, 100, null, 26, null, 300, null, 64, 64, 1, 64, null, None)
    assert st == capi.BUDE_ERR_INVALID_ARGUMENT
    assert b"at most 256" in capi.lib().bude_mi355x_last_error()


@pytest.mark.gpu
def test_noalloc_equals_allocating_entry_point_on_shrinking_decks(gpu_device):
    h = Hip()
    try:
        # first call: the bm2 deck (2,672 / 2,672 atoms, 44 types), a pose
        # subset read with the full row stride
        big = D.read_deck(GOLDEN / "bm2", nposes=8192)
        e1 = noalloc(h, big, 8, 64)
        ref1 = allocating(big, 8, 64)
        assert np.array_equal(e1, ref1), np.max(np.abs(e1 - ref1))
        # second call: smaller in every dimension (938 / 26 atoms, 34 types,
        # 1,000 of 4,096 poses: a ragged pose count)
        small = D.workload_deck("c0", nposes=4096)
        e2 = noalloc(h, small, 1, 256, nposes=1000)
        ref2 = allocating(small.head(1000), 1, 256)
        assert np.array_equal(e2, ref2), np.max(np.abs(e2 - ref2))
        # and a third with another work shape on the same small deck
        e3 = noalloc(h, small, 4, 128)
        ref3 = allocating(small, 4, 128)
        assert np.array_equal(e3, ref3)
        assert np.array_equal(e3[:1000], e2)
    finally:
        h.free()


@pytest.mark.gpu
def test_noalloc_rejects_a_bad_type_and_leaves_energies_untouched(gpu_device):
    h = Hip()
    try:
        d = D.workload_deck("c0", nposes=256)
        d.protein = d.protein.copy()
        d.protein["type"][7] = d.ntypes  # one past the forcefield
        with pytest.raises(capi.BudeInvalidArgument, match="outside the forcefield"):
            noalloc(h, d, 1, 64, sentinel=-7.0)
        e = h.download(h.bufs[-1], 256)
        assert np.all(e == -7.0)
        # the next valid call is unaffected
        ok = D.workload_deck("c0", nposes=256)
        assert np.array_equal(noalloc(h, ok, 1, 64), allocating(ok, 1, 64))
    finally:
        h.free()
