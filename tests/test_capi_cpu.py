"""The C-ABI library loads, exports every declared entry point, and rejects
bad arguments without touching a GPU."""
import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT
from minibude_amd import capi, model
from minibude_amd.deck import synth_deck


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        names |= set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bude_\w+)\s*\(", text, re.M))
    return names


def test_header_declares_what_the_binding_binds():
    assert declared_functions() == set(capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.bude_mi355x_abi_version() == 1


def test_ppwi_list():
    out = (ctypes.c_size_t * 16)()
    n = ctypes.c_size_t()
    capi.check(capi.lib().bude_mi355x_ppwi_list(out, 16, ctypes.byref(n)))
    assert list(out[: n.value]) == [1, 2, 4, 8, 16, 32, 64, 128]


def test_compatible_rejects_unsupported_shapes():
    ok = ctypes.c_int(1)
    assert capi.lib().bude_mi355x_compatible(0, 3, 64, ctypes.byref(ok)) == capi.BUDE_ERR_UNSUPPORTED
    assert ok.value == 0
    assert capi.lib().bude_mi355x_compatible(0, 8, 96, ctypes.byref(ok)) == capi.BUDE_ERR_UNSUPPORTED
    assert b"wgsize" in capi.lib().bude_mi355x_last_error()


def test_create_validates_deck_before_any_device_work():
    d = synth_deck(16, 4, 8, seed=3)
    d.protein["type"][5] = 99
    with pytest.raises(capi.BudeInvalidArgument, match="protein atom 5 has type 99"):
        model.Context(d)
    d = synth_deck(16, 4, 8, seed=3)
    d.ligand["type"][1] = -1
    with pytest.raises(capi.BudeInvalidArgument, match="ligand atom 1"):
        model.Context(d)


def test_no_device_is_reported_not_faked():
    if model.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.BudeNoDevice):
        model.Context(synth_deck(16, 4, 8, seed=3))


def test_model_rejects_unknown_ppwi():
    with pytest.raises(ValueError, match="PPWI 3"):
        model.Mi355xBude(3)
    assert model.Mi355xBude(8).name() == "mi355x"
