"""minibude_amd -- an MI355X-native implementation of miniBUDE's fasten_main.

The compute lives in libbude_mi355x (hand-written gfx950 HIP kernels behind
the C ABI in include/bude_mi355x.h).  This package is the host-side mirror of
the reference's plugin interface (src/bude.h) and driver helpers:

    capi     ctypes binding of the C ABI (fails loudly if the library is absent)
    deck     deck I/O in the reference format, seeded synthetic decks
    model    Params / Sample / Mi355xBude (Bude<PPWI>) / resident Context
    metrics  the reference driver's validation and FLOP accounting
"""
from .capi import ATOM_DTYPE, FF_DTYPE, BudeError  # noqa: F401

__version__ = "0.1.0"
