"""Small hand-built decks for closed-form and edge-case tests."""
from __future__ import annotations

import numpy as np

from minibude_amd.capi import ATOM_DTYPE, FF_DTYPE
from minibude_amd.deck import Deck

NORMAL, HB_E, HB_F = 45, 69, 70


def ff(*rows) -> np.ndarray:
    """rows of (hbtype, radius, hphb, elsc)."""
    return np.array([tuple(r) for r in rows], dtype=FF_DTYPE)


def atoms(*rows) -> np.ndarray:
    """rows of (x, y, z, type)."""
    return np.array([tuple(r) for r in rows], dtype=ATOM_DTYPE)


def poses(*rows) -> np.ndarray:
    """rows of (tilt, roll, pan, tx, ty, tz) -> (6, n) float32."""
    return np.ascontiguousarray(np.array(rows, dtype=np.float32).T)


def pair_deck(pff, lff, distances, axis=0) -> Deck:
    """One protein atom (type 0) at the origin and one ligand atom (type 1) at
    the origin, with identity-rotation poses translating the ligand to each
    distance along `axis`."""
    rows = []
    for d in distances:
        t = [0.0, 0.0, 0.0]
        t[axis] = d
        rows.append((0.0, 0.0, 0.0, *t))
    return Deck(protein=atoms((0, 0, 0, 0)), ligand=atoms((0, 0, 0, 1)), forcefield=ff(pff, lff),
                poses=poses(*rows), name="pair")
