"""Invariants of the pose energy that pin the oracle where no golden vector
reaches: additivity, permutation, rigid motion, periodicity, far field."""
import numpy as np
import pytest

import oracle_lib as O
from minibude_amd import deck as deckmod
from minibude_amd.deck import Deck


@pytest.fixture(scope="module")
def small():
    return deckmod.synth_deck(natpro=300, natlig=12, nposes=64, seed=11)


def _with(d, **kw):
    f = dict(protein=d.protein, ligand=d.ligand, forcefield=d.forcefield, poses=d.poses)
    f.update(kw)
    return Deck(**f)


def test_additive_over_ligand_atoms(small):
    a = _with(small, ligand=small.ligand[:5].copy())
    b = _with(small, ligand=small.ligand[5:].copy())
    np.testing.assert_allclose(O.fasten_f64(a) + O.fasten_f64(b), O.fasten_f64(small), rtol=1e-12, atol=1e-9)


def test_additive_over_protein_atoms(small):
    a = _with(small, protein=small.protein[:123].copy())
    b = _with(small, protein=small.protein[123:].copy())
    np.testing.assert_allclose(O.fasten_f64(a) + O.fasten_f64(b), O.fasten_f64(small), rtol=1e-12, atol=1e-9)


def test_permutation_invariant(small):
    rng = np.random.default_rng(0)
    d = _with(small, protein=small.protein[rng.permutation(small.natpro)].copy(),
              ligand=small.ligand[rng.permutation(small.natlig)].copy())
    np.testing.assert_allclose(O.fasten_f64(d), O.fasten_f64(small), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(O.fasten_f32(d), O.fasten_f32(small), rtol=2e-5, atol=2e-3)


def test_rigid_translation(small):
    # moving the protein by s and every pose translation by s keeps E
    s = np.array([0.5, -1.25, 2.0], np.float32)   # shifted coordinates round to FP32 once
    pro = small.protein.copy()
    for k, c in enumerate("xyz"):
        pro[c] += s[k]
    poses = small.poses.copy()
    poses[3:] += s[:, None]
    d = _with(small, protein=pro, poses=poses)
    np.testing.assert_allclose(O.fasten_f64(d), O.fasten_f64(small), rtol=2e-5, atol=2e-5)


def test_angle_period(small):
    poses = small.poses.astype(np.float64)
    poses[:3] += 2 * np.pi
    d = _with(small, poses=poses.astype(np.float32))
    # the angle shift is rounded to FP32, so allow the rotation error it causes
    np.testing.assert_allclose(O.fasten_f64(d), O.fasten_f64(small), rtol=2e-4, atol=2e-3)


def test_far_field_is_exactly_zero(small):
    poses = small.poses.copy()
    poses[3] += 1000.0
    d = _with(small, poses=poses)
    assert np.all(O.fasten_f64(d) == 0.0)
    assert np.all(O.fasten_f32(d) == 0.0)
