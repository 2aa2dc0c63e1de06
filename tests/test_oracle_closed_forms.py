"""Pin the oracle's pair energy to hand-evaluated closed forms, one case per
zone and per forcefield branch of src/serial/fasten.hpp:80-123.

Pose energy of one (protein, ligand) pair = 0.5 * (steric + charge + dslv).
Every expected value below is worked by hand from the reference formula.
"""
import numpy as np
import pytest

import oracle_lib as O
from handdecks import HB_E, HB_F, NORMAL, pair_deck

# protein: radius 1.5, hphb -1.0, elsc 0.5; ligand: radius 1.0, hphb -0.5, elsc 0.4
# radij = 2.5, chrg_init = 0.2, both hphb < 0 -> distdslv = NPNPDIST = 5.5,
# p_hphb = -1.0, l_hphb = -0.5 (no sign flips), dslv_init = -1.5.
P = (NORMAL, 1.5, -1.0, 0.5)
L = (NORMAL, 1.0, -0.5, 0.4)

CASES = [
    # d, (steric, charge, dslv)
    # d=1: b=-1.5 zone1: steric (1-1/2.5)*76 = 45.6; charge 0.2*45 = 9; dslv -1.5
    (1.0, (45.6, 9.0, -1.5)),
    # d=3: b=0.5: charge 0.2*(1-0.25)*45 = 6.75; dslv -1.5*(1-0.5/5.5) = -15/11
    (3.0, (0.0, 6.75, -15.0 / 11.0)),
    # d=5: b=2.5 >= elcdst=2: charge 0; dslv -1.5*(1-2.5/5.5) = -9/11
    (5.0, (0.0, 0.0, -9.0 / 11.0)),
    # d=9: b=6.5 beyond every cutoff
    (9.0, (0.0, 0.0, 0.0)),
]


@pytest.mark.parametrize("d,terms", CASES)
def test_pair_terms_f64(d, terms):
    got = O.pair_terms_f64(_row(P), _row(L), d)
    np.testing.assert_allclose(got, terms, rtol=1e-6, atol=1e-7)  # inputs are FP32-rounded


def _row(r):
    from handdecks import ff
    return ff(r)[0:1]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pose_energy_matches_closed_form(axis):
    deck = pair_deck(P, L, [c[0] for c in CASES], axis=axis)
    want = np.array([0.5 * sum(t) for _, t in CASES])
    np.testing.assert_allclose(O.fasten_f64(deck), want, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(O.fasten_f32(deck), want, rtol=2e-6, atol=1e-6)


def test_type_E_makes_charge_attractive():
    # either side 'E' (69): charge = -|chrg_init * f|.  At d=3: -6.75 (was +6.75)
    for p, l in (((HB_E, 1.5, -1.0, 0.5), L), (P, (HB_E, 1.0, -0.5, 0.4))):
        np.testing.assert_allclose(O.pair_terms_f64(_row(p), _row(l), 3.0), (0.0, -6.75, -15.0 / 11.0), rtol=1e-6)
    # negative chrg_init with E stays negative: elsc_p = -0.5 -> q = -0.2 -> -6.75
    np.testing.assert_allclose(O.pair_terms_f64(_row((HB_E, 1.5, -1.0, -0.5)), _row(L), 3.0)[1], -6.75, rtol=1e-6)
    # negative chrg_init without E: -0.2*0.75*45 = -6.75 as well (sign kept)
    np.testing.assert_allclose(O.pair_terms_f64(_row((NORMAL, 1.5, -1.0, -0.5)), _row(L), 3.0)[1], -6.75, rtol=1e-6)


def test_FF_pair_uses_long_charge_range():
    # both 'F' (70): elcdst = 4, elcdst1 = 0.25
    p, l = (HB_F, 1.5, -1.0, 0.5), (HB_F, 1.0, -0.5, 0.4)
    # d=3 (b=.5): 0.2*(1-0.125)*45 = 7.875; d=5.5 (b=3): 0.2*0.25*45 = 2.25; d=7 (b=4.5): 0
    for d, q in ((3.0, 7.875), (5.5, 2.25), (7.0, 0.0)):
        assert O.pair_terms_f64(_row(p), _row(l), d)[1] == pytest.approx(q, rel=1e-6, abs=1e-7)
    # only one F: short range, d=5 (b=2.5) -> 0
    assert O.pair_terms_f64(_row(p), _row(L), 5.0)[1] == 0.0


def test_desolvation_branches():
    # p > 0, l < 0: distdslv = NPPDIST = 1, l_hphb flips sign: dslv_init = 0.8 + 0.5
    p, l = (NORMAL, 1.5, 0.8, 0.0), (NORMAL, 1.0, -0.5, 0.0)
    assert O.pair_terms_f64(_row(p), _row(l), 2.0)[2] == pytest.approx(1.3)          # zone1
    assert O.pair_terms_f64(_row(p), _row(l), 3.0)[2] == pytest.approx(1.3 * 0.5)    # b=.5: 1-.5/1
    assert O.pair_terms_f64(_row(p), _row(l), 3.6)[2] == 0.0                         # b=1.1 > 1
    # p < 0, l > 0: distdslv = 1, p_hphb flips sign: dslv_init = 1.0 + 0.5
    p, l = (NORMAL, 1.5, -1.0, 0.0), (NORMAL, 1.0, 0.5, 0.0)
    assert O.pair_terms_f64(_row(p), _row(l), 3.0)[2] == pytest.approx(1.5 * 0.5)
    # both >= 0: distdslv = -FLT_MAX, term never on
    p, l = (NORMAL, 1.5, 1.0, 0.0), (NORMAL, 1.0, 0.5, 0.0)
    assert O.pair_terms_f64(_row(p), _row(l), 1.0)[2] == 0.0
    # protein hphb == 0: gated off even though the ligand is negative
    p, l = (NORMAL, 1.5, 0.0, 0.0), (NORMAL, 1.0, -0.5, 0.0)
    assert O.pair_terms_f64(_row(p), _row(l), 1.0)[2] == 0.0
    # ligand hphb == 0 with protein < 0: distdslv = NPPDIST, dslv_init = -1.0 + 0
    p, l = (NORMAL, 1.5, -1.0, 0.0), (NORMAL, 1.0, 0.0, 0.0)
    assert O.pair_terms_f64(_row(p), _row(l), 3.0)[2] == pytest.approx(-1.0 * 0.5)


def test_rotation_convention():
    # R = Rz(pan) * Ry(roll) * Rx(tilt) (src/serial/fasten.hpp:39-50)
    def apply(angles, v):
        t = np.zeros(12)
        O.oracle().bude_oracle_transform_f64(*angles, 0.0, 0.0, 0.0, t.ctypes.data)
        return t.reshape(3, 4)[:, :3] @ np.asarray(v, float)

    h = np.pi / 2
    np.testing.assert_allclose(apply((0, 0, h), (1, 0, 0)), (0, 1, 0), atol=1e-15)   # pan: x -> y
    np.testing.assert_allclose(apply((h, 0, 0), (0, 1, 0)), (0, 0, 1), atol=1e-15)   # tilt: y -> z
    np.testing.assert_allclose(apply((0, h, 0), (0, 0, 1)), (1, 0, 0), atol=1e-15)   # roll: z -> x
    a, b, c = 0.3, -1.1, 2.0
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    t = np.zeros(12)
    O.oracle().bude_oracle_transform_f64(a, b, c, 7.0, 8.0, 9.0, t.ctypes.data)
    np.testing.assert_allclose(t.reshape(3, 4)[:, :3], Rz @ Ry @ Rx, atol=1e-15)
    np.testing.assert_allclose(t.reshape(3, 4)[:, 3], (7, 8, 9))
