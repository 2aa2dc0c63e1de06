"""Deck format round trips, the driver's deck checks, and the seeded
synthetic generator's determinism and distributions."""
import numpy as np
import pytest

from minibude_amd import deck as D


def test_generator_is_deterministic_and_seeded():
    a = D.synth_deck(200, 20, 300, seed=7)
    b = D.synth_deck(200, 20, 300, seed=7)
    c = D.synth_deck(200, 20, 300, seed=8)
    for f in ("protein", "ligand", "forcefield", "poses"):
        assert getattr(a, f).tobytes() == getattr(b, f).tobytes()
        assert getattr(a, f).tobytes() != getattr(c, f).tobytes()


def test_pose_stream_is_prefix_stable():
    full = D.synth_poses(5, 1000)
    assert np.array_equal(D.synth_poses(5, 100, first=450), full[:, 450:550])


def test_distributions():
    d = D.synth_deck(4096, 96, 20000, seed=1)
    r = np.sqrt(d.ligand["x"].astype(float) ** 2 + d.ligand["y"] ** 2 + d.ligand["z"] ** 2)
    assert r.max() <= 10.0 and r.max() > 8.5
    for c in "xyz":
        assert -20 <= d.protein[c].min() < -19 and 19 < d.protein[c].max() < 20
    assert set(d.forcefield["hbtype"].tolist()) <= {45, 69, 70}
    assert 1.0 <= d.forcefield["radius"].min() and d.forcefield["radius"].max() < 2.2
    assert np.all(np.abs(d.forcefield["hphb"]) <= 2) and np.all(np.abs(d.forcefield["elsc"]) <= 1)
    assert set(d.protein["type"].tolist()) == set(range(34))
    assert np.all(np.abs(d.poses[:3]) <= np.pi) and np.all(np.abs(d.poses[3:]) <= 5)
    assert abs(d.poses[:3].mean()) < 0.05 and abs(d.poses[3:].mean()) < 0.1


def test_round_trip_and_n_slicing(tmp_path):
    d = D.synth_deck(50, 7, 40, seed=2)
    d.ref_energies = np.arange(40, dtype=np.float32) * 1.5
    sha = D.write_deck(tmp_path, d)
    assert set(sha) == {"protein.in", "ligand.in", "forcefield.in", "poses.in", "ref_energies.out"}
    back = D.read_deck(tmp_path)
    assert back.protein.tobytes() == d.protein.tobytes()
    assert back.poses.tobytes() == d.poses.tobytes()
    sub = D.read_deck(tmp_path, nposes=10)
    # -n takes the first n of each of the 6 rows (src/main.cpp:305-311)
    assert np.array_equal(sub.poses, d.poses[:, :10])
    assert (tmp_path / "poses.in").stat().st_size == 6 * 40 * 4


def test_driver_deck_checks(tmp_path):
    d = D.synth_deck(50, 7, 40, seed=2)
    d.ref_energies = np.zeros(20, np.float32)
    D.write_deck(tmp_path, d)
    with pytest.raises(ValueError, match="less than poses"):
        D.read_deck(tmp_path)
    with pytest.raises(ValueError, match="exceeded max poses"):
        D.read_deck(tmp_path, nposes=41, require_ref=False)
    (tmp_path / "poses.in").write_bytes(b"\0" * 4 * 7)
    with pytest.raises(ValueError, match="not divisible"):
        D.read_deck(tmp_path, require_ref=False)
