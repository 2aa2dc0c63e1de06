"""Decks from real molecules (the makedeck role): bhff/mol2 parsing, the pose
sampler and the bude-makedeck command line, on the CPU.

Pins:
* the bm2 deck (tests/golden/bm2, made by makedeck from
  2_heavy_by-atom_2016-v1.bhff and bude_1gcz_proA.mol2 with seed 42,
  data/bm2/params.txt) is rebuilt byte for byte: forcefield, protein,
  ligand and all 65,536 poses in makedeck's order;
* the reference's only makedeck unit test (makedeck/test-bude.cpp:94-111);
* the pose draws against 32-bit MT19937 words from numpy's RandomState
  (the same generator, seeded the same way) through the divide-and-reject
  mapping;
* the same protein in its two mol2 dialects (Sybyl-typed, 9 columns, and
  residue/atom-named, 10 columns) gives the same heavy atoms;
* counts read off the files' text (declared rows, hydrogen rows,
  conformation blocks) and the error cases.
The bhff/mol2 fixtures are files of makedeck/raw/.
"""
import math
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from minibude_amd import capi
from minibude_amd import makedeck as M

MK = GOLDEN / "makedeck"
BHFF_ATOM = MK / "2_heavy_by-atom_2016-v1.bhff"
BHFF_RESIDUE = MK / "4_heavy_by-residue_2016-v1.bhff"
BHFF_OLD = MK / "0_force_field.bhff"
PROTEIN = MK / "1212_bude_1gcz_proA.mol2"
PROTEIN_SYB = MK / "1213_bude_1gcz_proA_syb.mol2"
LIGAND_CONF = MK / "1186_EXPERIMENTAL.mol2"
LIGAND_SMALL = MK / "0_1GCZ_R01_S0001_C0000_L00001_S0001_C0000_G0010_E0001.mol2"
UNTYPED = MK / "1184_2ZTA_chainA_syb.mol2"
TOOL = ROOT / "minibude_amd" / "bin" / "bude-makedeck"


def content(path):
    out = []
    for line in open(path, errors="replace"):
        t = line.strip()
        if t and t[0] not in "#%":
            out.append(t)
    return out


def declared(path):
    """(residue groups, rows) as the file's 'NAME COUNT' lines declare them."""
    groups = {}
    for t in content(path):
        c = t.split()
        if len(c) == 2:
            groups[c[0]] = groups.get(c[0], 0) + int(c[1])
    return len(groups), sum(groups.values())


@pytest.mark.parametrize("bhff", [BHFF_ATOM, BHFF_RESIDUE, BHFF_OLD], ids=lambda p: p.name)
def test_bhff_parses_to_declared_rows(bhff):
    with M.Forcefield(bhff) as ff:
        assert (ff.nresidues, len(ff.params())) == declared(bhff)


def test_bm2_deck_rebuilt_byte_for_byte():
    d = M.build_deck(BHFF_ATOM, PROTEIN, PROTEIN, nposes=65536, seed=42)
    bm2 = GOLDEN / "bm2"
    assert d.forcefield.tobytes() == (bm2 / "forcefield.in").read_bytes()
    assert d.protein.tobytes() == (bm2 / "protein.in").read_bytes()
    assert d.ligand.tobytes() == (bm2 / "ligand.in").read_bytes()
    assert d.poses.tobytes() == (bm2 / "poses.in").read_bytes()


def test_reference_unit_test_one_pose():
    # makedeck/test-bude.cpp:94-111: one value per field -> the single pose is all ones
    grid = [np.array([180.0 / math.pi], np.float32)] * 3 + [np.array([1.0], np.float32)] * 3
    assert np.array_equal(M.generate_poses(1, 42, grid), np.ones((6, 1), np.float32))


def below(words, n):
    scaling = 0xFFFFFFFF // n
    past = n * scaling
    while True:
        w = int(next(words))
        if w < past:
            return w // scaling


@pytest.mark.parametrize("seed", [42, 7, 2**32 + 5])
def test_first_pose_follows_mt19937_words(seed):
    grid = [M.GRID_DEGREES] * 3 + [M.GRID_TRANSLATIONS] * 3
    rs = np.random.RandomState(seed % 2**32)  # std::mt19937 takes the seed modulo 2^32
    words = iter(rs.randint(0, 2**32, size=64, dtype=np.uint64))
    idx = [below(words, len(g)) for g in grid]
    want = [np.float32(float(grid[k][i]) * (math.pi / 180.0)) if k < 3 else grid[k][i] for k, i in enumerate(idx)]
    got = M.generate_poses(1, seed, grid)[:, 0]
    assert np.array_equal(got, np.array(want, np.float32))


def test_poses_are_distinct_grid_points():
    p = M.generate_poses(8192, 3)
    assert len({tuple(c) for c in p.T.tolist()}) == 8192
    ang = np.float32([np.float32(float(d) * (math.pi / 180.0)) for d in M.GRID_DEGREES])
    assert np.isin(p[:3], ang).all() and np.isin(p[3:], M.GRID_TRANSLATIONS).all()
    # a grid of 2 x 2 x 2 x 1 x 1 x 2 = 16 combinations: asking for all 16 returns each once
    grid = [np.float32([0, 90])] * 3 + [np.float32([0])] * 2 + [np.float32([1, 2])]
    small = M.generate_poses(16, 9, grid)
    assert len({tuple(c) for c in small.T.tolist()}) == 16
    with pytest.raises(capi.BudeError, match="exceed"):
        M.generate_poses(17, 9, grid)


def test_two_mol2_dialects_agree():
    a, na = M.Forcefield(BHFF_ATOM).molecule(PROTEIN)          # Sybyl types, 9 columns
    b, nb = M.Forcefield(BHFF_RESIDUE).molecule(PROTEIN_SYB)   # residue + atom name, 10 columns
    assert len(a) == len(b) == 2672 and na == nb == 0
    for k in ("x", "y", "z"):
        assert np.array_equal(a[k], b[k])


def test_residue_forcefield_types_are_indices_within_the_residue():
    # first atom of the residue-named protein: N of PRO -> its row index inside the PRO group
    rows, lines = [], content(BHFF_RESIDUE)
    for i, t in enumerate(lines):
        c = t.split()
        if len(c) == 2 and c[0] == "PRO":
            rows = [lines[i + 1 + j].split()[0] for j in range(int(c[1]))]
            break
    b, _ = M.Forcefield(BHFF_RESIDUE).molecule(PROTEIN_SYB)
    assert int(b["type"][0]) == rows.index("N")
    assert int(b["type"].max()) < max(declared_group_sizes(BHFF_RESIDUE))


def declared_group_sizes(path):
    return [int(t.split()[1]) for t in content(path) if len(t.split()) == 2]


def test_hydrogens_skipped_and_conformations_counted():
    for mol2 in (LIGAND_SMALL, LIGAND_CONF):
        lines = content(mol2)
        i = lines.index("@<TRIPOS>ATOM") + 1
        j = next(k for k in range(i, len(lines)) if lines[k].startswith("@<TRIPOS>"))
        heavy = sum(1 for t in lines[i:j] if t.split()[5] not in ("H", "h"))
        nconf = sum(1 for t in lines[j:] if t.split()[:1] == ["@<BUDE>CONF"] and len(t.split()) == 2)
        atoms, got_conf = M.Forcefield(BHFF_ATOM).molecule(mol2)
        assert (len(atoms), got_conf) == (heavy, nconf)
    assert M.Forcefield(BHFF_ATOM).molecule(LIGAND_CONF)[1] == 2


def test_parse_errors_name_the_line(tmp_path):
    with pytest.raises(capi.BudeError, match="needs 9-column atom rows"):
        M.Forcefield(BHFF_ATOM).molecule(UNTYPED)
    with pytest.raises(capi.BudeError, match=r"1184_2ZTA_chainA_syb.mol2:\d+: no forcefield entry ACE.C"):
        M.Forcefield(BHFF_RESIDUE).molecule(UNTYPED)
    with pytest.raises(capi.BudeError, match="unable to open"):
        M.Forcefield(tmp_path / "missing.bhff")
    nomark = tmp_path / "nomark.mol2"
    nomark.write_text("@<TRIPOS>MOLECULE\nx\n")
    with pytest.raises(capi.BudeError, match="atom section marker"):
        M.Forcefield(BHFF_ATOM).molecule(nomark)
    bad = tmp_path / "bad.bhff"
    bad.write_text("# comment\nWLD 2\nC.3 - 1.42 -0.47 38.0 5.5 1.0 1.0 0.0\nC.2 - 1.42\n")
    with pytest.raises(capi.BudeError, match=r"bad.bhff:4: an atom row needs 9 columns"):
        M.Forcefield(bad)
    short = tmp_path / "short.bhff"
    short.write_text("WLD 3\nC.3 - 1.42 -0.47 38.0 5.5 1.0 1.0 0.0\n")
    with pytest.raises(capi.BudeError, match="declares 3 rows"):
        M.Forcefield(short)


def test_scaled_radius_and_hbtype_code(tmp_path):
    f = tmp_path / "one.bhff"
    f.write_text("WLD 2\nC.3 - 1.5 -0.25 38.0 5.5 1.0 2.0 0.5\nO.3 E 1.25 0.75 0 0 0 1.0 -1.0\n")
    p = M.Forcefield(f).params()
    assert p["hbtype"].tolist() == [ord("-"), ord("E")]
    assert p["radius"].tolist() == [np.float32(1.5) * np.float32(2.0), 1.25]
    assert p["hphb"].tolist() == [-0.25, 0.75] and p["elsc"].tolist() == [0.5, -1.0]


def run_tool(*args):
    return subprocess.run([str(TOOL), *map(str, args)], capture_output=True, text=True, timeout=60)


def test_tool_command_line_checks(tmp_path):
    r = run_tool("--help")
    assert r.returncode == 0
    for flag in ("--forcefield", "--protein", "--ligand", "--pose-seed", "--pose-length", "--out", "--force"):
        assert flag in r.stdout
    r = run_tool("-f", BHFF_ATOM, "-p", PROTEIN, "-l", PROTEIN, "-o", tmp_path / "d", "-n", "1000")
    assert r.returncode == 1 and "cannot be divided by WGSIZE(256)" in r.stderr
    r = run_tool("-f", BHFF_ATOM, "-p", PROTEIN, "-o", tmp_path / "d")
    assert r.returncode == 1 and "--ligand is a required parameter" in r.stderr
    r = run_tool("-f", tmp_path / "nope.bhff")
    assert r.returncode == 1 and "Unable to open" in r.stderr
    r = run_tool("-n", "x")
    assert r.returncode == 1 and "integer required for <nPoses>" in r.stderr
    r = run_tool("--bogus")
    assert r.returncode == 1 and "Unrecognized argument '--bogus'" in r.stderr
    (tmp_path / "exists").mkdir()
    r = run_tool("-f", BHFF_ATOM, "-p", PROTEIN, "-l", PROTEIN, "-o", tmp_path / "exists")
    assert r.returncode == 1 and "pass --force" in r.stderr
