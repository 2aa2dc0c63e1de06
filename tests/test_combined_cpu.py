"""Host side of the combined screen: C-struct layouts, the workload
generator, and argument checks that run before any device work."""
import ctypes

import numpy as np
import pytest

from minibude_amd import capi
from minibude_amd import combined as C
from minibude_amd import deck as D


def test_struct_layouts_match_the_header():
    assert ctypes.sizeof(C.CombinedConfig) == 56
    # pipeline_chunks sits in what was the padding before the 8-byte seed
    assert C.CombinedConfig.pipeline_chunks.offset == 44 and C.CombinedConfig.seed.offset == 48
    assert ctypes.sizeof(C.CombinedResult) == 64
    assert ctypes.sizeof(C.PairResult) == 20


def test_workload_is_seeded_and_in_range():
    a = C.combined_workload(nproteins=3, nligands=4, nposes=100, seed=7)
    b = C.combined_workload(nproteins=3, nligands=4, nposes=100, seed=7)
    assert all(np.array_equal(x, y) for x, y in zip(a.proteins, b.proteins))
    assert all(np.array_equal(x, y) for x, y in zip(a.poses, b.poses))
    assert all(900 <= len(p) <= 6000 for p in a.proteins)
    assert all(8 <= len(l) <= 96 for l in a.ligands)
    # ligand i's poses are the deck generator's stream seed*100003 + i + 1
    assert np.array_equal(a.poses[2], D.synth_poses(7 * 100003 + 3, 100))
    assert a.pairs == 12 and a.total_poses == 3 * 4 * 100
    assert a.interactions == sum(len(p) for p in a.proteins) * sum(len(l) * 100 for l in a.ligands)


def test_create_checks_config_and_ligands_first():
    w = C.combined_workload(nproteins=1, nligands=2, nposes=64, seed=7)
    with pytest.raises(capi.BudeInvalidArgument, match="keep_fraction"):
        C.Combined(w.forcefield, w.ligands, w.poses, C.default_config(keep_fraction=0.0))
    cfg = C.default_config()
    cfg.surrogate = 7
    with pytest.raises(capi.BudeUnsupported, match="surrogate"):
        C.Combined(w.forcefield, w.ligands, w.poses, cfg)
    with pytest.raises(capi.BudeInvalidArgument, match="grid_spacing"):
        C.Combined(w.forcefield, w.ligands, w.poses, C.default_config("field", grid_spacing=0.0))
    with pytest.raises(capi.BudeInvalidArgument, match="pipeline_chunks"):
        C.Combined(w.forcefield, w.ligands, w.poses, C.default_config("field", pipeline_chunks=17))
    bad = [l.copy() for l in w.ligands]
    bad[1]["type"][0] = 99
    with pytest.raises(capi.BudeInvalidArgument, match="ligand 1"):
        C.Combined(w.forcefield, bad, w.poses, C.default_config())
