"""Host side of the combined screen: C-struct layouts, the workload
generator, and argument checks that run before any device work."""
import ctypes
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT

from minibude_amd import capi
from minibude_amd import combined as C
from minibude_amd import deck as D
from minibude_amd import screen as S


def c_layout(struct: str, fields) -> dict:
    """sizeof and offsetof of a header struct, from the C compiler."""
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"bude_mi355x.h\"\nint main(void) {\n"
    src += f'  printf("size %zu\\n", sizeof({struct}));\n'
    for f in fields:
        src += f'  printf("{f} %zu\\n", offsetof({struct}, {f}));\n'
    src += "  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        (Path(d) / "l.c").write_text(src)
        subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(Path(d) / "l"), str(Path(d) / "l.c")],
                       check=True)
        out = subprocess.run([str(Path(d) / "l")], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


@pytest.mark.parametrize("cls,struct", [(C.CombinedConfig, "BudeCombinedConfig"),
                                        (C.CombinedResult, "BudeCombinedResult"),
                                        (C.PairResult, "BudePairResult"),
                                        (S.ScreenConfig, "BudeScreenConfig")])
def test_struct_layouts_match_the_header(cls, struct):
    names = [f for f, _ in cls._fields_]
    want = c_layout(struct, names)
    assert ctypes.sizeof(cls) == want["size"]
    for f in names:
        assert getattr(cls, f).offset == want[f], f


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
