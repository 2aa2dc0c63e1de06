"""The combined screen on the GPU: single-atom energy maps against the plain
kernel, the interpolating scorer against a numpy restatement, the energy
decomposition on grid points, per-pair selection / exact rescoring / recall
for both surrogates, and per-pair fallback."""
import numpy as np
import pytest

import oracle_lib as O
from minibude_amd import capi
from minibude_amd import combined as C
from minibude_amd import deck as D
from minibude_amd import screen as S
from minibude_amd.model import Context

pytestmark = pytest.mark.gpu


def small_workload(nposes=8192, nligands=5, nproteins=2, seed=11):
    return C.combined_workload(nproteins=nproteins, nligands=nligands, nposes=nposes, seed=seed,
                               natpro_range=(300, 900), natlig_range=(4, 24))


def kept(keep, n):
    """Poses a pair keeps: ceil(keep * n) with keep as the C float the library holds."""
    return max(1, int(np.ceil(float(np.float32(keep)) * n)))


def plain_energies(protein, ligand, ff, poses, device):
    with Context(D.Deck(protein, ligand, ff, np.ascontiguousarray(poses, np.float32)), device) as c:
        return c.fasten(1, 64).energies


def rotations(p):
    a = p[:3].astype(np.float64)
    sx, sy, sz = np.sin(a)
    cx, cy, cz = np.cos(a)
    return np.array([[cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
                     [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
                     [-sy, sx * cy, cx * cy]]).transpose(2, 0, 1)  # (n, 3, 3)


def trilinear(maps, lo, h, n, pos):
    u = (pos - lo) / h
    i0 = np.clip(np.floor(u).astype(np.int64), 0, np.array(n) - 2)
    f = np.clip(u - i0, 0.0, 1.0)
    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                    (f[:, 2] if dz else 1 - f[:, 2])
                out = out + w * maps[i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz]
    return out


def test_field_maps_are_single_atom_plain_energies(gpu_device):
    w = small_workload(nposes=2048, nligands=3, nproteins=1)
    cfg = C.default_config("field", grid_spacing=1.0, max_val_rmse=1e9, gate_z=0.0)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        c.screen(w.proteins[0])
        lo, h, n = c.grid()
        assert h == 1.0 and np.all(lo == np.round(lo))  # grid coordinates exact in FP32
        t = int(w.ligands[0]["type"][0])
        m = c.map(t)
    g = [lo[k] + h * np.arange(n[k], dtype=np.float32) for k in range(3)]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    poses = np.zeros((6, X.size), np.float32)
    poses[3], poses[4], poses[5] = X.ravel(), Y.ravel(), Z.ravel()
    atom = np.array([(0.0, 0.0, 0.0, t)], dtype=w.ligands[0].dtype)
    want = plain_energies(w.proteins[0], atom, w.forcefield, poses, gpu_device)
    # same pair loop, same group order, same point: the same FP32 value
    assert np.array_equal(m.ravel().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("groups,chunks,launch_tiles", [("8", 0, 0), ("8", 3, 0), ("8", 2, 5), ("1", 0, 0)])
def test_field_score_is_trilinear_sum_over_atoms(gpu_device, monkeypatch, groups, chunks, launch_tiles):
    # groups: the scorer's atoms split by map slot over the 8 XCDs (partial
    # sums added in a second kernel; with chunks, launches from tile0 > 0;
    # launch_tiles: each chunk scored in launches of that many tiles, a last
    # one shorter), or one group, the form without the split
    monkeypatch.setenv("BUDE_MI355X_FIELD_GROUPS", groups)
    if launch_tiles:
        monkeypatch.setenv("BUDE_MI355X_FIELD_LAUNCH_TILES", str(launch_tiles))
    w = small_workload(nposes=4096, nligands=5, nproteins=1)
    cfg = C.default_config("field", grid_spacing=0.5, max_val_rmse=1e9, gate_z=0.0, pipeline_chunks=chunks)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        c.screen(w.proteins[0])
        lo, h, n = c.grid()
        for l in range(len(w.ligands)):
            assert c.pair(l).fallback == 0
            got = c.scores(l)
            lig = w.ligands[l]
            p = w.poses[l]
            R = rotations(p)
            t = p[3:].T.astype(np.float64)
            want = np.zeros(p.shape[1])
            mag = np.zeros(p.shape[1])
            maps = {}
            for a in range(len(lig)):
                ty = int(lig["type"][a])
                if ty not in maps:  # the scorer reads the corner values as FP16
                    maps[ty] = np.clip(c.map(ty), -65504, 65504).astype(np.float16).astype(np.float64)
                x = np.array([lig["x"][a], lig["y"][a], lig["z"][a]], np.float64)
                v = trilinear(maps[ty], lo.astype(np.float64), h, n, R @ x + t)
                want += v
                mag += np.abs(v)
            # FP32 interpolation and sum vs FP64: relative to the sum of |terms|
            assert np.max(np.abs(got - want) / np.maximum(mag, 1.0)) < 1e-4, f"ligand {l}"


def test_field_is_exact_when_atoms_sit_on_grid_points(gpu_device):
    # identity rotations, translations and atom coordinates on the 0.5 grid:
    # the field score is the sum of exact single-atom energies, i.e. the
    # exact pose energy up to FP32 summation order
    ff = D.synth_forcefield(21, 34)
    protein = D.synth_protein(21, 700, 34)
    rng = np.random.default_rng(3)
    lig = np.zeros(12, dtype=protein.dtype)
    for k in "xyz":
        lig[k] = rng.integers(-12, 13, size=12) * 0.5
    lig["type"] = rng.integers(0, 34, size=12)
    n = 1024
    poses = np.zeros((6, n), np.float32)
    poses[3:] = rng.integers(-10, 11, size=(3, n)) * 0.5
    cfg = C.default_config("field", grid_spacing=0.5, max_val_rmse=1e9, gate_z=0.0, labels_per_pair=64)
    with C.Combined(ff, [lig], [poses], cfg, gpu_device) as c:
        c.screen(protein)
        got = c.scores(0)
        lo, h, _ = c.grid()
        maps = {int(t): c.map(int(t)).astype(np.float64) for t in np.unique(lig["type"])}
    exact = plain_energies(protein, lig, ff, poses, gpu_device)
    total, total16, mag = np.zeros(n), np.zeros(n), np.zeros(n)
    for a in range(len(lig)):
        pos = np.stack([lig[k][a] + poses[3 + i] for i, k in enumerate("xyz")], 1)
        ijk = np.rint((pos - lo) / h).astype(np.int64)
        v = maps[int(lig["type"][a])][ijk[:, 0], ijk[:, 1], ijk[:, 2]]
        total += v
        total16 += np.clip(v, -65504, 65504).astype(np.float16).astype(np.float64)
        mag += np.abs(v)
    mag = np.maximum(mag, 1.0)
    # the scorer reads the (FP16-stored) map values exactly at grid points ...
    assert np.max(np.abs(got - total16) / mag) < 1e-5
    # ... the sum of the single-atom energies over the ligand is the pose energy ...
    assert np.max(np.abs(exact - total) / mag) < 1e-5
    # ... so on grid points the field is exact up to the FP16 storage rounding
    assert np.max(np.abs(got - exact) / mag) < 2.0 ** -11 + 1e-5


# small training runs of the learned kinds (the mechanics, not the accuracy)
QUICK = {"mlp": dict(train_points=65536, batch=4096, steps=100),
         "pose_mlp": dict(labels_per_pair=2048, batch=1024, epochs=4)}


@pytest.mark.parametrize("surrogate", ["field", "mlp", "pose_mlp"])
def test_selection_rescoring_and_recall(gpu_device, surrogate):
    w = small_workload()
    kw = dict(keep_fraction=0.05, max_val_rmse=1e9, gate_z=0.0)
    kw.update(QUICK.get(surrogate, {}))
    cfg = C.default_config(surrogate, **kw)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        for p, protein in enumerate(w.proteins):
            r = c.screen(protein)
            assert r.fallback_pairs == 0 and r.nposes == sum(q.shape[1] for q in w.poses)
            c.exhaustive()
            rec = c.recall(0.01)
            for l in range(len(w.ligands)):
                idx, e = c.selected(l)
                score = c.scores(l)
                exact = c.exact(l)
                k = kept(0.05, len(score))
                thr = np.sort(score)[k - 1]
                # the selection is the lowest-k of the surrogate's scores (ties kept)
                assert set(idx.tolist()) == set(np.nonzero(score <= thr)[0].tolist()), (p, l)
                assert np.all(np.diff(idx.astype(np.int64)) > 0)
                # rescored energies are the batched kernel's exact energies, bit for bit
                assert np.array_equal(e.view(np.uint32), exact[idx].view(np.uint32)), (p, l)
                assert rec[l] == pytest.approx(S.recall(idx, exact, 0.01), abs=1e-6), (p, l)
            if p == 0:
                # and those equal the plain kernel's for one pair
                want = plain_energies(protein, w.ligands[0], w.forcefield, w.poses[0], gpu_device)
                assert np.array_equal(c.exact(0).view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("surrogate", ["field", "mlp", "pose_mlp"])
def test_sampled_pose_energies(gpu_device, surrogate):
    # the sampled-pose pass runs the split kernel (four lanes per pose): its
    # energies agree with the tiled kernel's to rounding and with the FP64
    # oracle to 16 eps of the cancellation magnitude
    w = C.combined_workload(nproteins=1, nligands=6, nposes=3000, seed=13, natpro_range=(300, 900),
                            natlig_range=(1, 30))
    kw = dict(labels_per_pair=700, max_val_rmse=1e9, gate_z=0.0)
    kw.update({"mlp": dict(train_points=65536, batch=4096, steps=10),
               "pose_mlp": dict(batch=256, epochs=1)}.get(surrogate, {}))
    eps = 2.0 ** -24
    with C.Combined(w.forcefield, w.ligands, w.poses, C.default_config(surrogate, **kw), gpu_device) as c:
        c.screen(w.proteins[0])
        c.exhaustive()
        for l in range(len(w.ligands)):
            idx, e = c.labels(l)
            assert len(idx) == 700 and len(set(idx.tolist())) == 700 and idx.max() < 3000
            sub = D.Deck(w.proteins[0], w.ligands[l], w.forcefield, np.ascontiguousarray(w.poses[l][:, idx]))
            e64, mag = O.fasten_f64_mag(sub, 0, len(idx))
            ok = mag > 0
            assert np.all(e[~ok] == 0.0)
            noise = np.abs(e.astype(np.float64) - e64)[ok] / mag[ok] / eps
            tiled = np.abs(e.astype(np.float64) - c.exact(l)[idx])[ok] / mag[ok] / eps
            print(f"{surrogate} pair {l} ({len(w.ligands[l])} atoms): split vs FP64 {noise.max():.2f} eps, "
                  f"vs tiled {tiled.max():.2f} eps", flush=True)
            assert noise.max() < 16.0 and tiled.max() < 16.0


def test_set_keep_rescreens_with_the_new_fraction(gpu_device):
    w = small_workload(nposes=4096, nligands=3, nproteins=2)
    cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        c.screen(w.proteins[0])
        c.exhaustive()
        r2 = c.recall(0.01)
        for keep in (0.05, 0.01):
            c.set_keep(keep)
            c.screen(w.proteins[0])  # same protein: its exact energies stay usable
            rk = c.recall(0.01)
            for l in range(3):
                idx, _ = c.selected(l)
                assert len(idx) >= kept(keep, 4096)
                score = c.scores(l)
                thr = np.sort(score)[kept(keep, 4096) - 1]
                assert set(idx.tolist()) == set(np.nonzero(score <= thr)[0].tolist())
            if keep > 0.02:
                assert np.all(rk >= r2)
        with pytest.raises(ValueError):
            c.set_keep(0.0)
        c.screen(w.proteins[1])  # another protein: the exact energies are stale
        with pytest.raises(ValueError, match="this protein"):
            c.recall(0.01)


def test_fallback_pairs_are_scored_exhaustively(gpu_device):
    w = small_workload(nposes=4096, nligands=3, nproteins=1)
    cfg = C.default_config("field", max_val_rmse=0.0, keep_fraction=0.02)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        r = c.screen(w.proteins[0])
        assert r.fallback_pairs == 3
        c.exhaustive()
        assert np.all(c.recall(0.01) == 1.0)
        for l in range(3):
            assert c.pair(l).fallback == 1
            exact = c.exact(l)
            assert np.array_equal(c.scores(l).view(np.uint32), exact.view(np.uint32))
            idx, e = c.selected(l)
            assert np.array_equal(e, exact[idx])


def test_mixed_fallback_and_ragged_edges(gpu_device):
    # pose counts of 1, 63, 64, 65 and 3000, a threshold between the pairs'
    # validation errors (some pairs fall back, others are screened), and a
    # keep fraction of 1 on the same ligand set
    w = small_workload(nposes=3000, nligands=5, nproteins=1, seed=14)
    poses = [w.poses[0][:, :1], w.poses[1][:, :63], w.poses[2][:, :64], w.poses[3][:, :65], w.poses[4]]
    cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0, labels_per_pair=48, keep_fraction=0.05)
    with C.Combined(w.forcefield, w.ligands, poses, cfg, gpu_device) as c:
        c.screen(w.proteins[0])
        cal = np.array([c.pair(l).val_rmse_calibrated for l in range(5)])
    thr = float(np.sort(cal)[2])  # the two largest fall back
    for keep in (0.05, 1.0):
        # a fresh object draws the same samples, so the same pairs fall back
        cfg = C.default_config("field", max_val_rmse=thr, gate_z=0.0, labels_per_pair=48, keep_fraction=keep)
        with C.Combined(w.forcefield, w.ligands, poses, cfg, gpu_device) as c:
            r = c.screen(w.proteins[0])
            c.exhaustive()
            fb = [c.pair(l).fallback for l in range(5)]
            assert r.fallback_pairs == sum(fb) and 0 < sum(fb) < 5
            for l in range(5):
                n = poses[l].shape[1]
                exact = c.exact(l)
                want = plain_energies(w.proteins[0], w.ligands[l], w.forcefield, poses[l], gpu_device)
                assert np.array_equal(exact.view(np.uint32), want.view(np.uint32)), l
                idx, e = c.selected(l)
                k = kept(keep, n)
                score = c.scores(l)
                assert set(idx.tolist()) == set(np.nonzero(score <= np.sort(score)[k - 1])[0].tolist()), (keep, l)
                assert np.array_equal(e.view(np.uint32), exact[idx].view(np.uint32)), (keep, l)
                if fb[l]:
                    assert np.array_equal(score.view(np.uint32), exact.view(np.uint32))
                if keep == 1.0:
                    assert len(idx) == n
            assert np.all(c.recall(0.01)[np.array(fb) == 1] == 1.0)


def test_errors_are_reported_not_faked(gpu_device):
    w = small_workload(nposes=256, nligands=2, nproteins=1)
    with C.Combined(w.forcefield, w.ligands, w.poses, C.default_config("field"), gpu_device) as c:
        for call in (lambda: c.pair(0), lambda: c.selected(0), lambda: c.scores(0)):
            with pytest.raises(ValueError, match="no protein screened"):
                call()
        bad = w.proteins[0].copy()
        bad["type"][3] = 999
        with pytest.raises(ValueError, match="protein atom 3"):
            c.screen(bad)
        c.screen(w.proteins[0])
        with pytest.raises(ValueError, match="exhaustive"):
            c.recall(0.01)
        with pytest.raises(ValueError, match="exhaustive"):
            c.exact(0)
        used = set(np.concatenate([l["type"] for l in w.ligands]).tolist())
        unused = next(t for t in range(len(w.forcefield)) if t not in used)
        with pytest.raises(ValueError, match="no ligand atom has this type"):
            c.map(unused)
        with pytest.raises(ValueError):
            c.pair(2)
    cfg = C.default_config("pose_mlp", labels_per_pair=64, batch=64, epochs=1)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        with pytest.raises(ValueError, match="not a field or learned-field screen"):
            c.grid()


def test_field_screen_recall(gpu_device):
    # moderately sized pairs: the field keeps the exact top 1 % in its lowest 2 %
    w = C.combined_workload(nproteins=1, nligands=8, nposes=65536, seed=12)
    cfg = C.default_config("field")
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        r = c.screen(w.proteins[0])
        c.exhaustive()
        rec = c.recall(0.01)
    print(f"field recall@1% (2% kept): mean {rec.mean():.4f} min {rec.min():.4f}; {r.as_dict()}", flush=True)
    assert r.fallback_pairs == 0
    assert rec.min() >= 0.99


def test_pose_mlp_combined_learns(gpu_device):
    # the per-protein pose MLP against the same screen untrained (one step at
    # a vanishing learning rate): training has to buy recall (its recall stays
    # far below what a screen needs: the gate sends such pairs to exhaustive
    # scoring; the learned per-atom field is the learned surrogate that
    # passes, tests/test_gpu_atomnet.py)
    w = C.combined_workload(nproteins=1, nligands=4, nposes=32768, seed=13, natpro_range=(1500, 2500))
    out = {}
    for name, kw in (("trained", dict(epochs=12, lr=4e-3)), ("untrained", dict(epochs=1, lr=1e-30))):
        cfg = C.default_config("pose_mlp", labels_per_pair=8192, batch=1024, max_val_rmse=1e9, gate_z=0.0, **kw)
        with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
            c.screen(w.proteins[0])
            c.exhaustive()
            out[name] = (c.recall(0.01), np.array([c.pair(l).val_rmse_calibrated for l in range(4)]))
        print(f"pose mlp {name}: recall {out[name][0]}, calibrated val_rmse {out[name][1]}", flush=True)
    assert np.all(out["trained"][1] < out["untrained"][1])
    assert out["trained"][0].mean() > max(0.15, 3 * out["untrained"][0].mean())


@pytest.mark.parametrize("keep", [0.05, 1.0])
def test_field_pipeline_chunks_give_the_same_screen(gpu_device, keep):
    # ligand chunks scored, selected and rescored as a pipeline (2, 5 and the
    # clamp of 16 to the 7 ligands) select the same poses as one chunk and
    # return the same exact energies, bit for bit; ragged pose counts put
    # chunk edges next to partial tiles
    w = small_workload(nposes=3000, nligands=7, nproteins=2, seed=15)
    poses = [w.poses[0][:, :1], w.poses[1][:, :63], w.poses[2][:, :65], w.poses[3], w.poses[4][:, :129],
             w.poses[5][:, :64], w.poses[6]]
    runs = {}
    for chunks in (1, 2, 5, 16):
        cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0, labels_per_pair=48, keep_fraction=keep,
                               pipeline_chunks=chunks)
        with C.Combined(w.forcefield, w.ligands, poses, cfg, gpu_device) as c:
            out = []
            for p, protein in enumerate(w.proteins):
                print(f"pipeline_chunks={chunks} keep={keep} protein {p}", flush=True)
                r = c.screen(protein)
                assert r.fallback_pairs == 0
                c.exhaustive()
                sel = [c.selected(l) for l in range(7)]
                for l, (idx, e) in enumerate(sel):
                    exact = c.exact(l)
                    assert np.array_equal(e.view(np.uint32), exact[idx].view(np.uint32)), (chunks, p, l)
                    if keep == 1.0:
                        assert len(idx) == poses[l].shape[1]
                assert r.nselected == sum(len(i) for i, _ in sel)
                out.append(sel)
            runs[chunks] = out
    for chunks in (2, 5, 16):
        for p in range(2):
            for l in range(7):
                assert np.array_equal(runs[chunks][p][l][0], runs[1][p][l][0]), (chunks, p, l)
                assert np.array_equal(runs[chunks][p][l][1].view(np.uint32), runs[1][p][l][1].view(np.uint32))


def test_field_pipeline_with_fallback_pairs(gpu_device):
    # a pair that falls back switches the screen to one selection over the
    # whole set after the chunked scoring; the other pairs are still rescored
    w = small_workload(nposes=2000, nligands=6, nproteins=1, seed=16)
    cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0, labels_per_pair=48, pipeline_chunks=3)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        c.screen(w.proteins[0])
        cal = np.array([c.pair(l).val_rmse_calibrated for l in range(6)])
    cfg = C.default_config("field", max_val_rmse=float(np.sort(cal)[3]), gate_z=0.0, labels_per_pair=48,
                           pipeline_chunks=3)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        r = c.screen(w.proteins[0])
        c.exhaustive()
        fb = [c.pair(l).fallback for l in range(6)]
        assert r.fallback_pairs == sum(fb) == 2
        for l in range(6):
            exact = c.exact(l)
            idx, e = c.selected(l)
            score = c.scores(l)
            k = kept(0.02, 2000)
            assert set(idx.tolist()) == set(np.nonzero(score <= np.sort(score)[k - 1])[0].tolist()), l
            assert np.array_equal(e.view(np.uint32), exact[idx].view(np.uint32)), l
            if fb[l]:
                assert np.array_equal(score.view(np.uint32), exact.view(np.uint32))


@pytest.mark.parametrize("chunks", [1, 3])
def test_tied_scores_are_all_rescored(gpu_device, chunks):
    # every pose repeated three times: tied scores, so a pair keeps more than
    # ceil(keep n) poses (all ties at the threshold); the device-built
    # rescoring tables must cover every kept pose
    w = small_workload(nposes=1200, nligands=4, nproteins=1, seed=17)
    poses = [np.ascontiguousarray(np.repeat(q[:, :400], 3, axis=1)) for q in w.poses]
    cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0, labels_per_pair=48, keep_fraction=0.0085,
                           gate_top_fraction=0.005, pipeline_chunks=chunks)
    with C.Combined(w.forcefield, w.ligands, poses, cfg, gpu_device) as c:
        r = c.screen(w.proteins[0])
        c.exhaustive()
        for l in range(4):
            idx, e = c.selected(l)
            score = c.scores(l)
            k = kept(0.0085, 1200)  # 11: the threshold falls inside a triple
            thr = np.sort(score)[k - 1]
            want = np.nonzero(score <= thr)[0]
            assert len(idx) == len(want) and len(idx) % 3 == 0 and len(idx) > k, (l, len(idx), k)
            assert np.array_equal(idx, want)
            assert np.array_equal(e.view(np.uint32), c.exact(l)[idx].view(np.uint32)), l
        assert r.nselected == sum(len(c.selected(l)[0]) for l in range(4))


def test_automatic_chunking_follows_set_keep(gpu_device):
    # 8 pairs of 262,144 poses: keeping every pose expects 32,768 rescoring
    # tiles, which the automatic choice splits into two pipeline chunks on a
    # 256-CU device (one chunk at 2 %); the chunk tables are rebuilt by
    # set_keep and the screens agree with the exhaustive energies either way
    w = small_workload(nposes=262144, nligands=8, nproteins=1, seed=18)
    cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0, labels_per_pair=64, keep_fraction=1.0)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        for keep in (1.0, 0.02, 1.0):
            c.set_keep(keep)
            print(f"keep {keep}", flush=True)
            r = c.screen(w.proteins[0])
            c.exhaustive()
            total = 0
            for l in range(8):
                idx, e = c.selected(l)
                exact = c.exact(l)
                score = c.scores(l)
                k = kept(keep, 262144)
                assert set(idx.tolist()) == set(np.nonzero(score <= np.sort(score)[k - 1])[0].tolist()), (keep, l)
                assert np.array_equal(e.view(np.uint32), exact[idx].view(np.uint32)), (keep, l)
                total += len(idx)
            assert r.nselected == total


@pytest.mark.parametrize("kind", ["field1", "field3", "fallback", "mlp", "mlp_widened", "pose_mlp"])
def test_batch_equals_one_by_one(gpu_device, kind):
    # screen_batch overlaps each protein's scoring with the previous one's
    # rescoring (or, with fallback pairs or the MLP, runs them one after
    # another): every protein's selection, exact energies, sampled poses and
    # counts equal those of screening the proteins one by one on a fresh
    # object, and the callback sees each protein as the one screened last
    w = small_workload(nposes=3000, nligands=5, nproteins=4, seed=19)
    L = len(w.ligands)
    kw = dict(labels_per_pair=48, gate_z=0.0)  # (48 samples: a gate would fall back at random)
    if kind == "pose_mlp":
        cfg = lambda: C.default_config("pose_mlp", labels_per_pair=256, batch=64, epochs=1, max_val_rmse=1e9,  # noqa: E731
                                       gate_z=0.0)
    elif kind == "mlp":
        cfg = lambda: C.default_config("mlp", labels_per_pair=256, train_points=65536, batch=4096,  # noqa: E731
                                       steps=50, gate_z=0.0)
    elif kind == "mlp_widened":  # pairs widened, none falling back: the proteins still overlap
        cfg = lambda: C.default_config("mlp", labels_per_pair=256, train_points=65536, batch=4096,  # noqa: E731
                                       steps=300, gate_z=2.0, widen_max=1.0)
    else:
        kw.update(max_val_rmse=0.0 if kind == "fallback" else 1e9, pipeline_chunks=3 if kind == "field3" else 1)
        cfg = lambda: C.default_config("field", **kw)  # noqa: E731

    def snapshot(c):
        return ([c.selected(l) for l in range(L)], [c.labels(l) for l in range(L)],
                [c.pair(l).as_dict() for l in range(L)])

    single, nsel = [], []
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg(), gpu_device) as c:
        for p in w.proteins:
            nsel.append(c.screen(p).nselected)
            single.append(snapshot(c))
        last_scores = [c.scores(l) for l in range(L)]
    got, scores_err, exact_ok = {}, [], []
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg(), gpu_device) as c:
        def done(i):
            print(f"{kind}: protein {i} complete", flush=True)
            got[i] = snapshot(c)
            if i == 1:  # exhaustive energies of protein 1 from inside the batch
                c.exhaustive()
                exact_ok.append(all(np.array_equal(c.exact(l)[idx].view(np.uint32), e.view(np.uint32))
                                    for l, (idx, e) in enumerate(got[i][0])))
            if i < 3 and kind in ("field1", "field3", "mlp_widened"):
                try:
                    c.scores(0)
                except capi.BudeError:
                    scores_err.append(i)

        res, wall = c.screen_batch(w.proteins, done)
        assert sorted(got) == [0, 1, 2, 3] and wall > 0
        assert exact_ok == [True]
        if kind == "mlp_widened":
            wide, fb = [r.widened_pairs for r in res], [r.fallback_pairs for r in res]
            print(f"widened {wide}, fallbacks {fb}, scores of the next protein begun at {scores_err}", flush=True)
            assert sum(wide) > 0
            # a protein without fallback pairs lets the next one begin before it ends
            assert scores_err == [i for i in range(3) if fb[i] == 0]
        if kind in ("field1", "field3"):
            assert scores_err == [0, 1, 2]  # the next protein's scoring had begun
        for i in range(4):
            assert res[i].nselected == nsel[i]
            for l in range(L):
                for a, b in zip(got[i][0][l] + got[i][1][l], single[i][0][l] + single[i][1][l]):
                    assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32)), (kind, i, l)
                assert got[i][2][l] == single[i][2][l]
        for l in range(L):  # after the batch the last protein is the one screened last, scores included
            assert np.array_equal(c.scores(l).view(np.uint32), last_scores[l].view(np.uint32))


def test_batch_edges_and_errors(gpu_device):
    # an empty batch, a batch of one, and a batch whose second protein is
    # malformed: the error is reported, nothing stays in flight, and the
    # object screens normally afterwards
    w = small_workload(nposes=2000, nligands=3, nproteins=2, seed=20)
    cfg = C.default_config("field", max_val_rmse=1e9, gate_z=0.0, labels_per_pair=48)
    with C.Combined(w.forcefield, w.ligands, w.poses, cfg, gpu_device) as c:
        assert c.screen_batch([])[0] == []
        seen = []
        res, _ = c.screen_batch(w.proteins[:1], seen.append)
        assert seen == [0] and res[0].nposes == 6000
        one = [c.selected(l) for l in range(3)]
        bad = w.proteins[1].copy()
        bad["type"][0] = 999
        with pytest.raises(capi.BudeInvalidArgument):
            c.screen_batch([w.proteins[0], bad])
        with pytest.raises(capi.BudeError, match="no protein screened"):
            c.selected(0)
        c.screen(w.proteins[0])
        for l in range(3):
            idx, e = c.selected(l)
            assert np.array_equal(idx, one[l][0]) and np.array_equal(e.view(np.uint32), one[l][1].view(np.uint32))
