#!/usr/bin/env python3
"""Benchmark of the plain fasten_main on one MI355X.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--ppwi P] [--wgsize W]

A step is one pass of fasten_main over every pose of the single-GPU
workload deck (4,096 protein atoms, 64 ligand atoms, 34 forcefield types,
1,048,576 poses, seeded synthetic deck "c1"), with the deck already resident
in HBM.  Prints ONE JSON line:

  value         whole-job G-interactions/s over the K timed steps (host wall
                clock of launch + stream synchronize, the reference's own
                kernelTimes bracket, src/hip/fasten.hpp:240-247)
  roofline      FP32 VALU roofline of the kernel: the reference's FLOP
                accounting (src/main.cpp:358-360) per launch / average
                hipEvent duration of the launch on the library's stream, vs
                the MI355X FP32 vector peak (157.3 TFLOP/s); beside it the
                same launch at 30 FLOP per interaction and at the FLOP the
                hardware counters record; `traffic` is the committed
                rocprofv3 --pmc summary's (named in `traffic_source`)
  cpu_baseline  the reference's own omp model (src/main.cpp + src/omp/,
                built from the reference sources into oracle/_ref) on the
                host cores: the CPU workload as configured (deck c0, ppwi 1,
                8 iterations), the same deck at ppwi 64, and the first
                65,536 poses of c1
  valid         energies of the last timed step against the reference
                energies of the deck: the reference metric (max_diff_% <
                0.025) on every pose whose FP32 energy the reference itself
                resolves (|E| >= 0.002 * M, M = 0.5 * sum |term| from the
                committed FP64 fixture), and |E - E_fp64| / M within 16 FP32
                epsilons on every pose (the reference's own FP32: 8.8)
  sweep         the 40-point ppwi x wgsize sweep of the plain kernel
  batched       the ragged screen (256 ligands of 8-96 atoms, 4,096-65,536
                poses each, one 4,096-atom protein) in one persistent
                hipGraph-replayed launch at the launch shape the batched
                autotuner picks on it, vs 256 separate plain launches
  screen        the learned prescreen (4,096-atom protein, 26-atom ligand,
                16.8M fresh poses, lowest 2 % rescored exactly): the pose MLP
                with the 12 configured inputs and with 21, each trained on
                8.4M plain-kernel energies (poses/s and top-1 % recall at
                2 % and at the fraction reaching 0.99, end-to-end time with
                labels and training, break-even pose count), then the
                learned per-atom field and the exact single-atom maps on the
                same pair
  combined      16 proteins x 512 ligands x 1,048,576 poses per pair
                (16 x 64 x 262,144 with --combined-reduced): the exact
                single-atom maps and the learned per-atom field, per-pair
                lowest 2 % rescored in one ragged batched launch per
                protein, all proteins in one batch call, vs exhaustive
                batched scoring (measured on the first protein, extrapolated
                by interactions): poses/s, fallbacks and per-pair top-1 %
                recall of the first protein's 512 pairs
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from minibude_amd import deck as D  # noqa: E402
from minibude_amd import metrics  # noqa: E402
from minibude_amd.model import Context, device_count, device_name  # noqa: E402

METRIC = "G-interactions/s plain vs batched; screened poses/s at \u22650.99 top-1% recall"
WORKLOAD = "c1"
DEFAULT_PPWI = 1
DEFAULT_WGSIZE = 256
# screen operating points: the workload's 2 % first, then larger fractions
# until the exact top-1 % recall reaches 0.99
SCREEN_KEEP = (0.02, 0.03, 0.04, 0.05, 0.06, 0.08)
# combined screen sizes: the workload's own (default), and a reduced one
COMBINED_DEFAULT = dict(nproteins=16, nligands=64, nposes=1 << 18)
COMBINED_FULL = dict(nproteins=16, nligands=512, nposes=1 << 20)
TRAFFIC_SUMMARY = ROOT / "profiles" / "fasten_plain_c1_traffic.json"
CPU_SAMPLE_POSES = 65536
CPU_ITERS = 2


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host() -> dict:
    info = Path("/proc/cpuinfo").read_text() if Path("/proc/cpuinfo").exists() else ""
    cpu = re.search(r"^model name\s*:\s*(.+)$", info, re.M)
    return {"v4": " avx512f " in info.replace("\n", " "), "model": cpu.group(1).strip() if cpu else "unknown CPU",
            "bind": os.environ.get("OMP_PROC_BIND", "unset"),
            "cores": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1), "nproc": os.cpu_count()}


def run_reference_omp(exe: Path, deck: D.Deck, ppwi: int, iters: int, cores: int) -> dict | None:
    """The reference driver's own omp model on a deck written to a temporary
    directory: per-iteration wall times from its raw_iterations (2 warmups
    first, src/main.cpp:160), its validation outcome, G-interactions/s of the
    timed iterations."""
    n = deck.nposes
    with tempfile.TemporaryDirectory() as tmp:
        D.write_deck(tmp, deck)
        cmd = [str(exe), "--deck", tmp, "-n", str(n), "-i", str(iters), "-p", str(ppwi)]
        log("cpu baseline: " + " ".join([exe.name] + cmd[3:]))
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
        wall = time.time() - t0
    if r.returncode not in (0, 1):
        log(f"cpu baseline failed ({r.returncode}): {r.stderr[-400:]}")
        return None
    raw = re.search(r"raw_iterations:\s*\[([^\]]*)\]", r.stdout)
    outcome = re.search(r"valid: (\w+), max_diff_%: ([0-9.]+)", r.stdout)
    if not raw:
        log(f"cpu baseline: no raw_iterations in output: {r.stdout[-400:]}")
        return None
    its = [float(x) for x in raw.group(1).split(",") if x.strip()]
    timed = its[2:] if len(its) > 2 else its
    ms = sum(timed) / len(timed)
    ips = n * deck.natlig * deck.natpro / (ms * 1e-3)
    out = {"ppwi": ppwi, "iterations": len(timed), "ms_per_iteration": round(ms, 3),
           "g_interactions_per_s": round(ips / 1e9, 4), "wall_s": round(wall, 1),
           "valid": outcome.group(1) == "true" if outcome else None,
           "max_diff_pct": float(outcome.group(2)) if outcome else None}
    log(f"cpu baseline: {out}")
    return out


def cpu_baseline(c1: D.Deck, c1_ref: np.ndarray) -> dict | None:
    """The reference's own omp model (src/main.cpp + src/omp/fasten.hpp, built
    from the reference sources into oracle/_ref) on the host cores: the
    CPU workload as configured (bm1-shaped synthetic deck c0: 938 protein /
    26 ligand atoms, 65,536 poses, seed 0; ppwi 1, 8 timed iterations after
    the driver's 2 warmups), the same deck at its best ppwi (64), and the
    first 65,536 poses of the GPU deck c1 at ppwi 64."""
    h = _host()
    exe = ROOT / "oracle" / "_ref" / ("omp-bude-v4" if h["v4"] else "omp-bude-v3")
    if not exe.exists():
        log(f"cpu baseline skipped: {exe} not built")
        return None
    c0 = D.workload_deck("c0")
    fx = ROOT / "tests" / "golden" / "c0" / "ref_energies.f32"
    c0.ref_energies = np.fromfile(fx, "<f4") if fx.exists() else None
    if c0.ref_energies is None:
        log("cpu baseline skipped: no c0 reference energies")
        return None
    conf = run_reference_omp(exe, c0, 1, 8, h["cores"])
    best = run_reference_omp(exe, c0, 64, 8, h["cores"])
    n = min(CPU_SAMPLE_POSES, c1.nposes)
    sub = c1.head(n)
    sub.ref_energies = c1_ref[:n]
    c1s = run_reference_omp(exe, sub, 64, CPU_ITERS, h["cores"])
    if conf is None:
        return None
    return {
        "value": conf["g_interactions_per_s"],
        "unit": "G-interactions/s",
        "cores": h["cores"],
        "kind": "reference",
        "sample": (f"c0: bm1-shaped synthetic deck (938 protein / 26 ligand atoms, 34 types, 65,536 poses, seed 0), "
                   f"the reference's omp model {exe.name} (src/main.cpp + src/omp/fasten.hpp, -O3 -ffast-math "
                   f"-fopenmp) at ppwi 1 and wgsize 1, 2 warmup + 8 timed iterations timed by its own "
                   f"raw_iterations ({conf['wall_s']} s wall), valid={conf['valid']} "
                   f"(max_diff_% {conf['max_diff_pct']}); {h['cores']} threads of {h['nproc']} on {h['model']}, "
                   f"OMP_PROC_BIND={h['bind']}"),
        "configured": conf,
        "best_ppwi": best,
        "c1_sample": None if c1s is None else dict(c1s, deck=f"first {n} poses of the c1 deck (4,096 / 64 atoms)"),
        "host": {"nproc": h["nproc"], "model": h["model"], "omp_num_threads": h["cores"], "omp_proc_bind": h["bind"]},
    }


def validation(e: np.ndarray, ref: np.ndarray) -> dict:
    fx = ROOT / "tests" / "golden" / WORKLOAD / "fp64.f32"
    max_diff, failed = metrics.validate(ref, e)
    out = {"reference_metric_all_poses": {"max_diff_pct": round(max_diff, 5), "failed": int(failed.size),
                                          "poses": int(e.size)}}
    if not fx.exists():
        out["valid"] = metrics.is_valid(max_diff)
        return out
    em = np.fromfile(fx, "<f4").reshape(-1, 2).astype(np.float64)
    e64, mag = em[:, 0], em[:, 1]
    cond = np.abs(e64) >= 0.002 * mag
    md_c, f_c = metrics.validate(ref[cond], e[cond])
    eps = 2.0 ** -24
    noise = float(np.max(np.abs(e - e64) / mag) / eps)
    ref_noise = float(np.max(np.abs(ref - e64) / mag) / eps)
    md_ref64, f_ref64 = metrics.validate(e64.astype(np.float32), ref)
    out.update({
        "reference_metric_resolvable_poses": {"max_diff_pct": round(md_c, 5), "failed": int(f_c.size),
                                              "poses": int(cond.sum())},
        "max_err_over_magnitude_eps": round(noise, 2),
        "reference_fp32_max_err_over_magnitude_eps": round(ref_noise, 2),
        "reference_fp32_vs_fp64_reference_metric": {"max_diff_pct": round(md_ref64, 5), "failed": int(f_ref64.size)},
        "fp64_max_rel_pct_resolvable": round(metrics.symmetric_max_rel_pct(e64[cond], e[cond]), 5),
    })
    out["valid"] = bool(metrics.is_valid(md_c) and noise < 16.0)
    return out


SWEEP_PPWI = (1, 2, 4, 8, 16, 32, 64, 128)
SWEEP_WGSIZE = (64, 128, 256, 512, 1024)


def sweep(ctx: Context, ref: np.ndarray | None, warmup: int, steps: int) -> dict:
    """The single-GPU workload's autotune sweep (the reference driver's
    ppwi x wgsize report, src/main.cpp:517-565): hipEvent ms per launch and
    validity of every point, the best by summed time as the driver picks it."""
    inter = None
    rows = []
    for p in SWEEP_PPWI:
        for wg in SWEEP_WGSIZE:
            s = ctx.fasten(p, wg, warmup=warmup, iterations=steps, want_energies=ref is not None)
            dev = float(np.mean(s.device_ms[warmup:]))
            inter = inter or metrics.interactions_per_pass(ctx.deck.nposes, ctx.deck.natlig, ctx.deck.natpro)
            v = validation(s.energies, ref)["valid"] if ref is not None else None
            rows.append({"ppwi": p, "wgsize": wg, "ms": round(dev, 3),
                         "g_interactions_per_s": round(inter / dev / 1e6, 1),
                         "tflops": round(metrics.flops_per_pass(ctx.deck.nposes, ctx.deck.natlig, ctx.deck.natpro, p)
                                         / (dev * 1e-3) / 1e12, 2), "valid": v})
            log(f"sweep: {rows[-1]}")
    best = min(rows, key=lambda r: r["ms"])
    return {"points": rows, "best": best}


def batched() -> dict:
    from minibude_amd.batch import Batch, ragged_workload
    ls = ragged_workload()
    log(f"batched: {len(ls.ligands)} ligands, {ls.total_poses} poses, {ls.interactions:.3e} interactions")
    with Batch(ls, 0) as b:
        # the launch shape is the project's autotuner's choice on this batch
        t0 = time.time()
        cfg, tune_ms = b.autotune()
        tune_s = time.time() - t0
        log(f"batched: autotuned in {tune_s:.1f} s: {cfg.as_dict()} ({tune_ms:.2f} ms)")
        ms = b.run(cfg, warmup=1, iterations=3)[1:]
        sep = b.run_separate(1, 256, warmup=1, iterations=1)[1:]
    t, ts = float(np.median(ms)), float(np.median(sep))
    out = {
        "workload": "c2: 256 ligands (8-96 atoms, 4,096-65,536 poses each) vs one 4,096-atom protein",
        "interactions": ls.interactions,
        "ms": round(t, 3),
        "g_interactions_per_s": round(ls.interactions / t / 1e6, 2),
        "separate_launches_ms": round(ts, 3),
        "separate_g_interactions_per_s": round(ls.interactions / ts / 1e6, 2),
        "speedup_vs_separate": round(ts / t, 2),
        "config": cfg.as_dict(),
        "autotune": {"seconds": round(tune_s, 2), "best_ms": round(tune_ms, 3),
                     "search": ("bude_mi355x_batch_autotune on this batch: 24 candidates (tile width 64/128/256 "
                                "poses, work-group 64/256, 4/8 waves per SIMD, heaviest-first or ligand order; "
                                "hipGraph replays), each timed once, the 4 fastest by the median of 3")},
    }
    log(f"batched: {out}")
    return out


def breakeven(setup_ms: float, exhaustive_ms: float, screen_ms: float, n: int) -> float | None:
    """Poses beyond which a surrogate screen (setup_ms of labels and training,
    then screen_ms per n poses) beats exhaustive scoring (exhaustive_ms per n)."""
    per = (exhaustive_ms - screen_ms) / n
    return round(setup_ms / per) if per > 0 else None


def screen(epochs: int = 0) -> dict:
    """The learned-prescreen workload (4,096-atom protein, 26-atom ligand,
    16.8M fresh poses): the pose MLP as configured (12 pose inputs) and with
    the 9 rotation-matrix entries added (21), each trained on 8.4M plain-
    kernel energies; the learned per-atom field and the exact single-atom
    maps through the combined screen on the same pair."""
    from minibude_amd import combined as C
    from minibude_amd import screen as S
    w = S.screen_workload()
    n = w.screen_poses.shape[1]
    exact, ex_ms = None, None
    nets = {}
    for feats in (12, 21):
        cfg = S.default_config(features=feats, **({"epochs": epochs} if epochs else {}))
        log(f"screen: {feats}-input MLP, training on {w.train_poses.shape[1]} poses, screening {n}")
        points = []
        with S.Screen(w.protein, w.ligand, w.forcefield, cfg, 0) as sc:
            st, _ = sc.train(w.train_poses)
            sc.load(w.screen_poses)
            if exact is None:
                exact, ex_ms = sc.exhaustive()
            # the workload's operating point (2 % rescored), then the smallest
            # fraction whose recall of the exact top 1 % reaches 0.99
            for keep in SCREEN_KEEP:
                sc.set_keep(keep)
                r = min((sc.run() for _ in range(3)), key=lambda x: x.total_ms)
                idx, _ = sc.selected()
                rec = S.recall(idx, exact, 0.01)
                points.append({
                    "keep_fraction": keep,
                    "poses_per_s": round(n / (r.total_ms * 1e-3), 1),
                    "screen_ms": round(r.total_ms, 3),
                    "surrogate_ms": round(r.score_ms, 3),
                    "select_ms": round(r.select_ms, 3),
                    "rescore_ms": round(r.rescore_ms, 3),
                    "speedup_vs_exhaustive": round(ex_ms / r.total_ms, 2),
                    "recall_top1pct": round(rec, 4),
                    "recall_top0.1pct": round(S.recall(idx, exact, 0.001), 4),
                    "fallback": bool(r.fallback),
                })
                log(f"screen ({feats} inputs): {points[-1]}")
                if keep != SCREEN_KEEP[0] and rec >= 0.99:
                    break
        setup = st.label_ms + st.train_ms
        nets[feats] = {
            "inputs": feats,
            "val_rmse_standardised": round(st.val_rmse, 4),
            "val_rmse_energy_units": round(st.val_rmse * st.sigma, 3),
            "target": f"min(E, {st.clip:.2f}) standardised by mu {st.mu:.2f}, sigma {st.sigma:.2f}",
            "epochs": int(cfg.epochs),
            "labels_and_training_s": round(setup * 1e-3, 2),
            "points": points,
            "end_to_end_ms_at_2pct": round(setup + points[0]["screen_ms"], 1),
            "breakeven_poses_at_2pct": breakeven(setup, ex_ms, points[0]["screen_ms"], n),
        }
    # the learned per-atom field and the exact maps on the same pair (the
    # combined screen with one protein and one ligand); every screen trains
    # (learned field) or builds the maps (field) for its protein
    others = {}
    for kind in ("mlp", "field"):
        with C.Combined(w.forcefield, [w.ligand], [w.screen_poses], C.default_config(kind), 0) as c:
            runs = [c.screen(w.protein) for _ in range(3)]
            r = min(runs, key=lambda x: x.total_ms)
            idx, _ = c.selected(0)
            pair = c.pair(0)
        setup = r.label_ms + r.train_ms
        scr = r.score_ms + r.select_ms + r.rescore_ms
        others[kind] = {
            "keep_fraction": 0.02,
            "screen_ms_excluding_setup": round(scr, 3),
            "poses_per_s_excluding_setup": round(n / (scr * 1e-3), 1),
            "speedup_vs_exhaustive_excluding_setup": round(ex_ms / scr, 2),
            "end_to_end_ms": round(r.total_ms, 3),
            "speedup_vs_exhaustive_end_to_end": round(ex_ms / r.total_ms, 2),
            "breakeven_poses": breakeven(setup, ex_ms, scr, n),
            "recall_top1pct": round(S.recall(idx, exact, 0.01), 4),
            "recall_top0.1pct": round(S.recall(idx, exact, 0.001), 4),
            "fallback": bool(pair.fallback),
            "gate_z": round(pair.gate_z, 2),
            "pair_keep_fraction": round(pair.keep_fraction, 5),
            "one_atom_rmse_energy_units": round(r.surrogate_rmse, 3) if kind == "mlp" else None,
            "phases_ms": {k: round(getattr(r, k), 3) for k in ("label_ms", "train_ms", "score_ms", "select_ms",
                                                               "rescore_ms")},
        }
        log(f"screen ({kind} surrogate): {others[kind]}")
    return {
        "workload": (f"c3: 4,096-atom protein, 26-atom ligand; pose MLP trained on {w.train_poses.shape[1]} plain-"
                     f"kernel energies, {n} fresh poses screened, lowest keep_fraction rescored exactly"),
        "exhaustive_ms": round(ex_ms, 3),
        "exhaustive_poses_per_s": round(n / (ex_ms * 1e-3), 1),
        "pose_mlp_configured_12_inputs": nets[12],
        "pose_mlp_21_inputs": nets[21],
        "learned_field": others["mlp"],
        "field_surrogate": others["field"],
    }


def copy_calibration(nbytes: int) -> None:
    """One device-to-device copy of nbytes (the runtime's copy kernel), so the
    FETCH_SIZE / WRITE_SIZE counters can be scaled by a kernel whose HBM
    traffic is known: nbytes read + nbytes written."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # D2D
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    log(f"copied {nbytes} bytes device to device")


def counter_summary() -> dict:
    """The committed rocprofv3 --pmc summary of the plain kernel on c1."""
    if TRAFFIC_SUMMARY.exists():
        try:
            return json.loads(TRAFFIC_SUMMARY.read_text())
        except ValueError:
            return {}
    return {}


def combined(reduced: bool, kinds=("field", "mlp"), nproteins: int = 0, recall_proteins: int = 1,
             steps: int = 0) -> dict:
    """The combined screen with each surrogate on the same workload: every
    protein screened in one batch call (timed), then exact energies of every
    pair of the first `recall_proteins` proteins (one exhaustive batched pass
    each, measured once and shared by the surrogates) for per-pair recall; the
    exhaustive time of the rest is extrapolated from those proteins'
    interaction rate."""
    from minibude_amd import combined as C
    from minibude_amd import screen as S
    shape = dict(COMBINED_DEFAULT if reduced else COMBINED_FULL)
    if nproteins:
        shape["nproteins"] = nproteins
    t0 = time.time()
    w = C.combined_workload(**shape)
    L = len(w.ligands)
    per_protein = [len(p) * sum(len(l) * q.shape[1] for l, q in zip(w.ligands, w.poses)) for p in w.proteins]
    log(f"combined: {len(w.proteins)} proteins x {L} ligands x {shape['nposes']} poses "
        f"({w.total_poses} pair-poses, {w.interactions:.3e} interactions), generated in {time.time() - t0:.1f} s")
    full = not reduced and not nproteins
    out = {
        "workload": (f"c4: {len(w.proteins)} proteins of 900-6000 atoms x {L} ligands of 8-96 atoms x "
                     f"{shape['nposes']} poses per pair" +
                     ("" if full else " (reduced from 16 proteins x 512 ligands x 2^20 poses)")),
        "pairs": w.pairs,
        "pair_poses": w.total_poses,
        "interactions": w.interactions,
    }
    exact: dict[int, list[np.ndarray]] = {}
    ex_ms: dict[int, float] = {}
    nrec = min(recall_proteins, len(w.proteins))
    for kind in kinds:
        cfg = C.default_config(kind, **({"steps": steps} if steps and kind == "mlp" else {}))
        parts = {"label_ms": 0.0, "train_ms": 0.0, "score_ms": 0.0, "select_ms": 0.0, "rescore_ms": 0.0}
        fb, wide, labels, sel, rmse = 0, 0, 0, 0, []
        rec, trusted, gate_z, keeps, one_ms = [], [], [], [], 0.0
        with C.Combined(w.forcefield, w.ligands, w.poses, cfg, 0) as c:
            # one untimed screen of the first protein sizes the object's
            # buffers, then every protein in one batch call (each one's
            # scoring beside the previous one's rescoring)
            c.screen(w.proteins[0])
            results, screen_ms = c.screen_batch(w.proteins)
            for r in results:
                fb += r.fallback_pairs
                wide += r.widened_pairs
                labels += r.labels
                sel += r.nselected
                rmse.append(r.surrogate_rmse)
                for k in parts:
                    parts[k] += getattr(r, k)
            # recall of every pair of the first proteins (a protein draws the
            # same samples and trains the same network alone or in a batch,
            # so its selection equals the batch's)
            for p in range(nrec):
                one_ms += c.screen(w.proteins[p]).total_ms
                if p not in exact:
                    ex_ms[p] = c.exhaustive()
                    exact[p] = [c.exact(l) for l in range(L)]
                for l in range(L):
                    pr = c.pair(l)
                    r_l = S.recall(c.selected(l)[0], exact[p][l], 0.01)
                    rec.append(r_l)
                    gate_z.append(pr.gate_z)
                    keeps.append(pr.keep_fraction)
                    if not pr.fallback:
                        trusted.append(r_l)
        rec = np.array(rec)
        ex_inter = sum(per_protein[p] for p in range(nrec))
        ex_total_ms = sum(ex_ms[p] for p in range(nrec)) * w.interactions / ex_inter
        res = {
            "poses_per_s": round(w.total_poses / (screen_ms * 1e-3), 1),
            "screen_ms": round(screen_ms, 2),
            "screen_ms_one_by_one_recall_proteins": round(one_ms, 2),
            "exhaustive_ms": round(ex_total_ms, 2),
            "exhaustive_poses_per_s": round(w.total_poses / (ex_total_ms * 1e-3), 1),
            "speedup_vs_exhaustive": round(ex_total_ms / screen_ms, 2),
            "recall_top1pct_mean": round(float(rec.mean()), 4),
            "recall_top1pct_min": round(float(rec.min()), 4),
            "recall_top1pct_p05": round(float(np.percentile(rec, 5)), 4),
            "trusted_recall_top1pct_min": round(float(min(trusted)), 4) if trusted else None,
            "pairs_recall_ge_0.99": round(float((rec >= 0.99).mean()), 4),
            "recall_pairs_measured": int(len(rec)),
            "fallback_pairs": fb,
            "fallback_fraction": round(fb / w.pairs, 4),
            "widened_pairs": wide,
            "widened_fraction": round(wide / w.pairs, 4),
            "keep_fraction_mean_measured": round(float(np.mean(keeps)), 5),
            "gate_z_min_measured": round(float(np.min(gate_z)), 2),
            "labels": labels,
            "rescored": sel,
            "phases_ms": {k: round(v, 2) for k, v in parts.items()},
            "config": {k: (round(v, 6) if isinstance(v, float) else v) for k, v in cfg.as_dict().items()},
            "exhaustive_note": (f"exhaustive batched scoring measured on the first {nrec} protein(s) "
                                f"({ex_inter:.3e} interactions), extrapolated to all by interactions"),
        }
        if kind == "mlp":
            res["surrogate"] = (f"learned per-atom field: per protein a learned {cfg.grid_spacing:g} A grid of 8 "
                                "features, trilinear, then a 2 x 32 ReLU MLP with the atom type one-hot (BF16 "
                                "MFMA), summed over the ligand's atoms; trained by this library's kernels on "
                                f"{cfg.train_points} exact one-atom energies (the plain kernel's pair loop) for "
                                f"{cfg.steps} Adam steps of {cfg.batch}")
            res["one_atom_rmse_energy_units"] = [round(x, 3) for x in rmse]
        log(f"combined {kind}: {res}")
        out[kind] = res
    return out


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ppwi", type=int, default=DEFAULT_PPWI)
    ap.add_argument("--wgsize", type=int, default=DEFAULT_WGSIZE)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-screen", action="store_true")
    ap.add_argument("--no-combined", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the ppwi x wgsize autotune sweep of the plain kernel (40 points)")
    ap.add_argument("--screen-epochs", type=int, default=0, help="training epochs of the learned prescreen "
                    "(default: its configuration's 60; fewer only to keep a kernel trace short)")
    ap.add_argument("--combined-reduced", action="store_true",
                    help="combined screen at 16 proteins x 64 ligands x 2^18 poses instead of the full 16 x 512 x 2^20")
    ap.add_argument("--combined-surrogates", default="field,mlp", help="comma list of field, mlp, pose_mlp")
    ap.add_argument("--combined-recall-proteins", type=int, default=1,
                    help="proteins whose pairs get exhaustive energies for recall")
    ap.add_argument("--combined-proteins", type=int, default=0, help="screen only the first N proteins")
    ap.add_argument("--combined-steps", type=int, default=0, help="training steps of the learned per-atom field "
                    "(default: its configuration's 3,000; fewer only to keep a counter run short)")
    ap.add_argument("--copy-calibration", type=int, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration) and exit")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return 0

    if a.gpus != 1:
        print(f"bench.py measures one GPU only: the plain kernel is data-parallel over poses and the library runs "
              f"on a single MI355X (--gpus {a.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        return 2
    if device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    deck = D.workload_deck(WORKLOAD)
    fx = ROOT / "tests" / "golden" / WORKLOAD / "ref_energies.f32"
    ref = np.fromfile(fx, "<f4") if fx.exists() else None
    log(f"deck {WORKLOAD}: {deck.natpro} protein, {deck.natlig} ligand, {deck.nposes} poses; "
        f"device {device_name(0)}")

    with Context(deck, 0) as ctx:
        log(f"uploaded in {ctx.context_ms:.1f} ms; {a.warmup} warmup + {a.steps} timed steps "
            f"(ppwi={a.ppwi}, wgsize={a.wgsize})")
        s = ctx.fasten(a.ppwi, a.wgsize, warmup=a.warmup, iterations=a.steps, want_energies=True)
        sweep_out = None if a.no_sweep else sweep(ctx, ref, 1, 1)
    host = np.array(s.kernel_ms[a.warmup:])
    dev = np.array(s.device_ms[a.warmup:])
    ms_per_step = float(host.mean())
    inter = metrics.interactions_per_pass(deck.nposes, deck.natlig, deck.natpro)
    value = inter / (ms_per_step * 1e-3) / 1e9
    flops = metrics.flops_per_pass(deck.nposes, deck.natlig, deck.natpro, a.ppwi)
    achieved = flops / (float(dev.mean()) * 1e-3) / 1e12
    peak = metrics.FP32_VECTOR_PEAK_TFLOPS
    # the same launch under two other accountings: without the per-pair
    # parameter derivation (10/PPWI) and ligand-transform share (2/PPWI)
    # that this kernel does once per (ligand atom, protein type) on the host,
    # and the FLOP the hardware counted (committed counter summary)
    flops30 = float(deck.nposes) * (28 + deck.natlig * 18 + deck.natlig * deck.natpro * 30)
    cs = counter_summary()
    counted = float(cs.get("counted_flop_per_interaction", 0.0)) * inter
    log(f"{ms_per_step:.3f} ms/step host, {dev.mean():.3f} ms/launch device, {value:.1f} G-interactions/s, "
        f"{achieved:.2f} TFLOP/s ({achieved / peak:.3f} of FP32 vector peak)")

    parity = None
    valid = None
    if ref is not None:
        parity = validation(s.energies, ref)
        valid = parity["valid"]
        log(f"validation: {parity}")

    cpu = None if a.no_cpu_baseline or ref is None else cpu_baseline(deck, ref)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "G-interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded deck, regenerated byte-identically; reference energies committed)",
        "config": {
            "workload": (f"{WORKLOAD}: plain fasten_main, {deck.natpro} protein atoms, {deck.natlig} ligand atoms, "
                         f"{deck.ntypes} forcefield types, {deck.nposes} poses"),
            "ppwi": a.ppwi,
            "wgsize": a.wgsize,
            "interactions_per_step": inter,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "traffic": cs.get("hbm_bytes_per_launch"),
            "traffic_source": ("committed rocprofv3 --pmc summary profiles/fasten_plain_c1_traffic.json (FETCH_SIZE "
                               "x 2.0 + WRITE_SIZE x 1.0, factors from a 512 MiB device copy), not measured in "
                               "this run"),
            "kernel_ms": round(float(dev.mean()), 4),
            "flop_accounting": "src/main.cpp:358-360 (30 + 10/PPWI per interaction)",
            "frac_without_pair_setup": round(flops30 / (float(dev.mean()) * 1e-3) / 1e12 / peak, 4),
            "frac_counted": round(counted / (float(dev.mean()) * 1e-3) / 1e12 / peak, 4) if counted else None,
            "frac_counted_of_unpacked_rate": (round(counted / (float(dev.mean()) * 1e-3) / 1e12 / (peak / 2), 4)
                                              if counted else None),
            "accountings": ("frac: the reference's 30 + 10/PPWI FLOP per interaction; frac_without_pair_setup: "
                            "30 per interaction (no per-pair parameter derivation, which this kernel does per "
                            "ligand atom and protein type on the host); frac_counted: the "
                            f"{cs.get('counted_flop_per_interaction')} FLOP per interaction the hardware counters "
                            "record (committed summary), against the packed-FP32 peak and against half of it "
                            "(packed FP32 is not used)"),
        },
        "cpu_baseline": cpu,
        "valid": valid,
        "parity": parity,
        "sweep": sweep_out,
        "batched": None if a.no_batched else batched(),
        "screen": None if a.no_screen else screen(a.screen_epochs),
        "combined": None if a.no_combined else combined(a.combined_reduced, tuple(a.combined_surrogates.split(",")),
                                                         a.combined_proteins, a.combined_recall_proteins,
                                                         a.combined_steps),
        "device": device_name(0),
    }
    print(json.dumps(out), flush=True)
    return 0 if valid in (None, True) else 1


if __name__ == "__main__":
    sys.exit(main())
