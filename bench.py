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
                the MI355X FP32 vector peak (157.3 TFLOP/s); `traffic` comes
                from the committed rocprofv3 --pmc summary when present
  cpu_baseline  the reference's own omp model (src/main.cpp + src/omp/,
                built from the reference sources into oracle/_ref) on the
                host cores, on the first 65,536 poses of the same deck
  valid         energies of the last timed step against the reference
                energies of the deck: the reference metric (max_diff_% <
                0.025) on every pose whose FP32 energy the reference itself
                resolves (|E| >= 0.002 * M, M = 0.5 * sum |term| from the
                committed FP64 fixture), and |E - E_fp64| / M within 16 FP32
                epsilons on every pose (the reference's own FP32: 8.8)
  batched       the ragged screen (256 ligands of 8-96 atoms, 4,096-65,536
                poses each, one 4,096-atom protein) in one persistent
                hipGraph-replayed launch vs 256 separate plain launches
  screen        the learned prescreen (surrogate trained on 8M plain-kernel
                energies, 16.8M fresh poses screened, lowest 2 % rescored
                exactly) vs exhaustive plain scoring: poses/s and top-1 %
                recall, at 2 % rescored and at the fraction reaching 0.99
  combined      16 proteins x 64 ligands x 262,144 poses per pair (512 x
                1,048,576 with --combined-full): both surrogates (per-protein
                MLP, single-atom energy maps), per-pair lowest 2 % rescored in
                one ragged batched launch per protein, vs exhaustive batched
                scoring of every pair: poses/s and per-pair top-1 % recall
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
# persistent ragged-batch configuration chosen by bude_mi355x_batch_autotune on
# the 256-ligand workload: ppwi, wgsize, waves per SIMD, heaviest-first, graph
BATCH_CONFIG = (1, 64, 8, 0, 1)
# screen operating points: the workload's 2 % first, then larger fractions
# until the exact top-1 % recall reaches 0.99
SCREEN_KEEP = (0.02, 0.03, 0.04, 0.05, 0.06, 0.08)
# combined screen sizes: the default keeps the run to seconds with recall on
# every pair; the full size is the workload's own
COMBINED_DEFAULT = dict(nproteins=16, nligands=64, nposes=1 << 18)
COMBINED_FULL = dict(nproteins=16, nligands=512, nposes=1 << 20)
COMBINED_FIELD_KEEPS = (0.01, 0.0125, 0.015)
TRAFFIC_SUMMARY = ROOT / "profiles" / "fasten_plain_c1_traffic.json"
CPU_SAMPLE_POSES = 65536
CPU_ITERS = 2


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(deck: D.Deck, ref: np.ndarray) -> dict | None:
    """The reference omp model on the first CPU_SAMPLE_POSES poses of the deck."""
    flags = Path("/proc/cpuinfo").read_text() if Path("/proc/cpuinfo").exists() else ""
    v4 = " avx512f " in flags.replace("\n", " ")
    cpu = re.search(r"^model name\s*:\s*(.+)$", flags, re.M)
    cpu = cpu.group(1).strip() if cpu else "unknown CPU"
    bind = os.environ.get("OMP_PROC_BIND", "unset")
    exe = ROOT / "oracle" / "_ref" / ("omp-bude-v4" if v4 else "omp-bude-v3")
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    if not exe.exists():
        log(f"cpu baseline skipped: {exe} not built")
        return None
    n = min(CPU_SAMPLE_POSES, deck.nposes)
    sub = deck.head(n)
    sub.ref_energies = ref[:n]
    with tempfile.TemporaryDirectory() as tmp:
        D.write_deck(tmp, sub)
        cmd = [str(exe), "--deck", tmp, "-n", str(n), "-i", str(CPU_ITERS), "-p", "64"]
        log("cpu baseline: " + " ".join(cmd[:1] + cmd[2:2] + cmd[3:]))
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
        wall = time.time() - t0
    if r.returncode not in (0, 1):
        log(f"cpu baseline failed ({r.returncode}): {r.stderr[-400:]}")
        return None
    # time from the reference's own per-iteration wall clock (raw_iterations,
    # warmups first): its CSV reports zero rates for a result that fails its
    # validation, which the reference's FP32 does on this ill-conditioned deck
    raw = re.search(r"raw_iterations:\s*\[([^\]]*)\]", r.stdout)
    outcome = re.search(r"valid: (\w+), max_diff_%: ([0-9.]+)", r.stdout)
    if not raw:
        log(f"cpu baseline: no raw_iterations in output: {r.stdout[-400:]}")
        return None
    its = [float(x) for x in raw.group(1).split(",") if x.strip()]
    timed = its[2:] if len(its) > 2 else its
    ms = sum(timed) / len(timed)
    ips = n * deck.natlig * deck.natpro / (ms * 1e-3)
    valid = outcome.group(1) if outcome else "?"
    log(f"cpu baseline: raw_iterations {its} ms -> {ips / 1e9:.2f} G-interactions/s (reference outcome "
        f"{outcome.group(0) if outcome else '?'})")
    return {
        "value": round(ips / 1e9, 4),
        "unit": "G-interactions/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"first {n} of {deck.nposes} poses of the {WORKLOAD} deck ({deck.natpro} protein / {deck.natlig} "
                   f"ligand atoms); reference omp model {exe.name} (src/main.cpp + src/omp/fasten.hpp, -O3 "
                   f"-ffast-math -fopenmp), ppwi=64, 2 warmup + {CPU_ITERS} timed passes timed by its own "
                   f"raw_iterations, {wall:.1f}s wall ({cores} threads of {os.cpu_count()} on {cpu}, "
                   f"OMP_PROC_BIND={bind}); its validation against the strict reference kernel: valid={valid}"),
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
    from minibude_amd.capi import BatchConfig
    ls = ragged_workload()
    log(f"batched: {len(ls.ligands)} ligands, {ls.total_poses} poses, {ls.interactions:.3e} interactions")
    cfg = BatchConfig(*BATCH_CONFIG)
    with Batch(ls, 0) as b:
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
    }
    log(f"batched: {out}")
    return out


def screen(epochs: int = 0) -> dict:
    from minibude_amd import screen as S
    w = S.screen_workload()
    cfg = S.default_config(**({"epochs": epochs} if epochs else {}))
    n = w.screen_poses.shape[1]
    log(f"screen: training on {w.train_poses.shape[1]} poses, screening {n}")
    points = []
    with S.Screen(w.protein, w.ligand, w.forcefield, cfg, 0) as sc:
        st, _ = sc.train(w.train_poses)
        sc.load(w.screen_poses)
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
            log(f"screen: {points[-1]}")
            if keep != SCREEN_KEEP[0] and rec >= 0.99:
                break
    # the same workload through the combined screen's field surrogate (one pair)
    from minibude_amd import combined as C
    with C.Combined(w.forcefield, [w.ligand], [w.screen_poses], C.default_config("field"), 0) as c:
        runs = [c.screen(w.protein) for _ in range(3)]
        r = min(runs, key=lambda x: x.total_ms)
        idx, _ = c.selected(0)
    field = {
        "keep_fraction": 0.02,
        "poses_per_s": round(n / (r.total_ms * 1e-3), 1),
        "screen_ms": round(r.total_ms, 3),
        "speedup_vs_exhaustive": round(ex_ms / r.total_ms, 2),
        "recall_top1pct": round(S.recall(idx, exact, 0.01), 4),
        "recall_top0.1pct": round(S.recall(idx, exact, 0.001), 4),
        "fallback": bool(r.fallback_pairs),
        "phases_ms": {k: round(getattr(r, k), 3) for k in ("label_ms", "train_ms", "score_ms", "select_ms",
                                                           "rescore_ms")},
    }
    log(f"screen (field surrogate): {field}")
    out = {
        "workload": (f"c3: 4,096-atom protein, 26-atom ligand; surrogate trained on {w.train_poses.shape[1]} plain-"
                     f"kernel energies, {n} fresh poses screened, lowest keep_fraction rescored exactly"),
        "exhaustive_ms": round(ex_ms, 3),
        "exhaustive_poses_per_s": round(n / (ex_ms * 1e-3), 1),
        "val_rmse": round(st.val_rmse, 4),
        "epochs": int(cfg.epochs),
        "train_s": round((st.label_ms + st.train_ms) * 1e-3, 2),
        "points": points,
        "field_surrogate": field,
    }
    return out


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


def traffic() -> float | None:
    if TRAFFIC_SUMMARY.exists():
        try:
            return float(json.loads(TRAFFIC_SUMMARY.read_text())["hbm_bytes_per_launch"])
        except Exception:
            return None
    return None


def combined(full: bool, kinds=("field", "mlp"), nproteins: int = 0) -> dict:
    """Both surrogates of the combined screen on the same workload: every
    protein screened (timed), exact energies of every pair for recall (all
    proteins at the default size, the first one at full size, where the
    exhaustive time of the rest is extrapolated from its interaction rate)."""
    from minibude_amd import combined as C
    shape = dict(COMBINED_FULL if full else COMBINED_DEFAULT)
    if nproteins:
        shape["nproteins"] = nproteins
    t0 = time.time()
    w = C.combined_workload(**shape)
    per_protein = [len(p) * sum(len(l) * q.shape[1] for l, q in zip(w.ligands, w.poses)) for p in w.proteins]
    log(f"combined: {len(w.proteins)} proteins x {len(w.ligands)} ligands x {shape['nposes']} poses "
        f"({w.total_poses} pair-poses, {w.interactions:.3e} interactions), generated in {time.time() - t0:.1f} s")
    out = {
        "workload": (f"c4: {len(w.proteins)} proteins of 900-6000 atoms x {len(w.ligands)} ligands of 8-96 atoms x "
                     f"{shape['nposes']} poses per pair" +
                     ("" if full and not nproteins else " (reduced from 16 proteins x 512 ligands x 2^20 poses)")),
        "pairs": w.pairs,
        "pair_poses": w.total_poses,
        "interactions": w.interactions,
    }
    for kind in kinds:
        cfg = C.default_config(kind)
        screen_ms, ex_ms, ex_inter, rec, fb, labels, sel = 0.0, 0.0, 0, [], 0, 0, 0
        parts = {"label_ms": 0.0, "train_ms": 0.0, "score_ms": 0.0, "select_ms": 0.0, "rescore_ms": 0.0}
        sweep = {k: [0.0, []] for k in (COMBINED_FIELD_KEEPS if kind == "field" else ())}
        keep0 = float(cfg.keep_fraction)
        one_by_one = 0.0
        with C.Combined(w.forcefield, w.ligands, w.poses, cfg, 0) as c:
            # the timed screen: every protein in one batch call (each one's
            # scoring beside the previous one's rescoring), after one untimed
            # screen of the first protein that sizes the object's buffers
            c.screen(w.proteins[0])
            results, screen_ms = c.screen_batch(w.proteins)
            for r in results:
                fb += r.fallback_pairs
                labels += r.labels
                sel += r.nselected
                for k in parts:
                    parts[k] += getattr(r, k)
            # recall from one-by-one screens of the same proteins (a protein
            # draws the same samples alone or in a batch, so their selections
            # equal the batch's), exact energies of every pair
            for p, protein in enumerate(w.proteins):
                if not full or p == 0:
                    one_by_one += c.screen(protein).total_ms
                    ex_ms += c.exhaustive()
                    ex_inter += per_protein[p]
                    rec.extend(c.recall(0.01).tolist())
                    # smaller rescored fractions on the same protein
                    for keep, acc in sweep.items():
                        c.set_keep(keep)
                        acc[0] += c.screen(protein).total_ms
                        acc[1].extend(c.recall(0.01).tolist())
                    if sweep:
                        c.set_keep(keep0)
        rec = np.array(rec)
        ex_total_ms = ex_ms * w.interactions / ex_inter
        res = {
            "poses_per_s": round(w.total_poses / (screen_ms * 1e-3), 1),
            "screen_ms": round(screen_ms, 2),
            "screen_ms_one_by_one": None if full else round(one_by_one, 2),
            "exhaustive_ms": round(ex_total_ms, 2),
            "exhaustive_poses_per_s": round(w.total_poses / (ex_total_ms * 1e-3), 1),
            "speedup_vs_exhaustive": round(ex_total_ms / screen_ms, 2),
            "recall_top1pct_mean": round(float(rec.mean()), 4),
            "recall_top1pct_min": round(float(rec.min()), 4),
            "recall_top1pct_p05": round(float(np.percentile(rec, 5)), 4),
            "pairs_recall_ge_0.99": round(float((rec >= 0.99).mean()), 4),
            "recall_pairs_measured": int(len(rec)),
            "fallback_pairs": fb,
            "labels": labels,
            "rescored": sel,
            "phases_ms": {k: round(v, 2) for k, v in parts.items()},
            "config": {**{k: (round(v, 6) if isinstance(v, float) else v) for k, v in cfg.as_dict().items()},
                       "keep_fraction": keep0},
        }
        if sweep:
            ex_meas = ex_ms  # the sweep ran on the proteins whose exhaustive time was measured
            res["keep_sweep"] = [{"keep_fraction": k, "speedup_vs_exhaustive": round(ex_meas / v[0], 2),
                                  "recall_top1pct_mean": round(float(np.mean(v[1])), 4),
                                  "recall_top1pct_min": round(float(np.min(v[1])), 4),
                                  "pairs_recall_ge_0.99": round(float((np.array(v[1]) >= 0.99).mean()), 4)}
                                 for k, v in sweep.items()]
        if full:
            res["exhaustive_note"] = "measured on the first protein's pairs, extrapolated by interactions"
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
    ap.add_argument("--sweep", action="store_true",
                    help="also run the ppwi x wgsize autotune sweep of the plain kernel (40 points)")
    ap.add_argument("--screen-epochs", type=int, default=0, help="training epochs of the learned prescreen "
                    "(default: its configuration's 60; fewer only to keep a kernel trace short)")
    ap.add_argument("--combined-full", action="store_true",
                    help="combined screen at 16 proteins x 512 ligands x 2^20 poses (recall on the first protein)")
    ap.add_argument("--combined-surrogates", default="field,mlp", help="comma list of field, mlp")
    ap.add_argument("--combined-proteins", type=int, default=0, help="screen only the first N proteins")
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
        sweep_out = sweep(ctx, ref, 1, 2) if a.sweep else None
    host = np.array(s.kernel_ms[a.warmup:])
    dev = np.array(s.device_ms[a.warmup:])
    ms_per_step = float(host.mean())
    inter = metrics.interactions_per_pass(deck.nposes, deck.natlig, deck.natpro)
    value = inter / (ms_per_step * 1e-3) / 1e9
    flops = metrics.flops_per_pass(deck.nposes, deck.natlig, deck.natpro, a.ppwi)
    achieved = flops / (float(dev.mean()) * 1e-3) / 1e12
    peak = metrics.FP32_VECTOR_PEAK_TFLOPS
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
            "traffic": traffic(),
            "kernel_ms": round(float(dev.mean()), 4),
            "flop_accounting": "src/main.cpp:358-360 (30 + 10/PPWI per interaction)",
        },
        "cpu_baseline": cpu,
        "valid": valid,
        "parity": parity,
        "sweep": sweep_out,
        "batched": None if a.no_batched else batched(),
        "screen": None if a.no_screen else screen(a.screen_epochs),
        "combined": None if a.no_combined else combined(a.combined_full, tuple(a.combined_surrogates.split(",")),
                                                         a.combined_proteins),
        "device": device_name(0),
    }
    print(json.dumps(out), flush=True)
    return 0 if valid in (None, True) else 1


if __name__ == "__main__":
    sys.exit(main())
