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
  valid         energies of the last timed step validated against the
                reference energies of the deck (max_diff_% < 0.025)
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

METRIC = json.loads((ROOT / "BASELINE.json").read_text())["metric"] if (ROOT / "BASELINE.json").exists() else \
    "G-interactions/s"
WORKLOAD = "c1"
DEFAULT_PPWI = 8
DEFAULT_WGSIZE = 256
TRAFFIC_SUMMARY = ROOT / "profiles" / "fasten_plain_c1_traffic.json"
CPU_SAMPLE_POSES = 65536
CPU_ITERS = 2


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(deck: D.Deck, ref: np.ndarray) -> dict | None:
    """The reference omp model on the first CPU_SAMPLE_POSES poses of the deck."""
    flags = Path("/proc/cpuinfo").read_text() if Path("/proc/cpuinfo").exists() else ""
    v4 = " avx512f " in flags.replace("\n", " ")
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
        cmd = [str(exe), "--deck", tmp, "-n", str(n), "-i", str(CPU_ITERS), "-p", "64", "--csv"]
        log("cpu baseline: " + " ".join(cmd[:1] + cmd[2:2] + cmd[3:]))
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
        wall = time.time() - t0
    if r.returncode not in (0, 1):
        log(f"cpu baseline failed ({r.returncode}): {r.stderr[-400:]}")
        return None
    rows = [ln for ln in r.stdout.splitlines() if re.match(r"^\d+,\d+,", ln)]
    if not rows:
        log(f"cpu baseline: no CSV row in output: {r.stdout[-400:]}")
        return None
    f = rows[0].split(",")
    ips = float(f[7])
    valid = "valid=1" in r.stdout
    return {
        "value": round(ips / 1e9, 4),
        "unit": "G-interactions/s",
        "cores": cores,
        "kind": "reference",
        "sample": (f"first {n} of {deck.nposes} poses of the {WORKLOAD} deck ({deck.natpro} protein / {deck.natlig} "
                   f"ligand atoms); reference omp model {exe.name} (src/main.cpp + src/omp/fasten.hpp, -O3 "
                   f"-ffast-math -fopenmp), ppwi=64, 2 warmup + {CPU_ITERS} timed passes, {wall:.1f}s wall, "
                   f"valid={valid}"),
    }


def traffic() -> float | None:
    if TRAFFIC_SUMMARY.exists():
        try:
            return float(json.loads(TRAFFIC_SUMMARY.read_text())["hbm_bytes_per_launch"])
        except Exception:
            return None
    return None


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ppwi", type=int, default=DEFAULT_PPWI)
    ap.add_argument("--wgsize", type=int, default=DEFAULT_WGSIZE)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU only: the plain kernel is data-parallel over poses and this build runs "
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

    valid, max_diff, sym = None, None, None
    if ref is not None:
        max_diff, failed = metrics.validate(ref, s.energies)
        valid = metrics.is_valid(max_diff)
        sym = metrics.symmetric_max_rel_pct(ref, s.energies)
        log(f"validation vs reference energies: valid={valid} max_diff_%={max_diff:.5f} symmetric={sym:.5f}")

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
        "max_diff_pct": None if max_diff is None else round(max_diff, 6),
        "symmetric_max_rel_pct": None if sym is None else round(sym, 6),
        "device": device_name(0),
    }
    print(json.dumps(out), flush=True)
    return 0 if valid in (None, True) else 1


if __name__ == "__main__":
    sys.exit(main())
