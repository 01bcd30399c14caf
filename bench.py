"""Benchmark: cell-updates/s of the explicit step on one MI355X.

Workload (default): 4 blocks of 1024x1024 cells in a 2x2 layout (4.19 M cells, 2 ghost layers),
seeded angled planar shock (Mach ~ U[1.5,3]) into gamma=1.4 gas at rho=1, p=1 with a 5 % 8-mode density
perturbation; MUSCL van Leer, HLLC, SSP-RK3, CFL 0.5, y-periodic (BC 201), inter-block BC 101, x-ends
BC 400; block ghosts exchanged on the device.  Both arithmetic forms run from the same initial state:
  plain : FP64 everywhere (bit-identical to the CPU oracle)
  mixed : FP32 reconstruction / limiter / Riemann flux, FP64 state, residual, RK update and reductions
`value` is the mixed form's cell-updates/s (1 cell-update = one interior cell advanced one full
step, all RK stages); the plain form, the speed-up and the L1/Linf density error of the mixed form
against the plain FP64 form are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STAGE = 168      # state in 32 + RK base 32 + state out 32 + geometry 72 (FP64)
BYTES_DT_PER_CELL = 80          # dt pass: state 32 + M 32 + dl 16


PROFILE = ROOT / "profiles" / "k_stage_planar_shock_4x1024_rk3.md"


def measured_traffic(n):
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (copy-kernel calibrated); None for other sizes or when the summary is absent."""
    if n != 1024 or not PROFILE.exists():
        return {}
    for line in PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return json.loads(line.split(":", 1)[1])
    return {}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def time_form(problem_fn, precision, steps, warmup):
    from q2d_amd import GpuSolver
    pb = problem_fn(precision)
    g = GpuSolver(pb)
    ncells = pb.ncells()
    g.run(warmup)                                  # graph capture + warm-up steps (synchronised)
    t0 = time.perf_counter()
    out = g.run(steps)                             # K steps, synchronised at the end
    t1 = time.perf_counter()
    rho = [g.state(b)[2:-2, 2:-2, 0].copy() for b in range(len(pb.blocks))]
    # per-launch stage-kernel durations (HIP events on the solver's stream)
    g.set_timing(True)
    g.run(5)
    nlaunch, ms = g.timing_stats()
    g.set_timing(False)
    n_rk = pb.cases[0].n_rk
    sec = t1 - t0
    res = {
        "cell_updates_per_s": ncells * steps / sec,
        "ms_per_step": 1e3 * sec / steps,
        "stage_kernel_ms": ms / max(1, nlaunch),
        "stage_launches_timed": nlaunch,
        "ncells": ncells,
        "n_rk": n_rk,
        "iter": out[0]["iter"],
        "tiling": g.tiling(),
        "work_items": g.info()[1],
    }
    g.close()
    return res, rho


def cpu_baseline(problem_fn, steps):
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    pb = problem_fn("fp64")
    o = oracle.Oracle(pb, threads=threads)
    t0 = time.perf_counter()
    o.run(steps, stop_on_todo=False)
    sec = time.perf_counter() - t0
    return {"value": pb.ncells() * steps / sec, "unit": "cell-updates/s", "cores": oracle.load().q2do_get_threads(),
            "kind": "port",
            "sample": f"CPU oracle (C restatement, OpenMP) on the same 4x{int(round((pb.ncells() / 4) ** 0.5))}^2 "
                      f"workload, {steps} SSP-RK3 steps ({sec:.1f} s); CPU: {cpu_info()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["both", "fp64", "mixed"], default="both",
                    help="profiling aid: run only one arithmetic form (no JSON line)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only the HBM reference copy kernel (1 GiB, 10 launches) and print GB/s")
    args = ap.parse_args()
    if args.copy_probe:
        from q2d_amd.capi import hbm_copy_gbps
        gbs, ms = hbm_copy_gbps()
        print(json.dumps({"copy_GBps": gbs, "ms_per_copy": ms, "bytes_per_copy": 2 << 30}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the multi-GPU sharded step is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from q2d_amd import cases

    def problem_fn(precision):
        return cases.planar_shock(seed=args.seed, n=args.n, steps=1 << 30, precision=precision)

    log(f"bench: planar shock 4 x {args.n}^2, SSP-RK3, warmup {args.warmup}, steps {args.steps}")
    if args.form != "both":
        r, _ = time_form(problem_fn, args.form, args.steps, args.warmup)
        log(f"bench: {args.form} {r['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s, stage kernel {r['stage_kernel_ms']:.4f} ms")
        return
    plain, rho64 = time_form(problem_fn, "fp64", args.steps, args.warmup)
    log(f"bench: plain fp64  {plain['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s")
    mixed, rho32 = time_form(problem_fn, "mixed", args.steps, args.warmup)
    log(f"bench: mixed       {mixed['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s")
    num = sum(np.abs(a - b).sum() for a, b in zip(rho32, rho64))
    den = sum(np.abs(b).sum() for b in rho64)
    l1 = num / den
    linf = max(np.abs(a - b).max() for a, b in zip(rho32, rho64)) / max(np.abs(b).max() for b in rho64)

    n_rk = plain["n_rk"]
    bytes_update = n_rk * BYTES_PER_CELL_STAGE + BYTES_DT_PER_CELL

    traffic = measured_traffic(args.n)

    def roofline(r, form):
        alg = BYTES_PER_CELL_STAGE * r["ncells"]            # bytes of one stage launch
        ach = alg / (r["stage_kernel_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic.get(form),
                "kernel": "k_stage", "kernel_ms": round(r["stage_kernel_ms"], 4),
                "alg_bytes_per_launch": alg}

    line = {
        "metric": "cell-updates/s batched-mixed vs plain FP64 at stated err; p50/p99 solve latency",
        "value": mixed["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flux / f64 state",
        "data": "synthetic (seeded planar shock)",
        "config": {"workload": f"planar shock, 4 blocks of {args.n}x{args.n} (2x2), MUSCL vanleer, HLLC, SSP-RK3, "
                               f"CFL 0.5, BC 101/201/400, device ghost exchange",
                   "cells": plain["ncells"], "parallelism": "single GPU"},
        "plain_fp64": {"value": plain["cell_updates_per_s"], "ms_per_step": plain["ms_per_step"],
                       "roofline": roofline(plain, "fp64"), "dtype": "f64"},
        "mixed_vs_plain": round(mixed["cell_updates_per_s"] / plain["cell_updates_per_s"], 3),
        "mixed_err_vs_fp64": {"rel_L1_rho": l1, "rel_Linf_rho": linf, "steps": args.warmup + args.steps},
        "alg_bytes_per_cell_update": bytes_update,
        "effective_GBps": {"plain": plain["cell_updates_per_s"] * bytes_update / 1e9,
                           "mixed": mixed["cell_updates_per_s"] * bytes_update / 1e9},
        "roofline": roofline(mixed, "mixed"),
    }
    if not args.no_cpu:
        log("bench: CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(problem_fn, args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
