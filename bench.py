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


def time_form(problem_fn, precision, steps, warmup, tune=False):
    from q2d_amd import GpuSolver
    pb = problem_fn(precision)
    g = GpuSolver(pb)
    ncells = pb.ncells()
    if tune:
        g.autotune()                               # tile shape measured on a scratch solver (cached)
    # Stage-kernel durations are measured on the timed steps themselves: with timing on, every stage
    # launch stamps its first workgroup start and last wavefront end on the device wall clock (two
    # atomics per workgroup); the graph replays are otherwise unchanged.
    g.set_timing(True)
    g.run(warmup)                                  # graph capture + warm-up steps (synchronised)
    g.set_timing(True)                             # clear the warm-up launches from the statistics
    t0 = time.perf_counter()
    out = g.run(steps)                             # K steps, synchronised at the end
    t1 = time.perf_counter()
    nlaunch, ms = g.timing_stats()
    g.set_timing(False)
    rho = [g.state(b)[2:-2, 2:-2, 0].copy() for b in range(len(pb.blocks))]
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


def other_workload(args):
    """--workload axi | batch: one JSON line for the laminar axisymmetric case (2048x512) or the batch of
    64 uneven planar / quasi-2D cases, both arithmetic forms from the same initial state."""
    from q2d_amd import cases
    if args.workload == "axi":
        def problem_fn(precision):
            return cases.axisymmetric_pipe(seed=args.seed, precision=precision, steps=1 << 30)
        desc = ("axisymmetric laminar Navier-Stokes pipe, 2048x512, tanh wall clustering (first cell 1e-5 L), "
                "Re_L 1e5, Sutherland, Pr 0.72, isothermal wall, axis BC, MUSCL vanleer, HLLC, SSP-RK3, local dt, "
                "CFL 0.5, VNN 0.2")
        tune = False
    else:
        def problem_fn(precision):
            return cases.batch_uneven(seed=args.seed, ncases=64, steps=1 << 30, precision=precision)
        desc = ("64 independent planar / quasi-2D cases, 128x64 .. 2048x1024 log-uniform, Riemann or uniform "
                "inflow, HLL or HLLC per case, SSP-RK2, one persistent launch per stage, autotuned tiles")
        tune = True
    log(f"bench: {args.workload}: warmup {args.warmup}, steps {args.steps}")
    plain, rho64 = time_form(problem_fn, "fp64", args.steps, args.warmup, tune)
    log(f"bench: plain fp64  {plain['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s  tiling {plain['tiling']}")
    mixed, rho32 = time_form(problem_fn, "mixed", args.steps, args.warmup, tune)
    log(f"bench: mixed       {mixed['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s  tiling {mixed['tiling']}")
    num = sum(np.abs(a - b).sum() for a, b in zip(rho32, rho64))
    den = sum(np.abs(b).sum() for b in rho64)
    linf = max(np.abs(a - b).max() for a, b in zip(rho32, rho64)) / max(np.abs(b).max() for b in rho64)
    print(json.dumps({
        "metric": "cell-updates/s batched-mixed vs plain FP64 at stated err",
        "value": mixed["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 flux / f64 state",
        "data": "synthetic (seeded)",
        "config": {"workload": desc, "cells": plain["ncells"], "parallelism": "single GPU"},
        "plain_fp64": {"value": plain["cell_updates_per_s"], "ms_per_step": plain["ms_per_step"],
                       "stage_kernel_ms": plain["stage_kernel_ms"], "tiling": plain["tiling"]},
        "mixed_stage_kernel_ms": mixed["stage_kernel_ms"], "mixed_tiling": mixed["tiling"],
        "mixed_vs_plain": round(mixed["cell_updates_per_s"] / plain["cell_updates_per_s"], 3),
        "mixed_err_vs_fp64": {"rel_L1_rho": num / den, "rel_Linf_rho": linf, "steps": args.warmup + args.steps},
    }))


def irs_workload(args):
    """--workload irs: implicit residual smoothing (irs-beta 0.5) in the reference's steady setting
    (compression ramp, MUSCL minmod, HLLE, SSP-RK3, local dt, CFL 0.9: test/2D/euler/ramp-channel/
    input.ini).  Step throughput with and without smoothing on a 4096x1024 ramp, both forms; then, unless
    --no-conv, steps and wall time to a 1e-6 drop of residual(1) on a 256x128 ramp without smoothing
    and with it, at CFL 0.9 and at CFL 1.8."""
    from q2d_amd import GpuSolver, cases

    def ramp(precision, beta, ni=4096, nj=1024, cfl=0.9):
        pb = cases.ramp_channel(ni=ni, nj=nj, wedge_deg=10.0, mach=2.5, precision=precision, n_rk=3,
                                limiter="minmod", riemann="HLLE", cfl=cfl, itermax=1 << 30)
        pb.cases[0].irs_beta = beta
        return pb

    thr = {}
    for prec in ("fp64", "mixed"):
        for beta in (0.0, 0.5):
            r, _ = time_form(lambda p, b=beta: ramp(p, b), prec, args.steps, args.warmup)
            thr[f"{prec}_beta{beta}"] = {"cell_updates_per_s": r["cell_updates_per_s"], "ms_per_step": r["ms_per_step"],
                                         "stage_kernel_ms": round(r["stage_kernel_ms"], 4)}
            log(f"bench: irs {prec} beta {beta}: {r['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s, "
                f"{r['ms_per_step']:.3f} ms/step, stage kernel {r['stage_kernel_ms']:.4f} ms")
    conv = []
    if not args.no_conv:
        for beta, cfl in ((0.0, 0.5), (0.0, 0.9), (0.5, 0.9), (0.5, 1.8)):
            pb = ramp("fp64", beta, ni=256, nj=128, cfl=cfl)
            pb.cases[0].res_drop = 1e-6
            g = GpuSolver(pb)
            t0 = time.perf_counter()
            while True:
                out = g.run(500, raise_on_error=False)[0]
                if out["todo"] == 3 or out["status"] != 0 or out["iter"] >= 60000:
                    break                          # converged, failed (NaN or p < 0: the reference stops), or capped
            status = out["status"]
            sec = time.perf_counter() - t0
            g.close()
            conv.append({"irs_beta": beta, "cfl": cfl, "steps": out["iter"], "seconds": round(sec, 3),
                         "converged": status == 0 and out["todo"] == 3, "status": status})
            log(f"bench: irs convergence beta {beta} cfl {cfl}: {out['iter']} steps, {sec:.2f} s, status {status}")
    m, p = thr["mixed_beta0.5"], thr["fp64_beta0.5"]
    print(json.dumps({
        "metric": "cell-updates/s with implicit residual smoothing (irs-beta 0.5)",
        "value": m["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 flux / f64 state",
        "data": "synthetic (compression ramp)",
        "config": {"workload": "compression ramp 4096x1024 (10 deg, Mach 2.5), MUSCL minmod, HLLE, SSP-RK3, local dt, "
                               "CFL 0.9, irs-beta 0.5", "cells": 4096 * 1024, "parallelism": "single GPU"},
        "throughput": thr,
        "irs_cost": {"fp64": round(thr["fp64_beta0.0"]["cell_updates_per_s"] / p["cell_updates_per_s"], 3),
                     "mixed": round(thr["mixed_beta0.0"]["cell_updates_per_s"] / m["cell_updates_per_s"], 3)},
        "convergence_256x128": conv,
    }))


def serve_workload(nreq, seed, log_fn=None):
    """Poisson stream of steady compression-ramp solves (wedge ~ U[5,15] deg, Mach ~ U[1.5,4], grids
    log-uniform 64x64 .. 4096x1024), each run in FP64 to a 1e-6 drop of residual(1); continuous batching
    on one device-resident solver under a 64 GiB HBM budget.  The arrival rate is set for ~70 % occupancy from
    a cost estimate (1,000 steps at 64x64 growing as cells^0.275, the oracle's measured convergence on
    this family, at 4e9 cell-updates/s)."""
    from q2d_amd import serving
    probe = serving.poisson_stream(nreq, rate=1.0, seed=seed)
    est = sum(r.problem.ncells() * 1000.0 * (r.problem.ncells() / 4096.0) ** 0.275 for r in probe) / 4e9
    rate = 0.7 * nreq / max(est, 1e-3)
    reqs = serving.poisson_stream(nreq, rate=rate, seed=seed)
    srv = serving.Server(budget_bytes=64 << 30, chunk_steps=50)
    res, stats = srv.serve(reqs, keep_state=False, log=log_fn)
    srv.close()
    stats.update({"arrival_rate_per_s": rate, "budget_GiB": 64, "chunk_steps": 50,
                  "grid_cells_min": min(r.problem.ncells() for r in reqs),
                  "grid_cells_max": max(r.problem.ncells() for r in reqs),
                  "steps_total": int(sum(x.iters for x in res))})
    return stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["both", "fp64", "mixed"], default="both",
                    help="profiling aid: run only one arithmetic form (no JSON line)")
    ap.add_argument("--requests", type=int, default=24, help="serving: requests in the Poisson stream")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-conv", action="store_true", help="irs: skip the convergence runs")
    ap.add_argument("--workload", choices=["shock", "axi", "batch", "serve", "irs"], default="shock",
                    help="shock: the headline workload (default); axi, batch: the laminar axisymmetric case and "
                         "the uneven 64-case batch, reported on their own JSON line")
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

    if args.workload == "serve":
        stats = serve_workload(args.requests, args.seed, log)
        print(json.dumps({"metric": "p50/p99 solve latency (serving, continuous batching)", "value": stats["p50_s"],
                          "unit": "s", "n_gpus": 1, "higher_is_better": False, "data": "synthetic (seeded)",
                          "dtype": "f64", "serving": stats}))
        return
    if args.workload == "irs":
        irs_workload(args)
        return
    if args.workload != "shock":
        other_workload(args)
        return

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
                "kernel": "k_stage_mx" if form == "mixed" else "k_stage", "kernel_ms": round(r["stage_kernel_ms"], 4),
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
    if not args.no_serve:
        log("bench: serving stream ...")
        line["serving"] = serve_workload(args.requests, args.seed)
        log(f"bench: serving p50 {line['serving']['p50_s']:.2f} s, p99 {line['serving']['p99_s']:.2f} s")
    if not args.no_cpu:
        log("bench: CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(problem_fn, args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
