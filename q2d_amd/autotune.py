"""Tile autotuner for the stage kernel.

The stage kernel's launch shape has two knobs (q2d_gpu_set_tiling): rows per work item H (how far one
wavefront marches down its 60-column strip: longer strips reuse the rolled j-stencil over more rows,
shorter ones give more items for small or uneven batches) and wavefronts per workgroup. The best pair
depends on the batch's mix of grid sizes, so it is measured on the device: every candidate runs a few
explicit steps of the actual problem on a scratch solver, and the fastest is kept. Results are cached
per problem signature (arithmetic form and the per-case grid sizes and schemes) in a JSON file
($Q2D_TUNE_CACHE, default ~/.cache/q2d_amd/tiles.json).
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from dataclasses import replace
from pathlib import Path

ROWS = (-1, 8, 16, 32, 64)          # -1: the library's occupancy-balanced row count per launch
WAVES = (1, 2, 4)


def cache_path():
    return Path(os.environ.get("Q2D_TUNE_CACHE", Path.home() / ".cache" / "q2d_amd" / "tiles.json"))


def signature(problem):
    parts = []
    for c, spec in enumerate(problem.cases):
        dims = sorted((b.ni, b.nj) for b in problem.blocks if b.case_id == c)
        parts.append((spec.precision, spec.riemann, spec.limiter, spec.model, spec.source, spec.n_rk,
                      spec.time_accurate, tuple(dims)))
    parts.sort()
    return hashlib.sha1(json.dumps(parts).encode()).hexdigest()


def _load():
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except Exception:
        return {}


def _store(key, value):
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        d = _load()
        d[key] = value
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(d, indent=1))
        tmp.replace(p)
    except OSError:
        pass


def tune(problem, steps=4, rows=ROWS, waves=WAVES, use_cache=True, verbose=False):
    """Return (rows_per_item, waves_per_group, timings) for `problem`, measuring every candidate on a
    scratch solver (the caller's solver state is untouched)."""
    from .capi import GpuSolver
    key = signature(problem)
    if use_cache:
        hit = _load().get(key)
        if hit:
            return hit["rows"], hit["waves"], hit.get("timings", {})
    # scratch copy whose cases never reach their stop test while being timed
    scratch = replace(problem, cases=[replace(c, itermax=1 << 30, res_threshold=-1.0, time_threshold=1e300)
                                      for c in problem.cases])
    g = GpuSolver(scratch)
    timings = {}
    try:
        for w in waves:
            for r in rows:
                g.set_tiling(r, w)
                g.run(1)                                      # graph capture for this shape
                t0 = time.perf_counter()
                out = g.run(steps)
                dt = time.perf_counter() - t0
                if any(o["status"] != 0 for o in out):
                    break
                timings[f"{r}x{w}"] = dt / steps
                if verbose:
                    print(f"autotune: rows {r:3d} waves {w}: {1e3 * dt / steps:.3f} ms/step", flush=True)
    finally:
        g.close()
    if not timings:
        return 16, 4, {}
    best = min(timings, key=timings.get)
    r, w = (int(v) for v in best.split("x"))
    if use_cache:
        _store(key, {"rows": r, "waves": w, "timings": timings})
    return r, w, timings
