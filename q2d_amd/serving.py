"""Serving: continuous batching of independent steady solves on one device-resident solver.

Requests (one case each) arrive over time.  The server admits a waiting request as soon as its HBM
footprint fits the budget (q2d_gpu_add_cases: its state is uploaded once and stays resident), advances
every admitted case together in chunks of graph-replayed explicit steps (one launch sequence per step
for the whole mix, so the requests' work overlaps on the device), and retires a case the moment the
device's per-case stop test fires: its residual(1) dropped by `res_drop` relative to its first step,
or it reached its iteration cap.  The retired state is downloaded and the case removed, freeing its
memory for the queue.  Latency is arrival to retirement; throughput counts only the cell-updates of
admitted work.

Because every case's arithmetic is independent of the others in the launch, a request's result (state,
iteration count) is bit-identical whether it runs alone or in any mix.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace

import numpy as np

from . import cases
from .capi import GpuSolver


def block_planes(spec):
    """FP64 planes of one block of a case, as q2d_gpu_add_cases allocates them: 30 for an FP64 Euler block,
    37 for a mixed-form Euler block (its FP32 geometry), 39 for a laminar block, plus 4 with residual
    smoothing (the stage's scaled residual)."""
    planes = 39 if spec.model > 0 else (37 if spec.precision == "mixed" else 30)
    return planes + (4 if spec.irs_beta > 0 else 0)


def footprint_bytes(problem):
    """Device bytes the blocks of `problem` occupy: planes of (nj + 4) rows of round16(ni + 4) doubles."""
    total = 0
    for b in problem.blocks:
        pitch = (b.ni + 4 + 15) // 16 * 16
        total += 8 * block_planes(problem.cases[b.case_id]) * pitch * (b.nj + 4)
    return total


@dataclass
class Request:
    rid: int
    arrival: float                # seconds after the stream starts
    problem: cases.Problem        # exactly one case
    meta: dict = field(default_factory=dict)


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    iters: int
    status: int
    converged: bool
    cells: int
    state: list                   # interior state of each block at retirement

    @property
    def latency(self):
        return self.finished - self.arrival


def ramp_request(seed, smin=0.0, smax=1.0, precision="fp64", res_drop=1e-6, itermax=60000):
    """One compression-ramp solve: wedge ~ U[5, 15] deg, inflow Mach ~ U[1.5, 4], grid log-uniform on the
    diagonal from 64x64 (s = 0) to 4096x1024 (s = 1), local dt, SSP-RK3, run to a `res_drop` drop of
    residual(1) (capped at `itermax` steps).  FP64 by default: with FP32 face fluxes the residual of a
    steady solve stalls on the flux rounding well before a 1e-6 drop."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(smin, smax)
    wedge = rng.uniform(5.0, 15.0)
    mach = rng.uniform(1.5, 4.0)
    ni = max(64, int(round(64 * 64 ** s / 8.0)) * 8)
    nj = max(64, int(round(64 * 16 ** s / 8.0)) * 8)
    pb = cases.ramp_channel(ni=ni, nj=nj, wedge_deg=wedge, mach=mach, precision=precision, n_rk=3,
                            limiter="minmod", riemann="HLLE", itermax=itermax)
    pb.cases[0] = replace(pb.cases[0], res_drop=res_drop)
    pb.meta.update({"ni": ni, "nj": nj, "wedge": wedge, "mach": mach})
    return pb


def poisson_stream(n, rate, seed=0, **kw):
    """n ramp requests with exponential inter-arrival times of mean 1/rate seconds."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(k, t, ramp_request(int(rng.integers(0, 2 ** 31)), **kw)))
    return out


class Server:
    """Continuous-batching request loop on one GpuSolver."""

    def __init__(self, budget_bytes=64 << 30, chunk_steps=50):
        self.budget = int(budget_bytes)
        self.chunk = int(chunk_steps)
        self.g = GpuSolver(cases.Problem())

    def close(self):
        self.g.close()

    def serve(self, requests, keep_state=True, clock=time.perf_counter, log=None):
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if footprint_bytes(r.problem) > self.budget:
                raise ValueError(f"request {r.rid} alone exceeds the HBM budget")
        queue, active, results = [], {}, []
        used = 0
        cell_updates = 0
        t0 = clock()
        nxt = 0
        while nxt < len(pending) or queue or active:
            now = clock() - t0
            while nxt < len(pending) and pending[nxt].arrival <= now:
                queue.append(pending[nxt])
                nxt += 1
            while queue and used + footprint_bytes(queue[0].problem) <= self.budget:   # FIFO admission
                r = queue.pop(0)
                c, b0 = self.g.add_cases(r.problem)
                fp = footprint_bytes(r.problem)
                used += fp
                active[c] = (r, clock() - t0, b0, len(r.problem.blocks), fp, 0)
            if not active:
                if nxt < len(pending):
                    time.sleep(max(0.0, min(pending[nxt].arrival - (clock() - t0), 0.05)))
                continue
            out = self.g.run(self.chunk, raise_on_error=False)
            tnow = clock() - t0
            for c in list(active):
                r, tadm, b0, nb, fp, last_iter = active[c]
                o = out[c]
                cell_updates += r.problem.ncells() * (o["iter"] - last_iter)
                active[c] = (r, tadm, b0, nb, fp, o["iter"])
                if o["todo"] == 3 or o["status"] != 0:
                    st = [self.g.state(b0 + k)[2:-2, 2:-2].copy() for k in range(nb)] if keep_state else []
                    conv = o["status"] == 0 and o["iter"] < r.problem.cases[0].itermax
                    results.append(Result(r.rid, r.arrival, tadm, tnow, o["iter"], o["status"], conv,
                                          r.problem.ncells(), st))
                    self.g.remove_case(c)
                    used -= fp
                    del active[c]
                    if log:
                        log(f"serve: request {r.rid} ({r.problem.blocks[0].ni}x{r.problem.blocks[0].nj}) "
                            f"done at {tnow:.2f} s, {o['iter']} steps, latency {tnow - r.arrival:.2f} s")
        makespan = clock() - t0
        results.sort(key=lambda x: x.rid)
        # time-to-converged percentiles are over the requests that converged; a request retired at its
        # iteration cap (or failed) is counted separately and does not enter them
        lat = np.array([x.latency for x in results if x.converged])
        lat_all = np.array([x.latency for x in results])
        stats = {
            "requests": len(results),
            "converged": int(sum(x.converged for x in results)),
            "stalled": [x.rid for x in results if not x.converged],
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "mean_s": float(lat.mean()) if len(lat) else math.nan,
            "p99_all_s": float(np.percentile(lat_all, 99)) if len(lat_all) else math.nan,
            "makespan_s": makespan,
            "cell_updates": int(cell_updates),
            "cell_updates_per_s": cell_updates / makespan if makespan > 0 else 0.0,
        }
        return results, stats
