"""ctypes mirror of include/q2d_gpu.h and a thin Python solver object.

The HIP library (q2d_amd/lib/libq2d_gpu.so) is the only compute path; if it
cannot be loaded, or no GPU is present, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

GC = 2
NPRIM = 4

Q2D_OK = 0
ERRORS = {
    -1: "Q2D_E_ARG", -2: "Q2D_E_HIP", -3: "Q2D_E_NOMEM", -4: "Q2D_E_STATE",
    -5: "Q2D_E_INTERFACE", -6: "Q2D_E_UNSUPPORTED", -7: "Q2D_E_NODEVICE",
}
Q2D_E_STATE = -4
Q2D_E_INTERFACE = -5

LIMITERS = {"first-order": 0, "iord": 0, "minmod": 1, "vanleer": 2, "vanalbada": 3, "mc": 4, "superbee": 5, "LIMO3": 6}
RIEMANN = {"HLLC": 0, "HLLE": 1, "HLLEM": 2, "HLLE++": 3, "HLLC+": 4}
SOURCES = {"none": 0, "axisymmetric": 1, "area_variation": 2}
PRECISION = {"fp64": 0, "mixed": 1}
INTVARS = {"cons": 0, "prim": 1}


class Q2DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Scheme(C.Structure):
    _fields_ = [("n_rk", C.c_int32), ("limiter", C.c_int32), ("riemann", C.c_int32), ("time_accurate", C.c_int32),
                ("model", C.c_int32), ("source", C.c_int32), ("precision", C.c_int32), ("ramp_iter", C.c_int32),
                ("cfl", C.c_double), ("vnn", C.c_double), ("gas_R", C.c_double), ("gas_cp", C.c_double),
                ("mu_ref", C.c_double), ("T_ref", C.c_double), ("S_suth", C.c_double), ("Pr", C.c_double),
                ("irs_beta", C.c_double), ("space_sd", C.c_int32), ("integration_vars", C.c_int32)]


class BlockDesc(C.Structure):
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("case_id", C.c_int32), ("reserved", C.c_int32),
                ("P", C.c_void_p), ("dl", C.c_void_p), ("M", C.c_void_p), ("sur", C.c_void_p),
                ("face_i", C.c_void_p), ("face_j", C.c_void_p), ("node", C.c_void_p),
                ("src_a", C.c_void_p), ("src_b", C.c_void_p)]


class BCEntry(C.Structure):
    _fields_ = [("b", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("f", C.c_int32), ("type", C.c_int32),
                ("bs", C.c_int32), ("is_", C.c_int32), ("js", C.c_int32), ("fs", C.c_int32),
                ("d11", C.c_int32), ("d12", C.c_int32), ("d21", C.c_int32), ("d22", C.c_int32),
                ("case_id", C.c_int32), ("reserved", C.c_int32 * 2),
                ("mach", C.c_double), ("T", C.c_double), ("p", C.c_double), ("rel_fac", C.c_double),
                ("alpha", C.c_double)]


class CaseDesc(C.Structure):
    _fields_ = [("scheme", Scheme), ("itermax", C.c_int32), ("reserved", C.c_int32),
                ("res_threshold", C.c_double), ("time_threshold", C.c_double), ("res_drop", C.c_double),
                ("time0", C.c_double), ("thermo", C.c_void_p), ("thermo_n", C.c_int32), ("reserved2", C.c_int32)]


class Runtime(C.Structure):
    _fields_ = [("cfl", C.c_double), ("vnn", C.c_double), ("ramp_iter", C.c_int32), ("itermax", C.c_int32),
                ("res_threshold", C.c_double), ("time_threshold", C.c_double), ("res_drop", C.c_double),
                ("limiter", C.c_int32), ("riemann", C.c_int32), ("space_sd", C.c_int32), ("reserved", C.c_int32)]


def runtime_struct(base, **changes):
    """A Runtime from `base` (a Runtime) with fields changed; limiter and riemann by name or id."""
    rt = Runtime.from_buffer_copy(base)
    for k, v in changes.items():
        if k == "limiter" and isinstance(v, str):
            v = LIMITERS[v]
        if k == "riemann" and isinstance(v, str):
            v = RIEMANN[v]
        if k == "space_sd":
            v = int(v)
        if not hasattr(rt, k):
            raise KeyError(f"{k} is not a runtime key")
        setattr(rt, k, v)
    return rt


class ProblemC(C.Structure):
    _fields_ = [("nblocks", C.c_int32), ("nbc", C.c_int32), ("ncases", C.c_int32), ("reserved", C.c_int32),
                ("blocks", C.POINTER(BlockDesc)), ("bc", C.POINTER(BCEntry)), ("cases", C.POINTER(CaseDesc))]


class StepOut(C.Structure):
    _fields_ = [("dtglobal", C.c_double), ("time", C.c_double), ("residual", C.c_double * 4),
                ("iter", C.c_int32), ("todo", C.c_int32), ("status", C.c_int32),
                ("fail_b", C.c_int32), ("fail_i", C.c_int32), ("fail_j", C.c_int32), ("reserved", C.c_int32 * 2)]

    def as_dict(self):
        return {"dtglobal": self.dtglobal, "time": self.time, "residual": list(self.residual), "iter": self.iter,
                "todo": self.todo, "status": self.status, "fail": (self.fail_b, self.fail_i, self.fail_j)}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def to_problem_c(problem):
    """Build the C description of a cases.Problem; returns (ProblemC, keepalive)."""
    nb = len(problem.blocks)
    blocks = (BlockDesc * nb)()
    for k, b in enumerate(problem.blocks):
        d = blocks[k]
        d.ni, d.nj, d.case_id = b.ni, b.nj, b.case_id
        d.P, d.dl, d.M, d.sur = _ptr(b.P), _ptr(b.geo.dl), _ptr(b.geo.M), _ptr(b.geo.sur)
        d.face_i, d.face_j, d.node = _ptr(b.geo.fi), _ptr(b.geo.fj), _ptr(b.geo.node)
        d.src_a, d.src_b = _ptr(b.src_a), _ptr(b.src_b)
    nbc = len(problem.bc)
    bcs = (BCEntry * max(1, nbc))()
    for k, r in enumerate(problem.bc):
        e = bcs[k]
        e.b, e.i, e.j, e.f, e.type = r["b"], r["i"], r["j"], r["f"], r["type"]
        e.bs, e.is_, e.js, e.fs = r.get("bs", 0), r.get("is", 0), r.get("js", 0), r.get("fs", 0)
        e.d11, e.d12, e.d21, e.d22 = r.get("d11", 1), r.get("d12", 0), r.get("d21", 0), r.get("d22", 1)
        e.case_id = r["case_id"]
        e.mach, e.T, e.p = r.get("mach", 0.0), r.get("T", 0.0), r.get("p", 0.0)
        e.rel_fac, e.alpha = r.get("rel_fac", 1.0), r.get("alpha", 0.0)
    nc = len(problem.cases)
    cases = (CaseDesc * nc)()
    for k, c in enumerate(problem.cases):
        s = cases[k].scheme
        s.n_rk, s.limiter, s.riemann = c.n_rk, LIMITERS[c.limiter], RIEMANN[c.riemann]
        s.time_accurate, s.model, s.source = int(c.time_accurate), c.model, SOURCES[c.source]
        s.precision, s.ramp_iter = PRECISION[c.precision], c.ramp_iter
        s.cfl, s.vnn, s.gas_R, s.gas_cp = c.cfl, c.vnn, c.gas_R, c.gas_cp
        s.mu_ref, s.T_ref, s.S_suth, s.Pr = c.mu_ref, c.T_ref, c.S_suth, c.Pr
        s.irs_beta = c.irs_beta
        s.space_sd = int(c.space_sd)
        s.integration_vars = INTVARS[c.integration_vars]
        cases[k].itermax, cases[k].res_threshold, cases[k].time_threshold = c.itermax, c.res_threshold, c.time_threshold
        cases[k].res_drop = c.res_drop
        cases[k].time0 = c.time0
        if c.thermo is not None:
            if c.thermo.dtype != np.float64 or c.thermo.ndim != 2 or c.thermo.shape[0] != 4 or \
                    not c.thermo.flags.c_contiguous:
                raise ValueError("CaseSpec.thermo must be a C-contiguous float64 array of shape (4, n)")
            cases[k].thermo, cases[k].thermo_n = _ptr(c.thermo), c.thermo.shape[1]
    pc = ProblemC()
    pc.nblocks, pc.nbc, pc.ncases = nb, nbc, nc
    pc.blocks = C.cast(blocks, C.POINTER(BlockDesc))
    pc.bc = C.cast(bcs, C.POINTER(BCEntry))
    pc.cases = C.cast(cases, C.POINTER(CaseDesc))
    return pc, (blocks, bcs, cases, problem)


_LIB = None
LIB_PATH = Path(os.environ.get("Q2D_GPU_LIB", Path(__file__).resolve().parent / "lib" / "libq2d_gpu.so"))

# exported symbols and their signatures (declared in include/q2d_gpu.h)
SIGNATURES = {
    "q2d_version": ([], C.c_char_p),
    "q2d_last_error": ([], C.c_char_p),
    "q2d_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "q2d_gpu_create": ([C.POINTER(ProblemC), C.POINTER(C.c_void_p)], C.c_int),
    "q2d_gpu_destroy": ([C.c_void_p], C.c_int),
    "q2d_gpu_add_cases": ([C.c_void_p, C.POINTER(ProblemC), C.POINTER(C.c_int32)], C.c_int),
    "q2d_gpu_remove_case": ([C.c_void_p, C.c_int32], C.c_int),
    "q2d_gpu_case_count": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)], C.c_int),
    "q2d_gpu_step": ([C.c_void_p, C.POINTER(StepOut)], C.c_int),
    "q2d_gpu_run": ([C.c_void_p, C.c_int32, C.POINTER(StepOut)], C.c_int),
    "q2d_gpu_capture": ([C.c_void_p], C.c_int),
    "q2d_gpu_step_noalloc": ([C.c_void_p, C.POINTER(StepOut)], C.c_int),
    "q2d_gpu_set_runtime": ([C.c_void_p, C.c_int32, C.POINTER(Runtime)], C.c_int),
    "q2d_gpu_set_fusion": ([C.c_void_p, C.c_int32], C.c_int),
    "q2d_gpu_get_runtime": ([C.c_void_p, C.c_int32, C.POINTER(Runtime)], C.c_int),
    "q2d_gpu_download_state": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "q2d_gpu_upload_state": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "q2d_gpu_download_dt": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "q2d_gpu_set_timing": ([C.c_void_p, C.c_int32], C.c_int),
    "q2d_gpu_timing_stats": ([C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)], C.c_int),
    "q2d_gpu_set_tiling": ([C.c_void_p, C.c_int32, C.c_int32], C.c_int),
    "q2d_gpu_get_tiling": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "q2d_gpu_info": ([C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "q2d_hbm_copy": ([C.c_int64, C.c_int32, C.POINTER(C.c_double)], C.c_int),
}


def hbm_copy_gbps(nbytes=1 << 30, reps=10):
    """Achieved bandwidth of the library's plain copy kernel: (read + write bytes) / time."""
    lib = load_library()
    ms = C.c_double(0.0)
    _check(lib, lib.q2d_hbm_copy(int(nbytes), int(reps), C.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def load_library(path=None):
    """Load the HIP library and declare every C-ABI signature. Raises if absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"HIP library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _LIB = lib
    return lib


def _check(lib, rc):
    if rc != Q2D_OK:
        raise Q2DError(rc, lib.q2d_last_error().decode())


class GpuSolver:
    """Device-resident explicit solver for a cases.Problem (all cases at once)."""

    def __init__(self, problem):
        self.lib = load_library()
        n = C.c_int(0)
        _check(self.lib, self.lib.q2d_device_count(C.byref(n)))
        if n.value < 1:
            raise Q2DError(-7, "no GPU")
        self.problem = problem
        pc, keep = to_problem_c(problem)
        h = C.c_void_p()
        _check(self.lib, self.lib.q2d_gpu_create(C.byref(pc), C.byref(h)))
        self.h = h
        self.ncases = len(problem.cases)
        self.blocks = list(problem.blocks)        # blocks by global index (grows with add_cases)
        self._out = (StepOut * max(1, self.ncases))()
        self._cases = [problem.case(k) for k in range(self.ncases)]   # per case, None once removed

    def add_cases(self, problem):
        """Append the cases of `problem` to the live solver; returns (first case index, first block index)."""
        pc, keep = to_problem_c(problem)
        first = C.c_int32(0)
        nb0 = len(self.blocks)
        _check(self.lib, self.lib.q2d_gpu_add_cases(self.h, C.byref(pc), C.byref(first)))
        self.blocks.extend(problem.blocks)
        self.ncases += len(problem.cases)
        self._cases.extend(problem.case(k) for k in range(len(problem.cases)))
        self._out = (StepOut * max(1, self.ncases))()
        return first.value, nb0

    def remove_case(self, case):
        _check(self.lib, self.lib.q2d_gpu_remove_case(self.h, int(case)))
        self._cases[int(case)] = None

    def live_problem(self):
        """The cases the handle holds now (added ones included, removed ones not), as one Problem."""
        from .cases import Problem
        pb = Problem()
        for c in self._cases:
            if c is not None:
                pb.add(c)
        return pb

    def device_bytes(self):
        nc, nb, by = C.c_int32(0), C.c_int32(0), C.c_int64(0)
        _check(self.lib, self.lib.q2d_gpu_case_count(self.h, C.byref(nc), C.byref(nb), C.byref(by)))
        return by.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.q2d_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, raise_on_error=True):
        rc = self.lib.q2d_gpu_step(self.h, self._out)
        if raise_on_error:
            _check(self.lib, rc)
        return [o.as_dict() for o in self._out[:self.ncases]]

    def step_noalloc(self, raise_on_error=True):
        """One step through the allocation-free entry point (file-scope device scratch)."""
        rc = self.lib.q2d_gpu_step_noalloc(self.h, self._out)
        if raise_on_error:
            _check(self.lib, rc)
        return [o.as_dict() for o in self._out[:self.ncases]]

    def set_fusion(self, visc=True):
        """Mixed laminar cases: viscous terms inside the stage kernel (True) or in k_visc (False)."""
        _check(self.lib, self.lib.q2d_gpu_set_fusion(self.h, int(bool(visc))))

    def runtime(self, case):
        """The runtime keys of `case` (a Runtime)."""
        rt = Runtime()
        _check(self.lib, self.lib.q2d_gpu_get_runtime(self.h, int(case), C.byref(rt)))
        return rt

    def set_runtime(self, case, **changes):
        """Change runtime keys of `case` between steps (input.ini re-read: cfl, vnn, ramp_iter, itermax,
        res_threshold, time_threshold, res_drop, limiter, riemann, space_sd)."""
        rt = runtime_struct(self.runtime(case), **changes)
        _check(self.lib, self.lib.q2d_gpu_set_runtime(self.h, int(case), C.byref(rt)))

    def run(self, nsteps, raise_on_error=True):
        rc = self.lib.q2d_gpu_run(self.h, int(nsteps), self._out)
        if raise_on_error:
            _check(self.lib, rc)
        return [o.as_dict() for o in self._out[:self.ncases]]

    def capture(self):
        """Capture the step graphs now (q2d_gpu_capture): the next run() spends no host time on capture."""
        _check(self.lib, self.lib.q2d_gpu_capture(self.h))

    def state(self, block):
        b = self.blocks[block]
        P = np.empty((b.nj + 4, b.ni + 4, 4))
        _check(self.lib, self.lib.q2d_gpu_download_state(self.h, block, P.ctypes.data_as(C.c_void_p)))
        return P

    def set_state(self, block, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        _check(self.lib, self.lib.q2d_gpu_upload_state(self.h, block, P.ctypes.data_as(C.c_void_p)))

    def dtlocal(self, block):
        b = self.blocks[block]
        d = np.empty((b.nj, b.ni))
        _check(self.lib, self.lib.q2d_gpu_download_dt(self.h, block, d.ctypes.data_as(C.c_void_p)))
        return d

    def set_timing(self, on):
        _check(self.lib, self.lib.q2d_gpu_set_timing(self.h, int(bool(on))))

    def timing_stats(self):
        n = C.c_int64(0)
        ms = C.c_double(0.0)
        _check(self.lib, self.lib.q2d_gpu_timing_stats(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def set_tiling(self, rows_per_item=0, waves_per_group=0):
        _check(self.lib, self.lib.q2d_gpu_set_tiling(self.h, int(rows_per_item), int(waves_per_group)))

    def autotune(self, steps=4, use_cache=True, verbose=False):
        """Pick the stage-kernel tiling for this problem (q2d_amd.autotune) and apply it."""
        from .autotune import tune
        r, w, _ = tune(self.live_problem(), steps=steps, use_cache=use_cache, verbose=verbose)
        self.set_tiling(r, w)
        return r, w

    def tiling(self):
        r = C.c_int32(0)
        w = C.c_int32(0)
        _check(self.lib, self.lib.q2d_gpu_get_tiling(self.h, C.byref(r), C.byref(w)))
        return r.value, w.value

    def info(self):
        c = C.c_int64(0)
        w = C.c_int64(0)
        _check(self.lib, self.lib.q2d_gpu_info(self.h, C.byref(c), C.byref(w)))
        return c.value, w.value
