"""ctypes binding of oracle/build/libq2d_oracle.so (test infrastructure)."""
from __future__ import annotations

import ctypes as C
import subprocess
from contextlib import contextmanager
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None


def lib_path():
    return HERE / "build" / "libq2d_oracle.so"


def build():
    """Compile the oracle (gcc, see oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return lib_path()


def load(path=None):
    """The oracle library (the portable build; `path`: another build of the same source, such as the
    CPU-baseline timing build)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else lib_path()
    if not p.exists():
        build()
    lib = C.CDLL(str(p))
    from q2d_amd.capi import ProblemC, StepOut
    lib.q2do_create.argtypes = [C.POINTER(ProblemC)]
    lib.q2do_create.restype = C.c_void_p
    lib.q2do_destroy.argtypes = [C.c_void_p]
    lib.q2do_step.argtypes = [C.c_void_p, C.POINTER(StepOut)]
    lib.q2do_step.restype = C.c_int
    for f in ("q2do_get_state", "q2do_set_state", "q2do_get_dt", "q2do_get_residual", "q2do_set_residual"):
        getattr(lib, f).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        getattr(lib, f).restype = C.c_int
    lib.q2do_sd_tramel.argtypes = [C.c_double] * 5
    lib.q2do_sd_tramel.restype = C.c_double
    lib.q2do_irs_smooth.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.q2do_irs_smooth.restype = C.c_int
    lib.q2do_set_threads.argtypes = [C.c_int]
    lib.q2do_set_threads.restype = C.c_int
    lib.q2do_get_threads.restype = C.c_int
    lib.q2do_limiter.argtypes = [C.c_int, C.c_double, C.c_double]
    lib.q2do_limiter.restype = C.c_double
    dp = C.POINTER(C.c_double)
    lib.q2do_reconstruct.argtypes = [C.c_int] + [C.c_double] * 10 + [dp, dp]
    lib.q2do_riemann.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, C.c_double, C.c_double, C.c_double, dp]
    lib.q2do_roe_averages.argtypes = [C.c_double, C.c_double, dp, dp, dp]
    lib.q2do_convective_flux.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp, C.c_double, C.c_double,
                                         C.c_double, dp, dp]
    lib.q2do_convective_flux.restype = C.c_int
    lib.q2do_pow.argtypes = [C.c_double, C.c_double]
    lib.q2do_pow.restype = C.c_double
    lib.q2do_plant.argtypes = [C.c_int]
    lib.q2do_plant.restype = C.c_int
    lib.q2do_bc_inflow_flux.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, dp, dp] + \
        [C.c_double] * 8 + [dp]
    lib.q2do_bc_inflow_flux.restype = C.c_int
    lib.q2do_bc_outflow_flux.argtypes = [C.c_double, C.c_double, C.c_int, dp] + [C.c_double] * 5 + [dp]
    lib.q2do_bc_outflow_flux.restype = C.c_int
    from q2d_amd.capi import Runtime
    lib.q2do_set_runtime.argtypes = [C.c_void_p, C.c_int, C.POINTER(Runtime)]
    lib.q2do_set_runtime.restype = C.c_int
    lib.q2do_unit_tables.argtypes = [dp, C.c_int]
    lib.q2do_thermo_eval.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.q2do_thermo_eval.restype = C.c_double
    lib.q2do_prim2cons.argtypes = [C.c_double, C.c_double, dp, dp]
    lib.q2do_cons2prim.argtypes = [C.c_double, C.c_double, dp, C.c_double, dp]
    if path is None:
        _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# planted mistakes (q2d_oracle.h Q2DO_PLANT_*): a test shows it detects each one
PLANTS = {"axi_sign": 1, "axi_energy": 2, "inflow_angle": 3, "inflow_rho": 4, "outflow_riem": 5,
          "outflow_isen": 6, "ramp": 7, "tab_node": 8, "tab_tofe": 9}


@contextmanager
def planted(name):
    """Run the oracle with one planted mistake (tests only)."""
    lib = load()
    old = lib.q2do_plant(PLANTS[name])
    try:
        yield
    finally:
        lib.q2do_plant(old)


_UNIT_TAB = None


@contextmanager
def unit_tables(tab):
    """Run the unit entry points (riemann, roe_averages, convective_flux, the BC fluxes, thermo) in table
    mode with `tab` ((4, n) planes h, cp, mu, k), or in the closed form for None."""
    global _UNIT_TAB
    lib = load()
    old = _UNIT_TAB
    _UNIT_TAB = None if tab is None else np.ascontiguousarray(tab, dtype=np.float64)
    lib.q2do_unit_tables(None if _UNIT_TAB is None else _dp(_UNIT_TAB), 0 if _UNIT_TAB is None else _UNIT_TAB.shape[1])
    try:
        yield
    finally:
        _UNIT_TAB = old
        lib.q2do_unit_tables(None if old is None else _dp(old), 0 if old is None else old.shape[1])


THERMO = {"h": 0, "cp": 1, "mu": 2, "k": 3, "gamma": 4, "e": 5, "T_of_e": 6, "sound": 7}


def thermo(what, R, cp, a, b=0.0):
    """One closure value (THERMO names), in the mode unit_tables set."""
    return load().q2do_thermo_eval(THERMO[what], R, cp, float(a), float(b))


def prim2cons(R, cp, prim):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    u = np.zeros(4)
    load().q2do_prim2cons(R, cp, _dp(prim), _dp(u))
    return u


def cons2prim(R, cp, cons, T0):
    cons = np.ascontiguousarray(cons, dtype=np.float64)
    q = np.zeros(4)
    load().q2do_cons2prim(R, cp, _dp(cons), float(T0), _dp(q))
    return q


def pow_(x, y):
    """exact_pow of the oracle (the `**` of the characteristic outflow)."""
    return load().q2do_pow(float(x), float(y))


def inflow_flux(limiter, riemann, R, cp, face, Bp, Ip, n1, n2, int_l, mach, T, p, rel_fac, alpha):
    """Supersonic_Inflow face flux (before the face sign); None where the reference falls back to symmetry."""
    Bp = np.ascontiguousarray(Bp, dtype=np.float64)
    Ip = np.ascontiguousarray(Ip, dtype=np.float64)
    F = np.zeros(4)
    rc = load().q2do_bc_inflow_flux(limiter, riemann, R, cp, face, _dp(Bp), _dp(Ip), n1, n2, int_l, mach, T, p,
                                    rel_fac, alpha, _dp(F))
    return None if rc else F


def outflow_flux(R, cp, face, Bp, n1, n2, int_l, p_exit, rel_fac):
    """Outflow face flux (before the face sign); None where the reference falls back to symmetry."""
    Bp = np.ascontiguousarray(Bp, dtype=np.float64)
    F = np.zeros(4)
    rc = load().q2do_bc_outflow_flux(R, cp, face, _dp(Bp), n1, n2, int_l, p_exit, rel_fac, _dp(F))
    return None if rc else F


class Oracle:
    """The CPU restatement run on a cases.Problem."""

    def __init__(self, problem, threads=None, lib=None):
        from q2d_amd.capi import StepOut, to_problem_c
        self.lib = lib if lib is not None else load()
        if threads:
            self.lib.q2do_set_threads(int(threads))
        self.problem = problem
        pc, self._keep = to_problem_c(problem)
        self.h = self.lib.q2do_create(C.byref(pc))
        if not self.h:
            raise RuntimeError("q2do_create failed")
        self._out = (StepOut * len(problem.cases))()

    def close(self):
        if self.h:
            self.lib.q2do_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self):
        rc = self.lib.q2do_step(self.h, self._out)
        return rc, [o.as_dict() for o in self._out]

    def set_runtime(self, case, **changes):
        """Change runtime keys of `case` between steps, as GpuSolver.set_runtime."""
        from q2d_amd.capi import LIMITERS, RIEMANN, Runtime, runtime_struct
        if not hasattr(self, "_rt"):
            self._rt = {}
        c = self.problem.cases[case]
        base = self._rt.get(case) or Runtime(
            cfl=c.cfl, vnn=c.vnn, ramp_iter=c.ramp_iter, itermax=c.itermax, res_threshold=c.res_threshold,
            time_threshold=c.time_threshold, res_drop=c.res_drop, limiter=LIMITERS[c.limiter],
            riemann=RIEMANN[c.riemann], space_sd=int(c.space_sd))
        rt = runtime_struct(base, **changes)
        self._rt[case] = rt
        if self.lib.q2do_set_runtime(self.h, int(case), C.byref(rt)) != 0:
            raise ValueError("q2do_set_runtime failed")

    def run(self, nsteps, stop_on_todo=True):
        outs = None
        for _ in range(nsteps):
            rc, outs = self.step()
            if rc != 0:
                return rc, outs
            if stop_on_todo and all(o["todo"] == 3 for o in outs):
                break
        return 0, outs

    def state(self, block):
        b = self.problem.blocks[block]
        P = np.empty((b.nj + 4, b.ni + 4, 4))
        self.lib.q2do_get_state(self.h, block, P.ctypes.data_as(C.c_void_p))
        return P

    def residual(self, block):
        """R(nprim, 1-gc:ni+gc, 1-gc:nj+gc) of a block as (nj+4, ni+4, 4)."""
        b = self.problem.blocks[block]
        R = np.empty((b.nj + 4, b.ni + 4, 4))
        self.lib.q2do_get_residual(self.h, block, R.ctypes.data_as(C.c_void_p))
        return R

    def set_residual(self, block, R):
        R = np.ascontiguousarray(R, dtype=np.float64)
        b = self.problem.blocks[block]
        assert R.shape == (b.nj + 4, b.ni + 4, 4)
        self.lib.q2do_set_residual(self.h, block, R.ctypes.data_as(C.c_void_p))

    def irs_smooth(self, case_id, beta):
        """Residual_Smoothing (Lib_IRS.f90:17-44) of one case on the residual arrays as set."""
        rc = self.lib.q2do_irs_smooth(self.h, int(case_id), float(beta))
        if rc != 0:
            raise RuntimeError(f"q2do_irs_smooth: {rc}")

    def dtlocal(self, block):
        b = self.problem.blocks[block]
        d = np.empty((b.nj, b.ni))
        self.lib.q2do_get_dt(self.h, block, d.ctypes.data_as(C.c_void_p))
        return d

    # unit functions
    def limiter(self, lid, x, y):
        return self.lib.q2do_limiter(lid, x, y)

    def riemann(self, rid, R, cp, L, Rs, nx, ny, beta=1.0):
        L = np.ascontiguousarray(L, dtype=np.float64)
        Rs = np.ascontiguousarray(Rs, dtype=np.float64)
        F = np.zeros(4)
        self.lib.q2do_riemann(rid, R, cp, _dp(L), _dp(Rs), beta, nx, ny, _dp(F))
        return F
