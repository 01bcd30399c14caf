"""Problem description and seeded workload generators.

A Problem is what the Fortran host hands to the C-ABI: blocks (state and
geometry in the reference's allocation layout), boundary records (the
domain%bc list read from bc.txt, IO_BC.f90:163-...) and one scheme per
independent case (the [Q2D-*] keys of input.ini).

Generators (all seeded, numpy Generator):
  riemann_planar     two-state Riemann problem on one block (512x8 by default)
  shu_osher          the reference's test/1D/ShuOsher-LIMO3 case (N=400, Air table)
  planar_shock       angled planar shock into a perturbed gas, 2x2 blocks
  batch_uneven       many independent planar / quasi-2D cases of uneven size
  ramp_channel       supersonic compression ramp (steady, local dt)
  axisymmetric_pipe  laminar axisymmetric Navier-Stokes, supersonic pipe inflow, isothermal wall
  laminar_area_channel  quasi-2D laminar Navier-Stokes with area variation (Definition A)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from .geometry import BlockGeometry, area_geometry, axisymmetric_geometry, ghost_metrics

GC = 2

# Air table of the reference tests (test/common/Air: W = 28.970418, cp = 1004.5).  R_u = 8314.51 is the
# universal gas constant that reproduces the reference binary's Shu-Osher baselines to their 10 printed digits
# (test/1D/ShuOsher-LIMO3/reference/baseline_*.dat); docs/theory/thermo.md:56 quotes 8314.46.
AIR_R = 8314.51 / 28.970418
AIR_CP = 1004.5


@dataclass
class CaseSpec:
    n_rk: int = 2
    limiter: str = "minmod"
    riemann: str = "HLLC"
    time_accurate: bool = True
    model: int = 0
    source: str = "none"
    precision: str = "fp64"
    ramp_iter: int = 0
    cfl: float = 0.5
    vnn: float = 0.3
    gas_R: float = 1.0
    gas_cp: float = 3.5
    mu_ref: float = 0.0
    T_ref: float = 1.0
    S_suth: float = 0.0
    Pr: float = 0.72
    itermax: int = 1 << 30
    res_threshold: float = -1.0
    time_threshold: float = 1e300
    res_drop: float = 0.0
    irs_beta: float = 0.0         # irs-beta: > 0 enables implicit residual smoothing
    space_sd: bool = False        # space-reconstruction = MUSCL-SD (Tramel's shock detector on the slopes)
    integration_vars: str = "cons"  # integration-variables ('prim' is rejected by the library)
    time0: float = 0.0            # domain%time at the start: 0 (new run), SOLUTIONTIME (restart), -1 (steady)
    thermo: object = None         # table mode: (4, n) float64 planes h, cp, mu, k on T = 0..n-1 K (q2d_amd.thermo)


@dataclass
class Block:
    geo: BlockGeometry
    P: np.ndarray                 # (nj+4, ni+4, 4): P(nprim, 1-gc:ni+gc, 1-gc:nj+gc)
    case_id: int = 0
    src_a: np.ndarray | None = None
    src_b: np.ndarray | None = None

    @property
    def ni(self):
        return self.geo.ni

    @property
    def nj(self):
        return self.geo.nj

    def interior(self):
        return self.P[2:-2, 2:-2]


@dataclass
class Problem:
    blocks: list = field(default_factory=list)
    bc: list = field(default_factory=list)
    cases: list = field(default_factory=list)
    meta: dict = field(default_factory=dict)

    def ncells(self):
        return sum(b.ni * b.nj for b in self.blocks)

    def case(self, k):
        """Case k alone, as a one-case Problem (its blocks, records and scheme)."""
        blocks = [replace(b, case_id=0) for b in self.blocks if b.case_id == k]
        bc = []
        for r in self.bc:
            if r["case_id"] == k:
                r2 = dict(r)
                r2["case_id"] = 0
                bc.append(r2)
        return Problem(blocks=blocks, bc=bc, cases=[self.cases[k]], meta={"from_case": k})

    def add(self, other: "Problem"):
        """Append the cases of another problem (batching)."""
        off = len(self.cases)
        for b in other.blocks:
            self.blocks.append(replace(b, case_id=b.case_id + off))
        for r in other.bc:
            r2 = dict(r)
            r2["case_id"] = r["case_id"] + off
            self.bc.append(r2)
        self.cases.extend(other.cases)
        return self


def cartesian_nodes(ni, nj, x0, x1, y0, y1):
    xs = x0 + (x1 - x0) * np.arange(ni + 1) / ni
    ys = y0 + (y1 - y0) * np.arange(nj + 1) / nj
    node = np.empty((nj + 1, ni + 1, 2))
    node[..., 0] = xs[None, :]
    node[..., 1] = ys[:, None]
    return node


def face_records(b, ni, nj, types, case_id=0, props=None):
    """Records of one block, ordered face 1..4 (IO_BC layout). types: {face: code}."""
    props = props or {}
    recs = []
    for f in (1, 2, 3, 4):
        t = types[f]
        cells = [(1 if f == 1 else ni, j) for j in range(1, nj + 1)] if f <= 2 else \
                [(i, 1 if f == 3 else nj) for i in range(1, ni + 1)]
        for (i, j) in cells:
            r = {"b": b, "i": i, "j": j, "f": f, "type": t, "case_id": case_id}
            r.update(props.get(f, {}))
            recs.append(r)
    return recs


def fill_initial_ghosts(P):
    """Ghost layers of an initial state: zero-gradient copy (overwritten by the first ghost fill)."""
    P[:2, :] = P[2:3, :]
    P[-2:, :] = P[-3:-2, :]
    P[:, :2] = P[:, 2:3]
    P[:, -2:] = P[:, -3:-2]
    return P


def finish(problem, geoms):
    """Ghost metrics from the boundary records, per case."""
    by_case = {}
    for k, b in enumerate(problem.blocks):
        by_case.setdefault(b.case_id, []).append(k)
    for c, ks in by_case.items():
        recs = [r for r in problem.bc if r["case_id"] == c]
        ghost_metrics([problem.blocks[k].geo for k in ks], recs)
    return problem


# ---------------------------------------------------------------------------
def riemann_planar(seed=0, ni=512, nj=8, steps=1000, limiter="minmod", riemann="HLLC", n_rk=2, cfl=0.5,
                   precision="fp64", bc_y=400, length=1.0):
    """Two-state Riemann problem on x in [0,length], diaphragm at x = length/2.

    rho_L, p_L, rho_R, p_R ~ U[0.1, 1] drawn in that order from default_rng(seed);
    u = v = 0, gamma = 1.4 (R = 1, cp = 3.5); square cells; BC 400 on all faces
    ('transmissive'); time-accurate global dt.
    """
    rng = np.random.default_rng(seed)
    rl, pl, rr, pr = rng.uniform(0.1, 1.0, 4)
    node = cartesian_nodes(ni, nj, 0.0, length, 0.0, length * nj / ni)
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    xc = 0.5 * (node[0, 1:, 0] + node[0, :-1, 0])
    left = xc < 0.5 * length
    P[2:-2, 2:-2, 0] = np.where(left, rl, rr)[None, :]
    P[2:-2, 2:-2, 3] = np.where(left, pl, pr)[None, :]
    fill_initial_ghosts(P)
    spec = CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, cfl=cfl, precision=precision, itermax=steps)
    pb = Problem(blocks=[Block(geo, P)], bc=face_records(1, ni, nj, {1: 400, 2: 400, 3: bc_y, 4: bc_y}),
                 cases=[spec], meta={"name": "riemann_planar", "seed": seed, "states": (rl, pl, rr, pr)})
    return finish(pb, [geo])


def shu_osher(n=400, limiter="minmod"):
    """test/1D/ShuOsher-LIMO3 (gen_case.py): x in [-5,5], one cell row, BC 400 (i) / 300 (j),
    HLLC, RK3, CFL 0.5, time-accurate to t = 1.8, Air table."""
    dx = (5.0 - (-5.0)) / n
    xn = np.array([-5.0 + i * dx for i in range(n + 1)])
    xn = np.array([float(f"{v: .15E}") for v in xn])            # ic.tec holds the nodes in %.15E
    node = np.empty((2, n + 1, 2))
    node[..., 0] = xn[None, :]
    node[0, :, 1] = 0.0
    node[1, :, 1] = 1.0
    geo = BlockGeometry(node)
    xc = [0.5 * (xn[i] + xn[i + 1]) for i in range(n)]
    P = np.zeros((1 + 4, n + 4, 4))
    for i, x in enumerate(xc):
        st = (3.857143, 2.629369, 0.0, 10.33333) if x < -4.0 else (1.0 + 0.2 * math.sin(5.0 * x), 0.0, 0.0, 1.0)
        P[2, i + 2, :] = [float(f"{v: .15E}") for v in st]
    fill_initial_ghosts(P)
    spec = CaseSpec(n_rk=3, limiter=limiter, riemann="HLLC", cfl=0.5, gas_R=AIR_R, gas_cp=AIR_CP,
                    time_threshold=1.8, itermax=1 << 30)
    pb = Problem(blocks=[Block(geo, P)], bc=face_records(1, n, 1, {1: 400, 2: 400, 3: 300, 4: 300}),
                 cases=[spec], meta={"name": "shu_osher", "n": n, "xc": np.array(xc)})
    return finish(pb, [geo])


def _rh_post_shock(Ms, gam, rho1, p1):
    a1 = math.sqrt(gam * p1 / rho1)
    r = (gam + 1.0) * Ms * Ms / ((gam - 1.0) * Ms * Ms + 2.0)
    rho2 = rho1 * r
    p2 = p1 * (1.0 + 2.0 * gam / (gam + 1.0) * (Ms * Ms - 1.0))
    u2 = Ms * a1 * (1.0 - 1.0 / r)
    return rho2, u2, p2


def planar_shock(seed=0, n=1024, steps=500, limiter="vanleer", riemann="HLLC", n_rk=3, cfl=0.5, precision="fp64",
                 layout=(2, 2)):
    """Angled planar shock into a perturbed gas on [0,2]^2 split into 2x2 blocks of n x n cells.

    theta ~ U[-15, 15] deg, Ms ~ U[1.5, 3], pre-shock rho = 1*(1 + 0.05*sum_k sin(2 pi (kx x + ky y) + phi_k)/8),
    p = 1, u = 0, gamma = 1.4 (R = 1); shock front through (0.5, 1).  Blocks joined by BC 101, y-periodic by
    BC 201, x-ends BC 400.  Time-accurate.
    """
    rng = np.random.default_rng(seed)
    theta = math.radians(rng.uniform(-15.0, 15.0))
    Ms = rng.uniform(1.5, 3.0)
    kx = rng.integers(1, 5, 8)
    ky = rng.integers(1, 5, 8)
    ph = rng.uniform(0.0, 2.0 * math.pi, 8)
    gam = 1.4
    rho2, u2, p2 = _rh_post_shock(Ms, gam, 1.0, 1.0)
    nbx, nby = layout
    L = 2.0
    hx, hy = L / nbx, L / nby
    blocks, geoms, recs = [], [], []
    bid = lambda bx, by: by * nbx + bx + 1
    for by in range(nby):
        for bx in range(nbx):
            node = cartesian_nodes(n, n, bx * hx, (bx + 1) * hx, by * hy, (by + 1) * hy)
            geo = BlockGeometry(node)
            xc = 0.25 * (node[1:, 1:, 0] + node[1:, :-1, 0] + node[:-1, 1:, 0] + node[:-1, :-1, 0])
            yc = 0.25 * (node[1:, 1:, 1] + node[1:, :-1, 1] + node[:-1, 1:, 1] + node[:-1, :-1, 1])
            d = (xc - 0.5) * math.cos(theta) + (yc - 1.0) * math.sin(theta)
            pert = np.zeros_like(xc)
            for k in range(8):
                pert += np.sin(2.0 * math.pi * (kx[k] * xc + ky[k] * yc) + ph[k])
            rho1 = 1.0 * (1.0 + 0.05 * pert / 8.0)
            P = np.zeros((n + 4, n + 4, 4))
            post = d < 0.0
            P[2:-2, 2:-2, 0] = np.where(post, rho2, rho1)
            P[2:-2, 2:-2, 1] = np.where(post, u2 * math.cos(theta), 0.0)
            P[2:-2, 2:-2, 2] = np.where(post, u2 * math.sin(theta), 0.0)
            P[2:-2, 2:-2, 3] = np.where(post, p2, 1.0)
            fill_initial_ghosts(P)
            blocks.append(Block(geo, P))
            geoms.append(geo)
    for by in range(nby):
        for bx in range(nbx):
            b = bid(bx, by)
            # face 1 / 2 : x neighbours or extrapolation
            for f in (1, 2, 3, 4):
                if f <= 2:
                    nbx_ = bx - 1 if f == 1 else bx + 1
                    for j in range(1, n + 1):
                        r = {"b": b, "i": 1 if f == 1 else n, "j": j, "f": f, "case_id": 0}
                        if 0 <= nbx_ < nbx:
                            r.update(type=101, bs=bid(nbx_, by), **{"is": n if f == 1 else 1}, js=j, fs=2 if f == 1 else 1)
                        else:
                            r.update(type=400)
                        recs.append(r)
                else:
                    nby_ = by - 1 if f == 3 else by + 1
                    periodic = not (0 <= nby_ < nby)
                    nby_ %= nby
                    for i in range(1, n + 1):
                        r = {"b": b, "i": i, "j": 1 if f == 3 else n, "f": f, "case_id": 0,
                             "type": 201 if periodic else 101, "bs": bid(bx, nby_), "is": i,
                             "js": n if f == 3 else 1, "fs": 4 if f == 3 else 3}
                        recs.append(r)
    spec = CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, cfl=cfl, precision=precision, itermax=steps)
    pb = Problem(blocks=blocks, bc=recs, cases=[spec],
                 meta={"name": "planar_shock", "seed": seed, "theta": theta, "Ms": Ms})
    return finish(pb, geoms)


def _round8(v):
    return max(8, int(round(v / 8.0)) * 8)


def uniform_inflow(ni, nj, mach, riemann="HLLC", area_a=None, precision="mixed", n_rk=2, steps=200,
                   limiter="vanleer", length=1.0):
    """Uniform inflow at Mach `mach` along x: BC 405 (face 1), 400 (face 2), 300 (faces 3, 4).
    With area_a, quasi-2D area A(x) = 1 + a sin^2(pi x) (nodal Cross_Section, area-variation source)."""
    node = cartesian_nodes(ni, nj, 0.0, length, 0.0, length * nj / ni)
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    gam = 1.4
    a = math.sqrt(gam * 1.0 / 1.0)
    P[2:-2, 2:-2] = [1.0, mach * a, 0.0, 1.0]
    fill_initial_ghosts(P)
    props = {1: {"mach": mach, "T": 1.0, "p": 1.0, "rel_fac": 1.0, "alpha": 0.0}}
    recs = face_records(1, ni, nj, {1: 405, 2: 400, 3: 300, 4: 300}, props=props)
    spec = CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, precision=precision, itermax=steps)
    blk = Block(geo, P)
    if area_a is not None:
        spec.source = "area_variation"
        A = 1.0 + area_a * np.sin(math.pi * node[..., 0]) ** 2
        blk.src_a, blk.src_b = None, None
        pb = Problem(blocks=[blk], bc=recs, cases=[spec], meta={"name": "uniform_inflow", "area_a": area_a})
        finish(pb, [geo])
        blk.src_a, blk.src_b = area_geometry(geo, A)
        return pb
    pb = Problem(blocks=[blk], bc=recs, cases=[spec], meta={"name": "uniform_inflow"})
    return finish(pb, [geo])


def batch_uneven(seed=0, ncases=64, steps=200, precision="mixed", smin=0.0, smax=4.0, n_rk=2):
    """Independent planar / quasi-2D cases with log-uniform sizes 128x64 .. 2048x1024 (s ~ U[smin, smax]:
    ni = round8(128*2^s), nj = round8(64*2^s)).  Half (seeded) are quasi-2D with A = 1 + a sin^2(pi x),
    a ~ U[0.2, 1].  Initial state: Riemann (as riemann_planar) or uniform inflow at Mach ~ U[0.5, 3].
    HLLE or HLLC per case."""
    rng = np.random.default_rng(seed)
    pb = Problem(meta={"name": "batch_uneven", "seed": seed, "cases": []})
    for c in range(ncases):
        s = rng.uniform(smin, smax)
        ni, nj = _round8(128 * 2 ** s), _round8(64 * 2 ** s)
        quasi = rng.uniform() < 0.5
        kind = "riemann" if rng.uniform() < 0.5 else "inflow"
        riemann = "HLLE" if rng.uniform() < 0.5 else "HLLC"
        mach = rng.uniform(0.5, 3.0)
        a = rng.uniform(0.2, 1.0)
        sub_seed = int(rng.integers(0, 2 ** 31))
        if kind == "riemann" and not quasi:
            sub = riemann_planar(seed=sub_seed, ni=ni, nj=nj, steps=steps, limiter="vanleer", riemann=riemann,
                                 n_rk=n_rk, precision=precision, bc_y=300)
        else:
            sub = uniform_inflow(ni, nj, max(mach, 1.05) if kind == "inflow" else 1.5, riemann=riemann,
                                 area_a=a if quasi else None, precision=precision, n_rk=n_rk, steps=steps)
        sub.cases[0].itermax = steps
        pb.add(sub)
        pb.meta["cases"].append({"ni": ni, "nj": nj, "quasi": quasi, "kind": kind, "riemann": riemann})
    return pb


def ramp_channel(ni=128, nj=64, wedge_deg=10.0, mach=2.5, x_ramp=0.2, length=1.0, height=0.5, precision="fp64",
                 n_rk=3, limiter="minmod", riemann="HLLE", cfl=0.5, res_drop=1e-6, itermax=200000):
    """Supersonic compression ramp: wall at y = 0 (BC 300) turning by `wedge_deg` at x = x_ramp, inflow BC 405
    (face 1), outflow BC 400 (face 2), top BC 400.  The mesh is sheared to follow the wall.  Local dt."""
    xs = length * np.arange(ni + 1) / ni
    t = math.tan(math.radians(wedge_deg))
    ywall = np.where(xs > x_ramp, (xs - x_ramp) * t, 0.0)
    node = np.empty((nj + 1, ni + 1, 2))
    eta = np.arange(nj + 1) / nj
    node[..., 0] = xs[None, :]
    node[..., 1] = ywall[None, :] + (height - ywall[None, :]) * eta[:, None]
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    gam = 1.4
    a = math.sqrt(gam)
    P[2:-2, 2:-2] = [1.0, mach * a, 0.0, 1.0]
    fill_initial_ghosts(P)
    props = {1: {"mach": mach, "T": 1.0, "p": 1.0, "rel_fac": 1.0, "alpha": 0.0}}
    recs = face_records(1, ni, nj, {1: 405, 2: 400, 3: 300, 4: 400}, props=props)
    spec = CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, time_accurate=False, cfl=cfl, precision=precision,
                    itermax=itermax)
    pb = Problem(blocks=[Block(geo, P)], bc=recs, cases=[spec],
                 meta={"name": "ramp_channel", "wedge_deg": wedge_deg, "mach": mach, "res_drop": res_drop})
    return finish(pb, [geo])


def _tanh_wall_cluster(nj, first):
    """Radial node coordinates on [0, 1] clustered at eta = 1 by a tanh law, first cell (at eta = 1) of size
    `first`:  y(eta) = tanh(beta eta) / tanh(beta), beta solved by bisection."""
    def first_cell(beta):
        return 1.0 - math.tanh(beta * (1.0 - 1.0 / nj)) / math.tanh(beta)
    if first >= 1.0 / nj:
        return np.arange(nj + 1) / nj
    lo, hi = 1e-6, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first_cell(mid) > first:
            lo = mid
        else:
            hi = mid
    beta = 0.5 * (lo + hi)
    eta = np.arange(nj + 1) / nj
    y = np.tanh(beta * eta) / math.tanh(beta)
    y[0], y[-1] = 0.0, 1.0
    return y


def axisymmetric_pipe(seed=0, ni=2048, nj=512, first_cell=1e-5, reynolds=1e5, mach=None, radius=0.25,
                      T_wall=1.0, precision="mixed", n_rk=3, limiter="vanleer", riemann="HLLC", cfl=0.5,
                      vnn=0.2, steps=2000):
    """Axisymmetric laminar Navier-Stokes: supersonic inflow (BC 405, face 1) into a pipe of length 1 and
    radius `radius`; axis r = 0 on j = 0 (BC 300, face 3), isothermal no-slip wall on r = radius (BC 302,
    face 4), supersonic outflow (BC 400, face 2).  Radial tanh clustering at the wall with a first cell of
    `first_cell` (lengths in units of the pipe length L = 1).  Non-dimensional gas R = 1, cp = 3.5 at
    rho = p = T = 1; Sutherland viscosity with S / T_ref = 110.4 / 273.15, Pr = 0.72, mu_inf from Re_L;
    wall at T_wall (default the free-stream temperature).  Mach ~ U[1.5, 3] from the seed unless given.
    Local time step, CFL 0.5, SSP-RK3, VNN 0.2.  Two departures from the reference defaults, both
    for stability of this explicit scheme: (1) the CFL limit is taken per direction (Mod_dt.f90:76-84),
    so on the near-square cells of the stretched mesh the 2D Courant number reaches 2 x 0.5 = 1, which
    SSP-RK2 with MUSCL does not survive (it fails after ~350 steps at 2048x512); SSP-RK3 does.
    (2) VNN 0.2 rather than 0.3 (Register_Numerics.f90:35): the VNN limit uses the kinematic
    viscosity while heat conduction diffuses gamma/Pr = 1.94 times faster, so 0.3 puts the wall cells
    past the diffusive stability limit and the step fails at the inflow/wall corner."""
    rng = np.random.default_rng(seed)
    if mach is None:
        mach = float(rng.uniform(1.5, 3.0))
    gam = 1.4
    a = math.sqrt(gam)
    U = mach * a
    mu_inf = U * 1.0 / reynolds
    xs = np.arange(ni + 1) / ni
    rs = radius * _tanh_wall_cluster(nj, first_cell / radius)
    node = np.empty((nj + 1, ni + 1, 2))
    node[..., 0] = xs[None, :]
    node[..., 1] = rs[:, None]
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    P[2:-2, 2:-2] = [1.0, U, 0.0, 1.0]
    fill_initial_ghosts(P)
    props = {1: {"mach": mach, "T": 1.0, "p": 1.0, "rel_fac": 1.0, "alpha": 0.0},
             4: {"T": T_wall}}
    recs = face_records(1, ni, nj, {1: 405, 2: 400, 3: 300, 4: 302}, props=props)
    spec = CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, time_accurate=False, model=1,
                    source="axisymmetric", precision=precision, cfl=cfl, vnn=vnn, mu_ref=mu_inf, T_ref=1.0,
                    S_suth=110.4 / 273.15, Pr=0.72, itermax=steps)
    blk = Block(geo, P)
    pb = Problem(blocks=[blk], bc=recs, cases=[spec],
                 meta={"name": "axisymmetric_pipe", "mach": mach, "reynolds": reynolds, "T_wall": T_wall})
    finish(pb, [geo])
    blk.src_a, blk.src_b = axisymmetric_geometry(geo)
    return pb


def laminar_area_channel(seed=0, ni=128, nj=40, area_a=0.5, first_cell=2e-4, reynolds=2e4, mach=None, height=0.25,
                         T_wall=1.0, precision="fp64", n_rk=3, limiter="vanleer", riemann="HLLC", cfl=0.5, vnn=0.2,
                         steps=200):
    """Quasi-2D laminar Navier-Stokes with area variation (Definition A): the planar half-channel of
    axisymmetric_pipe (supersonic inflow BC 405, outflow BC 400, symmetry plane BC 300 on y = 0, isothermal
    no-slip wall BC 302 on y = height, tanh clustering at the wall) with the nodal cross-section
    A(x) = 1 + a sin^2(pi x) of the reference's quasi-2D cases (INPUT/block<N>_area.dat layout,
    Lib_Source_Areavariation.f90:53-101).  Local dt, SSP-RK3, VNN 0.2 as in axisymmetric_pipe."""
    rng = np.random.default_rng(seed)
    if mach is None:
        mach = float(rng.uniform(1.5, 3.0))
    U = mach * math.sqrt(1.4)
    xs = np.arange(ni + 1) / ni
    ys = height * _tanh_wall_cluster(nj, first_cell / height)
    node = np.empty((nj + 1, ni + 1, 2))
    node[..., 0] = xs[None, :]
    node[..., 1] = ys[:, None]
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    P[2:-2, 2:-2] = [1.0, U, 0.0, 1.0]
    fill_initial_ghosts(P)
    props = {1: {"mach": mach, "T": 1.0, "p": 1.0, "rel_fac": 1.0, "alpha": 0.0}, 4: {"T": T_wall}}
    recs = face_records(1, ni, nj, {1: 405, 2: 400, 3: 300, 4: 302}, props=props)
    spec = CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, time_accurate=False, model=1,
                    source="area_variation", precision=precision, cfl=cfl, vnn=vnn, mu_ref=U / reynolds, T_ref=1.0,
                    S_suth=110.4 / 273.15, Pr=0.72, itermax=steps)
    blk = Block(geo, P)
    pb = Problem(blocks=[blk], bc=recs, cases=[spec],
                 meta={"name": "laminar_area_channel", "mach": mach, "reynolds": reynolds, "area_a": area_a})
    finish(pb, [geo])
    A = 1.0 + area_a * np.sin(math.pi * node[..., 0]) ** 2
    blk.src_a, blk.src_b = area_geometry(geo, A)
    return pb


def noh_cylindrical(nj=1000, ni=4, r_max=2.5, rho0=1.0, u0=1.0, p0=1e-9, t_final=1.0, precision="fp64"):
    """test/2D/euler/noh-cylindrical (gen_case.py, input.ini): cylindrical Noh implosion on an (x, r) strip,
    r in [0, r_max] resolved by nj cells, ni cells in x with dx = dr; rho = rho0, u = 0, v = -u0, p = p0.
    BC 300 on faces 1, 2 (x-invariance) and 3 (the axis), 400 on face 4.  First-order reconstruction,
    HLLE, SSP-RK3, CFL 0.5, time-accurate to time-threshold t_final, axisymmetric source, Air table gas."""
    dr = r_max / nj
    node = cartesian_nodes(ni, nj, 0.0, ni * dr, 0.0, r_max)
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    P[2:-2, 2:-2] = [rho0, 0.0, -u0, p0]
    fill_initial_ghosts(P)
    spec = CaseSpec(n_rk=3, limiter="first-order", riemann="HLLE", cfl=0.5, gas_R=AIR_R, gas_cp=AIR_CP,
                    source="axisymmetric", precision=precision, time_threshold=t_final, itermax=1 << 30)
    blk = Block(geo, P)
    pb = Problem(blocks=[blk], bc=face_records(1, ni, nj, {1: 300, 2: 300, 3: 300, 4: 400}), cases=[spec],
                 meta={"name": "noh_cylindrical", "rho0": rho0, "u0": u0, "p0": p0, "t_final": t_final})
    finish(pb, [geo])
    blk.src_a, blk.src_b = axisymmetric_geometry(geo)
    return pb


# The reference's 1D refinement ladder (test/1D/<case>/input.ini and INPUT/ic_x<L>.tec, MESH/mesh_x<L>.tec):
# x in [0, 1] with 100 L cells, one row of height 0.01, left / right states split at x0, Air table gas,
# SSP-RK3, time-accurate to time-threshold.  States as the ic files hold them (%.15E).
LADDER_1D = {
    "Ein91": dict(x0=0.5, t=0.15, left=(0.999999980682838, -2.0, 0.0, 0.4), right=(0.999999980682838, 2.0, 0.0, 0.4),
                  limiter="vanleer", riemann="HLLE", cfl=0.9, bc=(400, 400, 400, 400)),
    "Tor99": dict(x0=0.3, t=0.20, left=(0.999998975822492, 0.75, 0.0, 1.0), right=(0.124999889921731, 0.0, 0.0, 0.1),
                  limiter="vanleer", riemann="HLLE++", cfl=0.9, bc=(400, 400, 300, 300)),
    "Noh87": dict(x0=0.5, t=1.0, left=(1.0, 1.0, 0.0, 1e-15), right=(1.0, -1.0, 0.0, 1e-15),
                  limiter="first-order", riemann="HLLE", cfl=0.5, bc=(400, 400, 400, 400)),
    "Sod79": dict(x0=0.5, t=0.20, left=(0.999998975822492, 0.0, 0.0, 1.0), right=(0.124999889921731, 0.0, 0.0, 0.1),
                  limiter="mc", riemann="HLLC", cfl=0.9, bc=(300, 300, 300, 300)),
}


def ladder_1d(name, level=1, precision="fp64"):
    """One level of the reference's 1D ladder (LADDER_1D).  Sod79's reference solver is SLAU, which this
    path does not build: it runs here with HLLC."""
    c = LADDER_1D[name]
    n = 100 * level
    xn = np.array([float(f"{i / n: .15E}") for i in range(n + 1)])
    node = np.empty((2, n + 1, 2))
    node[..., 0] = xn[None, :]
    node[0, :, 1] = 0.0
    node[1, :, 1] = 0.01
    geo = BlockGeometry(node)
    xc = 0.5 * (xn[:-1] + xn[1:])
    P = np.zeros((5, n + 4, 4))
    P[2, 2:-2] = np.where((xc < c["x0"])[:, None], np.array(c["left"])[None, :], np.array(c["right"])[None, :])
    fill_initial_ghosts(P)
    spec = CaseSpec(n_rk=3, limiter=c["limiter"], riemann=c["riemann"], cfl=c["cfl"], gas_R=AIR_R, gas_cp=AIR_CP,
                    precision=precision, time_threshold=c["t"], itermax=1 << 30)
    b1, b2, b3, b4 = c["bc"]
    pb = Problem(blocks=[Block(geo, P)], bc=face_records(1, n, 1, {1: b1, 2: b2, 3: b3, 4: b4}), cases=[spec],
                 meta={"name": name, "level": level, "xc": xc})
    return finish(pb, [geo])


def outflow_channel(ni=96, nj=32, mach=0.5, p_exit=1.2, rel_fac=0.5, wedge_deg=6.0, backflow_cols=0,
                    precision="fp64", n_rk=3, limiter="vanleer", riemann="HLLC", time_accurate=False, steps=100):
    """A channel with the characteristic outflow BC 406 (exit pressure p_exit, relaxation rel_fac) on face 2:
    the sheared mesh of ramp_channel (so exit faces are not aligned with x), BC 405 imposing the
    stream at Mach `mach` on face 1, BC 300 on face 3, BC 400 on face 4.  Uniform stream, except that the
    last `backflow_cols` columns flow back into the domain (BC 406 then falls back to BC_Symmetry_Eul)."""
    pb = ramp_channel(ni=ni, nj=nj, wedge_deg=wedge_deg, mach=mach, precision=precision, n_rk=n_rk,
                      limiter=limiter, riemann=riemann, itermax=steps)
    pb.cases[0].time_accurate = time_accurate
    for r in pb.bc:
        if r["f"] == 2:
            r.update(type=406, p=p_exit, rel_fac=rel_fac)
    if backflow_cols:
        P = pb.blocks[0].P
        P[2:-2, -2 - backflow_cols:-2, 1] *= -0.5
    pb.meta.update(name="outflow_channel", p_exit=p_exit, rel_fac=rel_fac)
    return pb
