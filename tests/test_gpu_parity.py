"""HIP path against the oracle, through the C-ABI (needs the GPU)."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu


def rel_linf_cons(Pg, Po, R, cp):
    """Relative L-inf per conserved variable (rho, rho u, rho v, rho E): max|d| / max|ref|, 0/0 = 0."""
    def cons(P):
        rho, u, v, p = P[..., 0], P[..., 1], P[..., 2], P[..., 3]
        cv = cp - R
        E = cv * (p / (rho * R)) + 0.5 * (u * u + v * v)
        return np.stack([rho, rho * u, rho * v, rho * E], -1)
    a, b = cons(Pg), cons(Po)
    out = []
    for k in range(4):
        m = np.abs(b[..., k]).max()
        d = np.abs(a[..., k] - b[..., k]).max()
        out.append(0.0 if d == 0 else (d / m if m > 0 else np.inf))
    return out


def run_both(oracle_mod, pb, steps, graph=True):
    g = GpuSolver(pb)
    go = g.run(steps) if graph else [g.step() for _ in range(steps)][-1]
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0
    return g, go, o, oo


def test_c1_riemann_1000_steps(oracle_mod):
    """512x8 two-state Riemann problem, minmod/HLLC/SSP-RK2, 1000 steps: bit-identical to the oracle."""
    pb = cases.riemann_planar(seed=0)
    g, go, o, oo = run_both(oracle_mod, pb, 1000)
    Pg, Po = g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2]
    errs = rel_linf_cons(Pg, Po, 1.0, 3.5)
    assert max(errs) <= 1e-12, errs
    assert np.array_equal(Pg, Po), np.abs(Pg - Po).max()
    assert go[0]["iter"] == oo[0]["iter"] == 1000
    assert go[0]["dtglobal"] == oo[0]["dtglobal"]
    assert go[0]["time"] == oo[0]["time"]
    for k in range(4):
        assert abs(go[0]["residual"][k] - oo[0]["residual"][k]) <= 1e-12 * max(1e-300, abs(oo[0]["residual"][k]))


@pytest.mark.parametrize("lim", ["iord", "minmod", "vanleer", "vanalbada", "mc", "superbee", "LIMO3"])
@pytest.mark.parametrize("rie", ["HLLC", "HLLE"])
def test_limiter_riemann_matrix(oracle_mod, lim, rie):
    pb = cases.riemann_planar(seed=3, ni=200, nj=12, limiter=lim, riemann=rie, n_rk=3, bc_y=300)
    g, go, o, oo = run_both(oracle_mod, pb, 60)
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2])


def test_multiblock_planar_shock(oracle_mod):
    """2x2 blocks with BC 101 connections, BC 201 periodicity and BC 400 ends."""
    pb = cases.planar_shock(seed=2, n=96, steps=40)
    g, go, o, oo = run_both(oracle_mod, pb, 40)
    for b in range(4):
        assert np.array_equal(g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]), b


def test_ramp_local_dt(oracle_mod):
    """Sheared mesh, BC 405 / 400 / 300, local time stepping, HLLE."""
    pb = cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5)
    g, go, o, oo = run_both(oracle_mod, pb, 80)
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2])
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


def test_area_variation_source(oracle_mod):
    pb = cases.uniform_inflow(160, 40, 1.8, riemann="HLLC", area_a=0.6, precision="fp64", n_rk=2, steps=60)
    g, go, o, oo = run_both(oracle_mod, pb, 60)
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2])


def test_step_equals_graph_run(oracle_mod):
    pb = cases.riemann_planar(seed=5, ni=150, nj=10)
    g1 = GpuSolver(pb)
    for _ in range(30):
        g1.step()
    g2 = GpuSolver(pb)
    g2.run(30)
    assert np.array_equal(g1.state(0), g2.state(0))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graphs_captured_ahead_equal_lazy_capture(precision):
    """q2d_gpu_capture before the first step does nothing; after it, the graphs it captures replay the same
    steps as graphs captured inside q2d_gpu_run (state, dt, time and residual bit for bit)."""
    make = lambda: cases.planar_shock(seed=2, n=48, steps=10 ** 6, precision=precision)
    a, b = GpuSolver(make()), GpuSolver(make())
    a.capture()                                    # first step still pending: nothing to capture
    oa, ob = a.run(3), b.run(3)
    a.capture()
    oa, ob = a.run(21), b.run(21)
    for k in range(len(a.blocks)):
        assert np.array_equal(a.state(k), b.state(k))
    for x, y in zip(oa, ob):
        assert x["iter"] == y["iter"] == 24 and x["dtglobal"] == y["dtglobal"] and x["time"] == y["time"]
        assert x["residual"] == y["residual"]
    a.close()
    b.close()


def test_multi_step_graph_stops_at_itermax(oracle_mod):
    """A case whose itermax falls inside a multi-step graph replay stops there, like the oracle."""
    pb = cases.riemann_planar(seed=5, ni=150, nj=10, steps=21)
    g = GpuSolver(pb)
    out = g.run(30)
    assert out[0]["iter"] == 21 and out[0]["todo"] == 3
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(21)
    assert rc == 0 and oo[0]["iter"] == 21
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2])
    assert out[0]["dtglobal"] == oo[0]["dtglobal"] and out[0]["time"] == oo[0]["time"]


def two_rarefaction(U, p, cfl, limiter, precision):
    """64x8 tube, rho = 1, p small, u = -U left of the diaphragm and +U right of it: a vacuum opens."""
    pb = cases.riemann_planar(seed=3, ni=64, nj=8, steps=50, limiter=limiter, cfl=cfl, precision=precision)
    P = pb.blocks[0].P
    x = np.arange(P.shape[1])
    P[:, :, 1] = np.where(x[None, :] < P.shape[1] // 2, -U, U)
    P[:, :, 0] = 1.0
    P[:, :, 3] = p
    return pb


@pytest.mark.parametrize("U,p,cfl,limiter", [(5.0, 1e-3, 0.9, "superbee"), (2.0, 0.05, 1.5, "minmod")])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_failure_reported_like_oracle(oracle_mod, U, p, cfl, limiter, precision):
    """Negative pressure (Check_And_Fix_State's stop, Lib_Newstate.f90:327-331): the device reports it in
    the step and cell where the oracle stops (FP64: same status and cell; mixed: same step, a failure),
    and the case stops advancing."""
    o = oracle_mod.Oracle(two_rarefaction(U, p, cfl, limiter, "fp64"))
    rc, oo = o.run(50)
    assert rc == -4 and oo[0]["status"] == -4
    g = GpuSolver(two_rarefaction(U, p, cfl, limiter, precision))
    out = g.run(50, raise_on_error=False)
    assert out[0]["iter"] == oo[0]["iter"]
    if precision == "fp64":
        assert out[0]["status"] == oo[0]["status"] and out[0]["fail"] == oo[0]["fail"]
    else:
        assert out[0]["status"] in (-4, -5)
    again = g.run(5, raise_on_error=False)           # a failed case is frozen
    assert again[0]["iter"] == out[0]["iter"]


def test_deterministic_residual(oracle_mod):
    pb = cases.planar_shock(seed=4, n=64, steps=10)
    r1 = GpuSolver(pb).run(10)[0]["residual"]
    r2 = GpuSolver(pb).run(10)[0]["residual"]
    assert r1 == r2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_cell_wide_block(oracle_mod, precision):
    """The reference's 1D Shu-Osher case (400 x 1 cells, BC 400 ends, BC 300 on the y faces): a block one
    cell wide, where the y-face records read each other's ghost cells (two-phase ghost fill and BC flux).
    FP64: bit-identical to the oracle; mixed: relative L1(rho) <= 1e-5."""
    pb = cases.shu_osher(400, limiter="minmod")
    for c in pb.cases:
        c.precision = precision
    g = GpuSolver(pb)
    go = g.run(300)
    pb64 = cases.shu_osher(400, limiter="minmod")
    o = oracle_mod.Oracle(pb64)
    rc, oo = o.run(300, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == 300
    Pg, Po = g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2]
    if precision == "fp64":
        assert np.array_equal(Pg, Po), np.abs(Pg - Po).max()
        assert go[0]["time"] == oo[0]["time"]
    else:
        err = np.abs(Pg[..., 0] - Po[..., 0]).sum() / np.abs(Po[..., 0]).sum()
        assert 0.0 < err <= 1e-5, err
