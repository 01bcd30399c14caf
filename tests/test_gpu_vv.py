"""The HIP path on the cases that pin the oracle to the reference's own numbers and to closed forms
(needs the GPU): the cylindrical Noh implosion (axisymmetric Euler source), the 1D refinement ladder
(HLLE, HLLE++, first order, MC), the characteristic outflow BC 406 in all its branches, the supersonic
inflow BC 405 on a preserved stream, and the CFL ramp.  FP64 bit-identical to the oracle."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu


def both(oracle_mod, pb, steps=None):
    """Oracle to its stop (or `steps`), then the device for the same number of steps."""
    print(f"vv case: {pb.meta.get('name')} {[(b.ni, b.nj) for b in pb.blocks]}", flush=True)
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(steps if steps else 10 ** 6, stop_on_todo=steps is None)
    assert rc == 0
    n = oo[0]["iter"]
    g = GpuSolver(pb)
    go = g.run(n)
    assert [x["iter"] for x in go] == [x["iter"] for x in oo]
    for b in range(len(pb.blocks)):
        Pg, Po = g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]
        assert np.array_equal(Pg, Po), (b, np.abs(Pg - Po).max())
    for k in range(len(pb.cases)):
        assert go[k]["dtglobal"] == oo[k]["dtglobal"] and go[k]["time"] == oo[k]["time"]
    return g, go, o, oo


def test_noh_cylindrical_full_size(oracle_mod):
    """test/2D/euler/noh-cylindrical at its own size (4 x 1000 cells, to t = 1.0): bit-identical."""
    g, go, o, oo = both(oracle_mod, cases.noh_cylindrical())
    assert go[0]["todo"] == 3 and 1.0 <= go[0]["time"] < 1.01


@pytest.mark.parametrize("name", ["Ein91", "Tor99", "Noh87", "Sod79"])
@pytest.mark.parametrize("level", [1, 8])
def test_ladder_1d(oracle_mod, name, level):
    g, go, o, oo = both(oracle_mod, cases.ladder_1d(name, level))
    assert go[0]["todo"] == 3


@pytest.mark.parametrize("time_accurate", [False, True])
def test_outflow_406_branches(oracle_mod, time_accurate):
    """Subsonic exits with the back pressure above and below the critical pressure, a supersonic exit, and
    an exit whose flow enters the domain, batched in one handle."""
    pb = cases.outflow_channel(mach=0.5, p_exit=1.2, rel_fac=0.5, time_accurate=time_accurate)
    pb.add(cases.outflow_channel(mach=0.7, p_exit=0.2, rel_fac=1.0, time_accurate=time_accurate))
    pb.add(cases.outflow_channel(mach=2.5, p_exit=1.0, rel_fac=1.0, time_accurate=time_accurate))
    pb.add(cases.outflow_channel(mach=0.4, p_exit=1.0, rel_fac=0.8, backflow_cols=3, time_accurate=time_accurate))
    both(oracle_mod, pb, steps=100)


def test_outflow_406_mixed_form(oracle_mod):
    pb = cases.outflow_channel(mach=0.6, p_exit=1.1, rel_fac=0.5, precision="mixed", steps=100)
    g = GpuSolver(pb)
    g.run(100)
    ref = cases.outflow_channel(mach=0.6, p_exit=1.1, rel_fac=0.5, precision="fp64", steps=100)
    o = oracle_mod.Oracle(ref)
    rc, _ = o.run(100)
    assert rc == 0
    a, b = g.state(0)[2:-2, 2:-2, 0], o.state(0)[2:-2, 2:-2, 0]
    assert np.abs(a - b).sum() / np.abs(b).sum() <= 1e-5


@pytest.mark.parametrize("rel_fac", [1.0, 0.5])
def test_supersonic_inflow_free_stream(oracle_mod, rel_fac):
    pb = cases.uniform_inflow(120, 24, 2.2, riemann="HLLC", precision="fp64", n_rk=3, steps=50)
    pb.cases[0].time_accurate = False
    for r in pb.bc:
        if r["type"] == 405:
            r["rel_fac"] = rel_fac
    P0 = pb.blocks[0].P[2:-2, 2:-2].copy()
    g, go, o, oo = both(oracle_mod, pb, steps=50)
    scale = np.abs(P0).max(axis=(0, 1))
    scale[2] = scale[1]
    assert (np.abs(g.state(0)[2:-2, 2:-2] - P0).max(axis=(0, 1)) / scale).max() < 1e-13


def test_cfl_ramp_local_dt(oracle_mod):
    """cfl-rise-threshold = 100 (test/2D/euler/ramp-channel/input.ini) on the ramp, local dt, run past the
    end of the ramp: state and local dt bit-identical."""
    pb = cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5, cfl=0.9, itermax=130)
    pb.cases[0].ramp_iter = 100
    g, go, o, oo = both(oracle_mod, pb, steps=130)
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


def test_cfl_ramp_global_dt(oracle_mod):
    pb = cases.riemann_planar(seed=4, ni=256, nj=8, steps=80)
    pb.cases[0].ramp_iter = 50
    both(oracle_mod, pb, steps=80)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ramp_mixed_form_within_tolerance(oracle_mod, precision):
    """The mixed form's fused dt applies the ramp as the FP64 form does (CFL 0.5: without smoothing the
    ramp's CFL 0.9 is unstable, and the growing instability amplifies any rounding difference)."""
    pb = cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5, cfl=0.5, itermax=130, precision=precision)
    pb.cases[0].ramp_iter = 100
    g = GpuSolver(pb)
    g.run(130)
    ref = cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5, cfl=0.5, itermax=130)
    ref.cases[0].ramp_iter = 100
    o = oracle_mod.Oracle(ref)
    o.run(130)
    a, b = g.state(0)[2:-2, 2:-2, 0], o.state(0)[2:-2, 2:-2, 0]
    assert np.abs(a - b).sum() / np.abs(b).sum() <= (0.0 if precision == "fp64" else 1e-5)
