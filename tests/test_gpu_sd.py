"""MUSCL-SD (Tramel's detector in the stage kernel's row march and in the block-connection boundary
fluxes) and the HLLEM, HLLE++ and HLLC+ Riemann solvers on the device against the oracle, through the
C-ABI: FP64 bit-identical; mixed form within 1e-5 relative L1 density."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu


def run_both(oracle_mod, pb, steps):
    print(f"sd case: {pb.meta.get('name')} steps={steps}", flush=True)
    g = GpuSolver(pb)
    go = g.run(steps)
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0
    for b in range(len(pb.blocks)):
        Pg, Po = g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]
        assert np.array_equal(Pg, Po), (b, np.abs(Pg - Po).max())
    assert go[0]["iter"] == oo[0]["iter"] == steps
    assert go[0]["dtglobal"] == oo[0]["dtglobal"]
    return g, o


def test_sd_ramp_reference_setting(oracle_mod):
    """test/2D/euler/ramp-channel/input.ini: MUSCL-SD, minmod, HLLE, SSP-RK3, local dt, CFL 0.9,
    cfl-rise-threshold 100, irs-beta 0.5; run past the end of the CFL ramp."""
    pb = cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5, cfl=0.9, itermax=120)
    pb.cases[0].space_sd = True
    pb.cases[0].irs_beta = 0.5
    pb.cases[0].ramp_iter = 100
    g, o = run_both(oracle_mod, pb, 120)
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


@pytest.mark.parametrize("rie,lim", [("HLLC", "vanleer"), ("HLLE", "superbee"), ("HLLEM", "minmod"),
                                     ("HLLE++", "vanleer"), ("HLLC+", "minmod")])
def test_sd_multiblock_shock(oracle_mod, rie, lim):
    """The shock crosses the 101 connections between the 2x2 blocks: connection faces take the boundary
    cell's detector value."""
    pb = cases.planar_shock(seed=2, n=64, steps=30, riemann=rie, limiter=lim)
    pb.cases[0].space_sd = True
    run_both(oracle_mod, pb, 30)


@pytest.mark.parametrize("rie,sd", [("HLLEM", False), ("HLLEM", True), ("HLLE++", False), ("HLLE++", True),
                                    ("HLLC+", False), ("HLLC+", True)])
def test_hllem_ramp(oracle_mod, rie, sd):
    """HLLEM (Lib_Riemann_HLL.f90:355-432), HLLE++ (:435-530) and HLLC+ (:571-603), with and without
    MUSCL-SD, local dt."""
    pb = cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5, riemann=rie, itermax=60)
    pb.cases[0].space_sd = sd
    g, o = run_both(oracle_mod, pb, 60)
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


@pytest.mark.parametrize("rie", ["HLLC", "HLLE++"])
def test_sd_laminar(oracle_mod, rie):
    pb = cases.laminar_area_channel(seed=5, ni=124, nj=40, area_a=0.6, precision="fp64", steps=60, riemann=rie)
    pb.cases[0].space_sd = True
    run_both(oracle_mod, pb, 60)


def test_sd_batch_with_plain_cases(oracle_mod):
    pb = cases.planar_shock(seed=4, n=48, steps=20)
    pb.cases[0].space_sd = True
    pb.add(cases.riemann_planar(seed=4, ni=150, nj=10, steps=20))
    g = GpuSolver(pb)
    g.run(20)
    for c in range(2):
        o = oracle_mod.Oracle(pb.case(c))
        rc, _ = o.run(20, stop_on_todo=False)
        assert rc == 0
        blocks = [k for k, b in enumerate(pb.blocks) if b.case_id == c]
        for kb, b in enumerate(blocks):
            assert np.array_equal(g.state(b)[2:-2, 2:-2], o.state(kb)[2:-2, 2:-2]), (c, kb)


@pytest.mark.parametrize("rie,sd,irs", [("HLLE", True, 0.5), ("HLLEM", False, 0.0), ("HLLE++", True, 0.0),
                                        ("HLLC+", False, 0.5)])
def test_mixed_form_within_tolerance(oracle_mod, rie, sd, irs):
    """Mixed form (FP32 faces in the generic stage kernel, FP64 state): relative L1 density <= 1e-5 of the
    FP64 oracle; the first case is the reference's ramp-channel setting."""
    def make(prec):
        pb = cases.ramp_channel(ni=96, nj=48, wedge_deg=10.0, mach=2.0, riemann=rie, cfl=0.9 if irs else 0.5,
                                precision=prec, itermax=150)
        pb.cases[0].space_sd = sd
        pb.cases[0].irs_beta = irs
        return pb
    pb = make("mixed")
    print(f"mixed detector case: {rie} sd={sd} irs={irs}", flush=True)
    g = GpuSolver(pb)
    go = g.run(150)
    o = oracle_mod.Oracle(make("fp64"))
    rc, oo = o.run(150, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == 150
    a, r = g.state(0)[2:-2, 2:-2, 0], o.state(0)[2:-2, 2:-2, 0]
    err = np.abs(a - r).sum() / np.abs(r).sum()
    assert 0.0 < err <= 1e-5, err


@pytest.mark.parametrize("rie,sd", [("HLLC", True), ("HLLC+", False)])
def test_mixed_2x2_blocks_detector_within_tolerance(oracle_mod, rie, sd):
    """Mixed form on 2x2 blocks (BC 101 connections, y-periodic BC 201): the detector of MUSCL-SD and HLLC+
    reads across block connections in the FP32 boundary-flux kernel; relative L1 density <= 1e-5 of the
    FP64 oracle on every block."""
    def make(prec):
        pb = cases.planar_shock(seed=4, n=48, riemann=rie, precision=prec, steps=100)
        pb.cases[0].space_sd = sd
        return pb
    print(f"mixed 2x2 detector case: {rie} sd={sd}", flush=True)
    g = GpuSolver(make("mixed"))
    go = g.run(100)
    o = oracle_mod.Oracle(make("fp64"))
    rc, oo = o.run(100, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == 100
    for b in range(4):
        a, r = g.state(b)[2:-2, 2:-2, 0], o.state(b)[2:-2, 2:-2, 0]
        err = np.abs(a - r).sum() / np.abs(r).sum()
        assert 0.0 < err <= 1e-5, (b, err)
