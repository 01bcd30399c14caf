"""Laminar Navier-Stokes on the HIP path against the oracle, through the C-ABI (needs the GPU).

FP64 form: bit-identical state and local time step (interior diffusive fluxes, viscous symmetry and
isothermal-wall fluxes, axisymmetric viscous source with reflected wall ghosts, VNN limit).
Mixed form (FP32 convective and diffusive face fluxes, FP64 state and update): relative L1 density error
<= 1e-5 against the FP64 oracle.  The full-size axisymmetric pipe (2048x512, first cell 1e-5 L, Re_L 1e5)
is checked for a few steps in both forms, and a mixed laminar case on k_stage_mx keeps its bits beside
cases of the generic laminar kernel."""
import math

import numpy as np
import pytest

from q2d_amd import GpuSolver, cases
from q2d_amd.capi import Q2DError

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-5


def _planar_viscous(riemann="HLLE", limiter="vanleer", n_rk=2, theta=0.4):
    """Sheared, rotated planar channel: BC 400 ends, symmetry (300) below, isothermal wall (302) above."""
    ni, nj = 90, 40
    node = cases.cartesian_nodes(ni, nj, 0.0, 1.0, 0.0, 0.4)
    node[..., 1] += 0.05 * np.sin(math.pi * node[..., 0]) * (0.4 - node[..., 1])
    ct, st = math.cos(theta), math.sin(theta)
    rot = node.copy()
    rot[..., 0] = ct * node[..., 0] - st * node[..., 1]
    rot[..., 1] = st * node[..., 0] + ct * node[..., 1]
    from q2d_amd.geometry import BlockGeometry
    geo = BlockGeometry(rot)
    P = np.zeros((nj + 4, ni + 4, 4))
    c = 0.25 * (node[1:, 1:] + node[1:, :-1] + node[:-1, 1:] + node[:-1, :-1])
    x, y = c[..., 0], c[..., 1]
    u = 1.5 * (1.0 - (y / 0.4) ** 2) + 0.1 * np.sin(6 * x)
    v = 0.05 * np.cos(5 * x) * y
    P[2:-2, 2:-2, 0] = 1.0 + 0.1 * np.exp(-((x - 0.4) ** 2 + (y - 0.2) ** 2) / 0.01)
    P[2:-2, 2:-2, 1] = ct * u - st * v
    P[2:-2, 2:-2, 2] = st * u + ct * v
    P[2:-2, 2:-2, 3] = 1.0 + 0.05 * np.cos(3 * x)
    cases.fill_initial_ghosts(P)
    recs = cases.face_records(1, ni, nj, {1: 400, 2: 400, 3: 300, 4: 302}, props={4: {"T": 1.2}})
    spec = cases.CaseSpec(n_rk=n_rk, limiter=limiter, riemann=riemann, time_accurate=True, model=1, source="none",
                          cfl=0.5, vnn=0.3, mu_ref=2e-3, T_ref=1.0, S_suth=0.38, Pr=0.72, itermax=100)
    pb = cases.Problem(blocks=[cases.Block(geo, P)], bc=recs, cases=[spec])
    return cases.finish(pb, [geo])


def _run_both(oracle_mod, pb, steps):
    g = GpuSolver(pb)
    go = g.run(steps)
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0
    return g, go, o, oo


def test_axisymmetric_pipe_fp64_bit_exact(oracle_mod):
    """Axisymmetric laminar pipe (axis 300, isothermal wall 302, inflow 405, outflow 400), local dt."""
    pb = cases.axisymmetric_pipe(seed=1, ni=124, nj=40, first_cell=2e-4, reynolds=2e4, precision="fp64",
                                 steps=150)
    g, go, o, oo = _run_both(oracle_mod, pb, 150)
    assert go[0]["iter"] == oo[0]["iter"] == 150
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2]), \
        np.abs(g.state(0)[2:-2, 2:-2] - o.state(0)[2:-2, 2:-2]).max()
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))
    for k in range(4):
        assert abs(go[0]["residual"][k] - oo[0]["residual"][k]) <= 1e-12 * abs(oo[0]["residual"][k])


@pytest.mark.parametrize("rie,lim,n_rk", [("HLLE", "vanleer", 2), ("HLLC", "minmod", 3)])
def test_planar_viscous_fp64_bit_exact(oracle_mod, rie, lim, n_rk):
    pb = _planar_viscous(rie, lim, n_rk)
    g, go, o, oo = _run_both(oracle_mod, pb, 80)
    assert go[0]["dtglobal"] == oo[0]["dtglobal"]
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2]), \
        np.abs(g.state(0)[2:-2, 2:-2] - o.state(0)[2:-2, 2:-2]).max()


def test_axisymmetric_pipe_mixed_within_tolerance(oracle_mod):
    make = lambda prec: cases.axisymmetric_pipe(seed=2, ni=124, nj=40, first_cell=2e-4, reynolds=2e4,
                                                precision=prec, steps=200)
    g = GpuSolver(make("mixed"))
    go = g.run(200)
    o = oracle_mod.Oracle(make("fp64"))
    rc, oo = o.run(200, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == 200
    a, r = g.state(0)[2:-2, 2:-2, 0], o.state(0)[2:-2, 2:-2, 0]
    err = np.abs(a - r).sum() / np.abs(r).sum()
    assert 0 < err <= TOL_L1, err


def test_axisymmetric_pipe_full_size(oracle_mod):
    """The laminar workload at its full size (2048x512, tanh wall clustering with a first cell of 1e-5 L,
    Re_L 1e5, SSP-RK3, local dt): 8 steps, FP64 bit-identical to the oracle (state and local dt), the mixed
    form within its tolerance of the same oracle run."""
    steps = 8
    make = lambda prec: cases.axisymmetric_pipe(seed=0, precision=prec, steps=steps)
    pb = make("fp64")
    print(f"full-size pipe {pb.blocks[0].ni}x{pb.blocks[0].nj}, {steps} steps", flush=True)
    assert (pb.blocks[0].ni, pb.blocks[0].nj) == (2048, 512)
    g, go, o, oo = _run_both(oracle_mod, pb, steps)
    assert go[0]["iter"] == oo[0]["iter"] == steps
    ref = o.state(0)[2:-2, 2:-2]
    assert np.array_equal(g.state(0)[2:-2, 2:-2], ref), np.abs(g.state(0)[2:-2, 2:-2] - ref).max()
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))
    g.close()
    gm = GpuSolver(make("mixed"))
    gmo = gm.run(steps)
    assert gmo[0]["iter"] == steps and gmo[0]["status"] == 0
    a, r = gm.state(0)[2:-2, 2:-2, 0], ref[..., 0]
    err = np.abs(a - r).sum() / np.abs(r).sum()
    assert 0 < err <= TOL_L1, err
    gm.close()


def test_mixed_laminar_case_keeps_its_bits_beside_generic_cases():
    """A mixed laminar case that runs k_stage_mx (fused viscous terms and dt) gives the same state and local
    dt alone and in a batch with an HLLEM laminar case (the generic laminar kernel; the handle then runs the
    dt pass every step, evaluating the same formula) and an area-variation case (k_visc)."""
    steps = 40
    pipe = lambda: cases.axisymmetric_pipe(seed=4, ni=96, nj=32, first_cell=5e-4, reynolds=1e4, precision="mixed",
                                           steps=10 ** 6)
    solo = GpuSolver(pipe())
    so = solo.run(steps)
    assert so[0]["status"] == 0 and so[0]["iter"] == steps
    batch = pipe()
    hllem = cases.axisymmetric_pipe(seed=5, ni=64, nj=24, first_cell=5e-4, reynolds=1e4, precision="mixed",
                                    steps=10 ** 6)
    hllem.cases[0].riemann = "HLLEM"
    batch.add(hllem)
    batch.add(cases.laminar_area_channel(seed=6, ni=64, nj=24, area_a=0.4, precision="mixed", steps=10 ** 6))
    g = GpuSolver(batch)
    go = g.run(steps)
    assert all(o["status"] == 0 and o["iter"] == steps for o in go)
    assert np.array_equal(g.state(0), solo.state(0)), np.abs(g.state(0) - solo.state(0)).max()
    assert np.array_equal(g.dtlocal(0), solo.dtlocal(0))
    g.close()
    solo.close()


def test_laminar_area_variation_fp64_bit_exact(oracle_mod):
    """Quasi-2D laminar channel with area variation (V.G in the face and symmetry-plane stresses,
    Source_Area_Visc_Blk), local dt: bit-identical to the oracle."""
    pb = cases.laminar_area_channel(seed=5, ni=124, nj=40, area_a=0.6, precision="fp64", steps=150)
    g, go, o, oo = _run_both(oracle_mod, pb, 150)
    assert go[0]["iter"] == oo[0]["iter"] == 150
    assert np.array_equal(g.state(0)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2]), \
        np.abs(g.state(0)[2:-2, 2:-2] - o.state(0)[2:-2, 2:-2]).max()
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


def test_laminar_area_variation_mixed_within_tolerance(oracle_mod):
    make = lambda prec: cases.laminar_area_channel(seed=6, ni=124, nj=40, area_a=0.4, precision=prec, steps=200)
    g = GpuSolver(make("mixed"))
    go = g.run(200)
    o = oracle_mod.Oracle(make("fp64"))
    rc, oo = o.run(200, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == 200
    a, r = g.state(0)[2:-2, 2:-2, 0], o.state(0)[2:-2, 2:-2, 0]
    err = np.abs(a - r).sum() / np.abs(r).sum()
    assert 0 < err <= TOL_L1, err


def test_viscous_and_inviscid_cases_in_one_batch(oracle_mod):
    """A laminar and an Euler case share one launch; each equals its own oracle run bit for bit."""
    pv = cases.axisymmetric_pipe(seed=3, ni=64, nj=24, first_cell=5e-4, reynolds=1e4, precision="fp64", steps=60)
    pe = cases.riemann_planar(seed=4, ni=128, nj=8, steps=60)
    batch = cases.axisymmetric_pipe(seed=3, ni=64, nj=24, first_cell=5e-4, reynolds=1e4, precision="fp64", steps=60)
    batch.add(cases.riemann_planar(seed=4, ni=128, nj=8, steps=60))
    g = GpuSolver(batch)
    g.run(60)
    for k, pb in enumerate((pv, pe)):
        o = oracle_mod.Oracle(pb)
        rc, _ = o.run(60, stop_on_todo=False)
        assert rc == 0
        assert np.array_equal(g.state(k)[2:-2, 2:-2], o.state(0)[2:-2, 2:-2]), k


def _laminar(pb):
    for c in pb.cases:
        c.model, c.mu_ref, c.T_ref, c.S_suth, c.Pr, c.vnn = 1, 2e-3, 1.0, 0.38, 0.72, 0.2
    return pb


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_laminar_block_connections(oracle_mod, precision):
    """Laminar planar shock on 2x2 blocks: BC 101 seams and BC 201 periodicity carry the viscous connection
    flux (BC_Connection_Visc).  FP64: bit-identical to the oracle; mixed: relative L1(rho) <= 1e-5."""
    pb = _laminar(cases.planar_shock(seed=0, n=32, steps=40, precision=precision))
    g = GpuSolver(pb)
    go = g.run(40)
    o = oracle_mod.Oracle(_laminar(cases.planar_shock(seed=0, n=32, steps=40, precision="fp64")))
    rc, oo = o.run(40, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == 40
    if precision == "fp64":
        assert go[0]["dtglobal"] == oo[0]["dtglobal"]
        for b in range(4):
            assert np.array_equal(g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]), \
                (b, np.abs(g.state(b)[2:-2, 2:-2] - o.state(b)[2:-2, 2:-2]).max())
    else:
        num = sum(np.abs(g.state(b)[2:-2, 2:-2, 0] - o.state(b)[2:-2, 2:-2, 0]).sum() for b in range(4))
        den = sum(np.abs(o.state(b)[2:-2, 2:-2, 0]).sum() for b in range(4))
        assert 0 < num / den <= TOL_L1, num / den


def test_laminar_connection_needs_orientation():
    pb = _laminar(cases.planar_shock(seed=0, n=16, steps=5))
    for r in pb.bc:
        if r["type"] in (101, 201):
            r["d11"] = 0
    with pytest.raises(Q2DError) as ei:
        GpuSolver(pb)
    assert ei.value.code == -1
