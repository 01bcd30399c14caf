"""Implicit residual smoothing on the device (k_irs_i, k_irs_j) against the oracle's restatement of
Lib_IRS.f90, through the C-ABI.  FP64: bit-identical states, local dt and iteration control; mixed
form: relative L1 density <= 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu


def with_irs(pb, beta):
    for c in pb.cases:
        c.irs_beta = beta
    return pb


def run_both(oracle_mod, pb, steps):
    print(f"irs case: {pb.meta.get('name')} beta={pb.cases[0].irs_beta} steps={steps}", flush=True)
    g = GpuSolver(pb)
    go = g.run(steps)
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0
    return g, go, o, oo


def assert_bit_exact(pb, g, o, go, oo):
    for b in range(len(pb.blocks)):
        Pg, Po = g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]
        assert np.array_equal(Pg, Po), (b, np.abs(Pg - Po).max())
    for c in range(len(pb.cases)):
        assert go[c]["iter"] == oo[c]["iter"]
        assert go[c]["dtglobal"] == oo[c]["dtglobal"]
        for k in range(4):
            assert abs(go[c]["residual"][k] - oo[c]["residual"][k]) <= 1e-12 * max(1e-300, abs(oo[c]["residual"][k]))


def test_irs_multiblock_connections(oracle_mod):
    """2x2 blocks: ghosts from donors across BC 101, zero-gradient across 201 and 400; RK3, global dt."""
    pb = with_irs(cases.planar_shock(seed=2, n=96, steps=30), 0.5)
    g, go, o, oo = run_both(oracle_mod, pb, 30)
    assert_bit_exact(pb, g, o, go, oo)


def test_irs_ramp_local_dt(oracle_mod):
    """The reference's steady setting (irs-beta 0.5, local dt, HLLE, minmod): the fused next-step dt of
    the smoothing kernel's last stage included."""
    pb = with_irs(cases.ramp_channel(ni=96, nj=48, wedge_deg=12.0, mach=2.5, itermax=80), 0.5)
    g, go, o, oo = run_both(oracle_mod, pb, 80)
    assert_bit_exact(pb, g, o, go, oo)
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


@pytest.mark.parametrize("n_rk,beta", [(1, 0.3), (2, 2.0)])
def test_irs_rk_orders(oracle_mod, n_rk, beta):
    pb = with_irs(cases.riemann_planar(seed=3, ni=200, nj=12, n_rk=n_rk, bc_y=300, limiter="vanleer"), beta)
    g, go, o, oo = run_both(oracle_mod, pb, 60)
    assert_bit_exact(pb, g, o, go, oo)


def test_irs_laminar(oracle_mod):
    """Laminar channel with the area-variation source (viscous stage kernel, viscous boundary fluxes)."""
    pb = with_irs(cases.laminar_area_channel(seed=5, ni=124, nj=40, area_a=0.6, precision="fp64", steps=60), 0.5)
    g, go, o, oo = run_both(oracle_mod, pb, 60)
    assert_bit_exact(pb, g, o, go, oo)
    assert np.array_equal(g.dtlocal(0), o.dtlocal(0))


def test_irs_one_cell_wide(oracle_mod):
    """A block one cell wide (ndir = 1): the smoothing runs along i only."""
    pb = with_irs(cases.shu_osher(200, limiter="minmod"), 0.5)
    g, go, o, oo = run_both(oracle_mod, pb, 100)
    assert_bit_exact(pb, g, o, go, oo)
    assert go[0]["time"] == oo[0]["time"]


def test_irs_batch_with_plain_cases(oracle_mod):
    """Smoothed and plain cases of different sizes in one handle: each equals its own oracle run."""
    pb = with_irs(cases.ramp_channel(ni=64, nj=40, wedge_deg=8.0, mach=2.2, itermax=40), 0.5)
    pb.add(cases.riemann_planar(seed=4, ni=150, nj=10, steps=40))
    pb.add(with_irs(cases.planar_shock(seed=5, n=48, steps=40), 0.3))
    g = GpuSolver(pb)
    go = g.run(40)
    for c in range(3):
        sub = pb.case(c)
        o = oracle_mod.Oracle(sub)
        rc, oo = o.run(40, stop_on_todo=False)
        assert rc == 0
        blocks = [k for k, b in enumerate(pb.blocks) if b.case_id == c]
        for kb, b in enumerate(blocks):
            assert np.array_equal(g.state(b)[2:-2, 2:-2], o.state(kb)[2:-2, 2:-2]), (c, kb)
        assert go[c]["iter"] == oo[0]["iter"] == 40


@pytest.mark.parametrize("make", [
    lambda prec: cases.ramp_channel(ni=96, nj=48, wedge_deg=10.0, mach=2.0, precision=prec, itermax=150),
    lambda prec: cases.planar_shock(seed=1, n=96, steps=60, precision=prec),
])
def test_irs_mixed_within_tolerance(oracle_mod, make):
    pb = with_irs(make("mixed"), 0.5)
    steps = pb.cases[0].itermax
    g = GpuSolver(pb)
    go = g.run(steps)
    o = oracle_mod.Oracle(with_irs(make("fp64"), 0.5))
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == steps
    num = sum(np.abs(g.state(b)[2:-2, 2:-2, 0] - o.state(b)[2:-2, 2:-2, 0]).sum() for b in range(len(pb.blocks)))
    den = sum(np.abs(o.state(b)[2:-2, 2:-2, 0]).sum() for b in range(len(pb.blocks)))
    assert 0.0 < num / den <= 1e-5, num / den
