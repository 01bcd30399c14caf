"""Fused viscous terms of the mixed form (needs the GPU): for mixed laminar cases the laminar stage kernel
evaluates the diffusive face fluxes and the axisymmetric viscous source itself, in FP32 like the form's
convective faces, instead of reading them from k_visc's FP64 planes (q2d_gpu_set_fusion, on by default).
Both schedules stay within the mixed form's tolerance of the FP64 oracle (relative L1 density <= 1e-5)
on the axisymmetric pipe (axis, isothermal wall, reflected wall ghosts in the source), a planar channel
with a symmetry plane and a wall, the laminar shock on 2x2 blocks (viscous connections, periodicity) and
with residual smoothing; a batch with an area-variation case (which keeps k_visc) gives that case the
same bits under both schedules."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

from test_gpu_viscous import _laminar, _planar_viscous

pytestmark = pytest.mark.gpu
TOL_L1 = 1e-5


def with_prec(pb, prec):
    for c in pb.cases:
        c.precision = prec
    return pb


def run(pb, steps, fuse):
    g = GpuSolver(pb)
    g.set_fusion(fuse)
    out = g.run(steps)
    assert all(o["iter"] == steps and o["status"] == 0 for o in out)
    st = [g.state(b)[2:-2, 2:-2].copy() for b in range(len(pb.blocks))]
    g.close()
    return st


def rel_l1(a, b):
    return np.abs(a[..., 0] - b[..., 0]).sum() / np.abs(b[..., 0]).sum()


def within_tolerance(oracle_mod, make, steps):
    pb = make()
    print(f"fusion: {pb.meta.get('name')} {[(b.ni, b.nj) for b in pb.blocks]} {steps} steps", flush=True)
    o = oracle_mod.Oracle(with_prec(make(), "fp64"))
    rc, _ = o.run(steps, stop_on_todo=False)
    assert rc == 0
    ref = [o.state(b)[2:-2, 2:-2] for b in range(len(pb.blocks))]
    fused, plain = run(with_prec(make(), "mixed"), steps, True), run(with_prec(make(), "mixed"), steps, False)
    for b in range(len(ref)):
        ef, ep = rel_l1(fused[b], ref[b]), rel_l1(plain[b], ref[b])
        assert 0.0 < ef <= TOL_L1 and 0.0 < ep <= TOL_L1, (b, ef, ep)
        assert not np.array_equal(fused[b], plain[b])      # the fused schedule did run its own arithmetic


def test_axisymmetric_pipe(oracle_mod):
    within_tolerance(oracle_mod, lambda: cases.axisymmetric_pipe(seed=1, ni=124, nj=40, first_cell=5e-4,
                                                                 reynolds=1e4, precision="mixed", steps=10 ** 6), 60)


def test_planar_channel_symmetry_and_wall(oracle_mod):
    within_tolerance(oracle_mod, lambda: _planar_viscous(riemann="HLLC", limiter="minmod", n_rk=3), 50)


def test_laminar_shock_on_2x2_blocks(oracle_mod):
    within_tolerance(oracle_mod, lambda: _laminar(cases.planar_shock(seed=0, n=32, steps=10 ** 6)), 40)


def test_with_residual_smoothing(oracle_mod):
    def make():
        pb = cases.axisymmetric_pipe(seed=2, ni=96, nj=32, first_cell=5e-4, reynolds=1e4, precision="mixed",
                                     steps=10 ** 6)
        pb.cases[0].irs_beta = 0.5
        return pb
    within_tolerance(oracle_mod, make, 45)


def test_area_variation_case_keeps_k_visc():
    def make():
        pb = cases.axisymmetric_pipe(seed=3, ni=64, nj=24, first_cell=5e-4, reynolds=1e4, precision="mixed",
                                     steps=10 ** 6)
        pb.add(cases.laminar_area_channel(seed=6, ni=124, nj=40, area_a=0.4, precision="mixed", steps=10 ** 6))
        return pb
    a, b = run(make(), 40, True), run(make(), 40, False)
    assert np.array_equal(a[1], b[1])                       # the area-variation case: k_visc either way
    assert not np.array_equal(a[0], b[0])
