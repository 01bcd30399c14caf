"""Mixed-precision form (FP32 reconstruction/limiter/Riemann, FP64 state/update/dt/norms) against the
FP64 oracle: relative L1 density error <= 1e-5 (the stated tolerance), and the FP64 control flow
(dt, time, iteration count) unchanged in kind.  Each case also states its relative Linf density error
(max |d rho| / max |rho|), recorded as a test property and bounded by 1e-3 (a shock position off by
a cell would show as O(1e-1))."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-5
TOL_LINF = 1e-3


def rel_l1_rho(pb, g, o):
    num = den = 0.0
    for b in range(len(pb.blocks)):
        a = g.state(b)[2:-2, 2:-2, 0]
        r = o.state(b)[2:-2, 2:-2, 0]
        num += np.abs(a - r).sum()
        den += np.abs(r).sum()
    return num / den


def rel_linf_rho(pb, g, o):
    d = max(np.abs(g.state(b)[2:-2, 2:-2, 0] - o.state(b)[2:-2, 2:-2, 0]).max() for b in range(len(pb.blocks)))
    return d / max(np.abs(o.state(b)[2:-2, 2:-2, 0]).max() for b in range(len(pb.blocks)))


@pytest.mark.parametrize("make", [
    lambda: cases.planar_shock(seed=1, n=96, steps=60, precision="mixed"),
    lambda: cases.riemann_planar(seed=2, ni=256, nj=8, steps=300, precision="mixed"),
    lambda: cases.ramp_channel(ni=96, nj=48, wedge_deg=10.0, mach=2.0, precision="mixed", itermax=200),
    lambda: cases.uniform_inflow(128, 32, 1.6, riemann="HLLE", area_a=0.5, precision="mixed", steps=100),
])
def test_mixed_within_tolerance(oracle_mod, make, record_property):
    pb = make()
    steps = pb.cases[0].itermax
    g = GpuSolver(pb)
    go = g.run(steps)
    pb64 = make()
    for c in pb64.cases:
        c.precision = "fp64"
    o = oracle_mod.Oracle(pb64)
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0 and go[0]["iter"] == oo[0]["iter"] == steps
    err = rel_l1_rho(pb, g, o)
    linf = rel_linf_rho(pb, g, o)
    record_property("rel_L1_rho", err)
    record_property("rel_Linf_rho", linf)
    print(f"mixed vs FP64 oracle: rel L1(rho) {err:.3e}, rel Linf(rho) {linf:.3e}")
    assert err <= TOL_L1, err
    assert linf <= TOL_LINF, linf
    assert err > 0.0                      # it really is a different arithmetic


def test_mixed_step_equals_run():
    pb = cases.planar_shock(seed=3, n=64, steps=20, precision="mixed")
    g1 = GpuSolver(pb)
    for _ in range(20):
        g1.step()
    g2 = GpuSolver(pb)
    g2.run(20)
    for b in range(4):
        assert np.array_equal(g1.state(b), g2.state(b))
