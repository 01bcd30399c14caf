"""q2d_gpu_step_noalloc (file-scope device scratch, no allocation) against q2d_gpu_step (the handle's own
scratch), through the C-ABI (needs the GPU): bit-identical states and step results on two problems in a
row, the second smaller, in FP64 and in the mixed form (dense FP32 boundary array), with residual
smoothing (k_irs partials) and a laminar case (viscous boundary fluxes); a handle whose scratch exceeds
the array is rejected."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases
from q2d_amd.capi import Q2DError

pytestmark = pytest.mark.gpu


def problems(prec):
    big = cases.planar_shock(seed=3, n=96, steps=10 ** 6, precision=prec)
    big.add(cases.axisymmetric_pipe(seed=2, ni=64, nj=24, first_cell=5e-4, reynolds=1e4, precision=prec,
                                    steps=10 ** 6))
    small = cases.ramp_channel(ni=64, nj=32, precision=prec, itermax=10 ** 6)
    small.cases[0].irs_beta = 0.5
    return big, small


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_step_equals_step(prec):
    for pb in problems(prec):                    # two calls in a row, the second with a smaller input
        print(f"noalloc: {prec} {pb.meta.get('name')} {pb.ncells()} cells", flush=True)
        a, b = GpuSolver(pb), GpuSolver(pb)
        for k in range(6):
            oa = a.step()
            ob = b.step_noalloc() if k % 3 != 2 else b.step()     # the two entry points mix on one handle
            assert oa == ob, k
        for blk in range(len(pb.blocks)):
            assert np.array_equal(a.state(blk), b.state(blk)), blk
        a.close()
        b.close()


def test_noalloc_rejects_scratch_beyond_the_maximum():
    pb = cases.Problem()
    for k in range(9):                            # 9 x 60,002 boundary records of 80 B > 32 MiB
        pb.add(cases.riemann_planar(seed=k, ni=30000, nj=1, steps=10 ** 6))
    g = GpuSolver(pb)
    with pytest.raises(Q2DError) as e:
        g.step_noalloc()
    assert e.value.code == -3 and "allocation-free maximum" in str(e.value)
    out = g.step()                                # the handle is unaffected
    assert all(o["iter"] == 1 for o in out)
    g.close()
