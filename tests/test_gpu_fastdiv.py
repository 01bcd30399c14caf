"""The FP64 stage kernel's fast-division pass against the oracle (needs the GPU).

The FP64 stage kernel divides by the hardware reciprocal, two Newton steps and one remainder
correction, without the range-scaling steps of the compiler's IEEE division, on work items whose
numerators (0 or at least 2^-969), quotients (0 or at least 2^-1020), state values (density and
pressure 0 or at least 2^-100, every component below 2^100) and geometry (0 or within [2^-100, 2^100))
keep that sequence equal to the IEEE quotient; a work item that meets a value outside is computed again
with the IEEE division.  These cases put values on both sides of every bound into the state and the
geometry, and hold the device bit-identical to the oracle (state, dt, time, residual)."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu

# velocity magnitudes around the numerator and quotient bounds, down to the smallest subnormal
TINY = [5e-324, 2.0 ** -1074 * 3, 2.0 ** -1040, 2.0 ** -1022, 2.0 ** -1000, 2.0 ** -970, 2.0 ** -969,
        2.0 ** -968, 1e-290, 1e-250, 1e-200, 1e-150, 2.0 ** -101, 2.0 ** -99, 1e-20]


def _bit_exact(oracle_mod, pb, steps):
    g = GpuSolver(pb)
    go = g.run(steps)
    o = oracle_mod.Oracle(pb)
    rc, oo = o.run(steps, stop_on_todo=False)
    assert rc == 0
    for k in range(len(pb.cases)):
        assert go[k]["status"] == 0 and go[k]["iter"] == oo[k]["iter"] == steps
        assert go[k]["dtglobal"] == oo[k]["dtglobal"] and go[k]["time"] == oo[k]["time"]
        for v in range(4):
            assert abs(go[k]["residual"][v] - oo[k]["residual"][v]) <= 1e-12 * abs(oo[k]["residual"][v])
    for b in range(len(pb.blocks)):
        Pg, Po = g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]
        assert np.array_equal(Pg, Po), (b, np.abs(Pg - Po).max())
        if not pb.cases[pb.blocks[b].case_id].time_accurate:
            assert np.array_equal(g.dtlocal(b), o.dtlocal(b)), b
    g.close()


@pytest.mark.parametrize("lim,rie", [("vanleer", "HLLC"), ("vanalbada", "HLLE"), ("minmod", "HLLEM"),
                                     ("LIMO3", "HLLC"), ("superbee", "HLLE"), ("mc", "HLLC")])
def test_tiny_velocities_ahead_of_a_shock(oracle_mod, lim, rie):
    """2x2-block planar shock; the gas at rest ahead of it carries velocity components of either sign from
    the list above (subnormals, the bounds themselves and values just inside), so that some work items
    run the fast pass throughout and others meet a value outside its range."""
    pb = cases.planar_shock(seed=5, n=32, steps=1 << 30, limiter=lim, riemann=rie)
    rng = np.random.default_rng(11)
    for blk in pb.blocks:
        P = blk.P[2:-2, 2:-2]
        rest = (P[..., 1] == 0.0) & (P[..., 2] == 0.0)
        pick = rest & (rng.random(rest.shape) < 0.15)
        n = int(pick.sum())
        for c in (1, 2):
            mags = rng.choice(TINY, n)
            P[..., c][pick] = mags * rng.choice([-1.0, 1.0], n)
        cases.fill_initial_ghosts(blk.P)
    _bit_exact(oracle_mod, pb, 24)


@pytest.mark.parametrize("s_rho,s_p", [(1.0, 2.0 ** -99), (1.0, 2.0 ** -101), (2.0 ** 99, 2.0 ** 99),
                                       (2.0 ** 101, 2.0 ** 101)])
def test_density_and_pressure_scaled_to_the_state_bounds(oracle_mod, s_rho, s_p):
    """The 256x8 two-state Riemann problem with its pressure (and density) multiplied by a factor: a
    pressure of order 2^-99, or density and pressure of order 2^99, keep every work item on the fast
    pass; 2^-101 and 2^101 put every state outside its range, so that every item runs the IEEE pass.
    (Densities below 1e-20 are clamped by the reference's state check, so the density is not scaled down.)"""
    pb = cases.riemann_planar(seed=1, ni=256, nj=8, limiter="vanleer", riemann="HLLC", n_rk=3)
    for blk in pb.blocks:
        blk.P[..., 0] *= s_rho
        blk.P[..., 3] *= s_p
    _bit_exact(oracle_mod, pb, 40)


@pytest.mark.parametrize("length", [1e-32, 1e33])
def test_geometry_outside_the_range_takes_the_ieee_pass(oracle_mod, length):
    """The same problem on a domain of length 1e-32 (cells below 2^-100) or 1e33 (cells above 2^100): the
    blocks' geometry fails the upload check and every item runs the IEEE pass."""
    pb = cases.riemann_planar(seed=1, ni=128, nj=6, limiter="minmod", riemann="HLLE", n_rk=2, length=length)
    _bit_exact(oracle_mod, pb, 30)
