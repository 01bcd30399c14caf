"""The oracle against the reference binary's own output.

test/1D/ShuOsher-LIMO3/reference/baseline_{minmod,vanleer,mc,iord}.dat hold the
(x, rho) profiles the reference binary wrote at N=400, t=1.8 (HLLC, SSP-RK3,
conservative update, CFL 0.5, BC 400 on the i-ends and 300 on the j-faces,
Air table; gen_case.py).  They are printed with 11 significant digits, so a
restatement that follows the reference's arithmetic reproduces them to ~5e-11
relative.  This pins, end to end: MUSCL reconstruction and the four limiters,
riemann_HLLC with roe_averages, Convective_Flux, BC_Extrapolation,
BC_Symmetry_Eul, Ghost_Extrapolate / Ghost_Symmetry, Compute_dt (ndir = 1),
Set_Global_dt, the RK stage and the conservative new state, and the
time-threshold stop.
"""
from pathlib import Path

import numpy as np
import pytest

from q2d_amd import cases

GOLD = Path(__file__).resolve().parent / "golden" / "shu_osher"


@pytest.mark.parametrize("lim", ["minmod", "vanleer", "mc", "iord"])
def test_shu_osher_baseline(oracle_mod, lim):
    ref = np.loadtxt(GOLD / f"baseline_{lim}.dat")
    pb = cases.shu_osher(400, limiter=lim)
    o = oracle_mod.Oracle(pb)
    rc, outs = o.run(100000)
    assert rc == 0
    assert outs[0]["todo"] == 3 and outs[0]["time"] >= 1.8
    rho = o.state(0)[2, 2:-2, 0]
    np.testing.assert_allclose(pb.meta["xc"], ref[:, 0], rtol=0, atol=1e-12)
    rel = np.abs(rho - ref[:, 1]) / np.abs(ref[:, 1])
    assert rel.max() < 1e-10, rel.max()


def test_shu_osher_discriminates_limiters(oracle_mod):
    """Negative control: the minmod run is far from the van Leer baseline."""
    ref = np.loadtxt(GOLD / "baseline_vanleer.dat")[:, 1]
    pb = cases.shu_osher(400, limiter="minmod")
    o = oracle_mod.Oracle(pb)
    o.run(100000)
    rho = o.state(0)[2, 2:-2, 0]
    assert (np.abs(rho - ref) / ref).max() > 1e-2


def test_gas_constant_is_pinned(oracle_mod):
    """Negative control: R_u = 8314.46 (docs/theory/thermo.md:56) misses the baseline by ~1e-5."""
    ref = np.loadtxt(GOLD / "baseline_minmod.dat")[:, 1]
    pb = cases.shu_osher(400, limiter="minmod")
    pb.cases[0].gas_R = 8314.46 / 28.970418
    o = oracle_mod.Oracle(pb)
    o.run(100000)
    rho = o.state(0)[2, 2:-2, 0]
    assert (np.abs(rho - ref) / ref).max() > 1e-6
