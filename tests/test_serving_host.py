"""Host logic of the serving loop (no GPU): the request generator's settings, the HBM accounting rule and
the oblique-shock analysis that classifies requests that do not converge."""
import math

import numpy as np
import pytest

from q2d_amd import cases, serving


def test_oblique_shock_textbook_values():
    # weak oblique shocks, gamma = 1.4 (NACA 1135 charts): (M1, wedge) -> (beta, M2)
    for m1, th, beta, m2 in ((2.0, 10.0, 39.31, 1.641), (3.0, 20.0, 37.76, 1.994), (3.0, 30.0, 52.01, 1.406)):
        b, m = serving.oblique_shock(m1, th)
        assert abs(b - beta) < 0.02 and abs(m - m2) < 2e-3, (m1, th, b, m)
    # a vanishing wedge gives the Mach wave and no change of Mach number
    b, m = serving.oblique_shock(2.0, 1e-7)
    assert abs(b - math.degrees(math.asin(0.5))) < 1e-3 and abs(m - 2.0) < 1e-5
    # detachment: M1 = 1.5 turns the flow by at most 12.11 deg
    assert serving.oblique_shock(1.5, 12.0) is not None and serving.oblique_shock(1.5, 12.2) is None


def test_stall_report_flags_an_unstarted_channel():
    req = serving.Request(7, 0.0, serving.ramp_request(3))
    P = np.zeros((6, 8, 4))
    P[..., 0], P[..., 3] = 1.0, 1.0
    P[..., 1] = 2.0 * math.sqrt(1.4)          # Mach 2 everywhere ...
    res = serving.Result(7, 0.0, 0.0, 1.0, 60000, 0, False, 48, [P.copy()])
    rep = serving.stall_report(req, res)
    assert not rep["inflow_column_subsonic"] and rep["subsonic_fraction"] == 0.0
    P[:, 0, 1] = 0.5 * math.sqrt(1.4)         # ... but subsonic in the first column
    res.state = [P]
    rep = serving.stall_report(req, res)
    assert rep["inflow_column_subsonic"] and rep["cause"].startswith("unstarted")
    assert abs(rep["min_mach"] - 0.5) < 1e-12


def test_requests_use_the_reference_ramp_channel_numerics():
    """The reference ramp-channel case's steady numerics, without its shock detector."""
    pb = serving.ramp_request(5)
    c = pb.cases[0]
    assert (c.space_sd, c.irs_beta, c.ramp_iter, c.cfl, c.limiter, c.riemann, c.n_rk) == \
        (False, 0.5, 100, 0.9, "minmod", "HLLE", 3)
    assert c.res_drop == 1e-6 and not c.time_accurate


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_block_planes_rule(precision):
    spec = cases.CaseSpec(precision=precision)
    assert serving.block_planes(spec) == (37 if precision == "mixed" else 30)
    spec.irs_beta = 0.5
    assert serving.block_planes(spec) == (41 if precision == "mixed" else 34)
    lam = cases.CaseSpec(precision=precision, model=1)
    assert serving.block_planes(lam) == (49 if precision == "mixed" else 39)
    lam.irs_beta = 0.5
    assert serving.block_planes(lam) == (53 if precision == "mixed" else 43)
