"""Closed-form pins of the oracle's boundary fluxes 405 / 406, the CFL ramp and the axisymmetric Euler source
(CPU).  Each check is taken from the physics the formula states, not from the formula, and each is shown
to fail when a mistake is planted in the oracle (oracle.planted).

* exact_pow (the `**` of Outflow) is within one ulp of the C library's pow.
* Outflow (Lib_BC_Fluxes_Outflow.f90:89-165): a supersonic exit, or no back pressure, gives the physical
  flux of the boundary state; a subsonic exit with back pressure gives the flux of a face state that
  (i) has the pressure max(p_exit, p*) relaxed by rel_fac toward the boundary pressure, p* the critical
  pressure of the normal Mach number, (ii) has the boundary cell's entropy p / rho^gamma, (iii) keeps the
  outgoing Riemann invariant a + sign (gamma-1)/2 Un and (iv) the tangential velocity; an inflow gives the
  symmetry fallback.
* Supersonic_Inflow (Lib_BC_Fluxes_Inflow.f90:318-403): when the imposed state and the boundary state
  are supersonic into the domain the HLL-family solvers upwind, so the flux is the physical flux of the
  imposed state: p = rel p_in + (1-rel) p_B, rho = p / (R T_in), |V| = (rel M + (1-rel) M_B) a, at the
  flow angle.  End to end, a uniform stream equal to the inflow state is preserved, inflow column included.
* CFL ramp (Mod_dt.f90:43): on a preserved stream the local and global dt of iteration k < ramp are the
  unramped dt times k / ramp.
* Source_Axi_Eul_Blk (Lib_Source_Axisymmetry.f90:40-78): for a uniform state the interior fluxes cancel,
  so one Euler stage's residual is the cylindrical geometric source: the Euler equations in (x, r) are
  d(rU)/dt + d(rF)/dx + d(rG)/dr = (0, 0, p, 0), so the planar residual gains
  -(rho v, rho u v, rho v^2, (rho E + p) v) / r per unit area.
"""
import math

import numpy as np
import pytest

from q2d_amd import cases
from q2d_amd.geometry import BlockGeometry, axisymmetric_geometry

R_GAS, CP = 0.8, 2.8                               # gamma = 1.4 with R != 1, so a dropped R shows
GAM = CP / (CP - R_GAS)


def phys_flux(rho, u, v, p, n1, n2, il, R=R_GAS, cp=CP):
    un = u * n1 + v * n2
    h0 = cp * p / (rho * R) + 0.5 * (u * u + v * v)
    return np.array([rho * un, rho * u * un + p * n1, rho * v * un + p * n2, rho * un * h0]) * il


def ulps(a, b):
    return abs(int(np.array(a).view(np.int64)) - int(np.array(b).view(np.int64)))


def test_exact_pow_within_one_ulp_of_libm(oracle_mod):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-7, 7, 20000))
    y = rng.uniform(-8, 8, 20000)
    y[::3] = 3.5                                     # gamma / (gamma - 1) at gamma = 1.4
    y[1::3] = 1.0 / 7.0                              # (gamma - 1) / (2 gamma)
    worst = max(ulps(oracle_mod.pow_(a, b), a ** b) for a, b in zip(x, y))
    assert worst <= 1
    assert oracle_mod.pow_(2.0, 10.0) == 1024.0 and oracle_mod.pow_(9.0, 0.5) == 3.0


def random_face(rng):
    a = rng.uniform(0, 2 * math.pi)
    return math.cos(a), math.sin(a), rng.uniform(0.5, 2.0)


def outflow_state(rng, n1, n2, face, mach_n):
    """A boundary state leaving through `face` with normal Mach number mach_n and a random tangential part."""
    rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
    a = math.sqrt(GAM * p / rho)
    sgn = 1 - 2 * (face % 2)
    un = sgn * mach_n * a
    vt = rng.uniform(-1, 1) * a
    return np.array([rho, un * n1 - vt * n2, un * n2 + vt * n1, p])


@pytest.mark.parametrize("face", [1, 2, 3, 4])
def test_outflow_supersonic_exit_is_physical_flux(oracle_mod, face):
    rng = np.random.default_rng(face)
    for _ in range(50):
        n1, n2, il = random_face(rng)
        Bp = outflow_state(rng, n1, n2, face, rng.uniform(1.05, 3.0))
        F = oracle_mod.outflow_flux(R_GAS, CP, face, Bp, n1, n2, il, rng.uniform(0.5, 2.0), rng.uniform(0.2, 1.0))
        assert np.allclose(F, phys_flux(*Bp, n1, n2, il), rtol=1e-14, atol=1e-14)
        F0 = oracle_mod.outflow_flux(R_GAS, CP, face, outflow_state(rng, n1, n2, face, 0.5), n1, n2, il, 0.0, 1.0)
        assert F0 is not None                        # p_exit <= 0: extrapolation, checked below
    Bp = outflow_state(rng, 0.6, 0.8, face, 0.4)
    assert np.allclose(oracle_mod.outflow_flux(R_GAS, CP, face, Bp, 0.6, 0.8, 1.3, 0.0, 1.0),
                       phys_flux(*Bp, 0.6, 0.8, 1.3), rtol=1e-14, atol=1e-14)
    Bp[1:3] = -Bp[1:3]                               # flow entering: BC_Symmetry_Eul fallback
    assert oracle_mod.outflow_flux(R_GAS, CP, face, Bp, 0.6, 0.8, 1.3, 1.0, 1.0) is None


def subsonic_outflow_violations(oracle_mod, face, seed):
    """Largest relative violation of the four characteristic conditions over random subsonic exits."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    sgn = 1 - 2 * (face % 2)
    for _ in range(60):
        n1, n2, il = random_face(rng)
        Bp = outflow_state(rng, n1, n2, face, rng.uniform(0.05, 0.95))
        rho_b, u_b, v_b, p_b = Bp
        a_b = math.sqrt(GAM * p_b / rho_b)
        un_b = u_b * n1 + v_b * n2
        mn = abs(un_b / a_b)
        pcrit = p_b * ((1 + 0.5 * (GAM - 1) * mn * mn) / (1 + 0.5 * (GAM - 1))) ** (GAM / (GAM - 1))
        p_exit = pcrit * rng.uniform(0.3, 2.0)
        rel = rng.uniform(0.1, 1.0)
        F = oracle_mod.outflow_flux(R_GAS, CP, face, Bp, n1, n2, il, p_exit, rel)
        pf = rel * max(p_exit, pcrit) + (1 - rel) * p_b
        # the face state from the flux, given its pressure
        uf = (F[1] - pf * il * n1) / F[0]
        vf = (F[2] - pf * il * n2) / F[0]
        unf = uf * n1 + vf * n2
        rhof = F[0] / (unf * il)
        af = math.sqrt(GAM * pf / rhof)
        checks = [
            abs(pf / rhof ** GAM - p_b / rho_b ** GAM) / (p_b / rho_b ** GAM),              # entropy
            abs((af + sgn * 0.5 * (GAM - 1) * unf) - (a_b + sgn * 0.5 * (GAM - 1) * un_b)) / a_b,   # invariant
            abs((-uf * n2 + vf * n1) - (-u_b * n2 + v_b * n1)) / a_b,                         # tangential
            abs(F[3] - F[0] * (CP * pf / (rhof * R_GAS) + 0.5 * (uf * uf + vf * vf))) / abs(F[3]),  # energy
        ]
        worst = max(worst, max(checks))
    return worst


@pytest.mark.parametrize("face", [1, 2, 3, 4])
def test_outflow_subsonic_characteristic_conditions(oracle_mod, face):
    assert subsonic_outflow_violations(oracle_mod, face, 10 + face) < 1e-11


@pytest.mark.parametrize("plant", ["outflow_riem", "outflow_isen"])
def test_outflow_checks_detect_planted_mistakes(oracle_mod, plant):
    with oracle_mod.planted(plant):
        assert subsonic_outflow_violations(oracle_mod, 2, 7) > 1e-4


def inflow_violation(oracle_mod, face, seed, riemann):
    rng = np.random.default_rng(seed)
    worst = 0.0
    sgn = 1 - 2 * (face % 2)
    for _ in range(60):
        n1, n2, il = random_face(rng)
        mach, T, p, rel = rng.uniform(1.5, 4.0), rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), rng.uniform(0.3, 1.0)
        alpha = math.atan2(n2, n1) + rng.uniform(-0.3, 0.3)       # within 17 deg of the face normal
        if sgn > 0:
            alpha += math.pi                                       # into the domain through faces 2 and 4
        Bp = np.empty(4)
        Bp[0], Bp[3] = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        ab = math.sqrt(GAM * Bp[3] / Bp[0])
        mb = rng.uniform(2.0, 3.0)
        Bp[1], Bp[2] = mb * ab * math.cos(alpha), mb * ab * math.sin(alpha)
        Ip = Bp * rng.uniform(0.98, 1.02, 4)
        F = oracle_mod.inflow_flux(1, riemann, R_GAS, CP, face, Bp, Ip, n1, n2, il, mach, T, p, rel, alpha)
        ps = rel * p + (1 - rel) * Bp[3]
        rhos = ps / (R_GAS * T)
        vs = (rel * mach + (1 - rel) * mb) * math.sqrt(GAM * ps / rhos)
        Fx = phys_flux(rhos, vs * math.cos(alpha), vs * math.sin(alpha), ps, n1, n2, il)
        worst = max(worst, float(np.abs(F - Fx).max() / np.abs(Fx).max()))
    return worst


@pytest.mark.parametrize("face", [1, 2, 3, 4])
@pytest.mark.parametrize("riemann", [0, 1])
def test_supersonic_inflow_upwinds_the_imposed_state(oracle_mod, face, riemann):
    assert inflow_violation(oracle_mod, face, 20 + face, riemann) < 1e-13


@pytest.mark.parametrize("plant", ["inflow_angle", "inflow_rho"])
def test_inflow_check_detects_planted_mistakes(oracle_mod, plant):
    with oracle_mod.planted(plant):
        assert inflow_violation(oracle_mod, 1, 5, 0) > 1e-3


def freestream(ramp_iter=0, time_accurate=False, rel_fac=1.0, steps=120):
    pb = cases.uniform_inflow(48, 16, 2.2, riemann="HLLC", precision="fp64", n_rk=3, steps=steps)
    pb.cases[0].time_accurate = time_accurate
    pb.cases[0].ramp_iter = ramp_iter
    pb.cases[0].gas_R, pb.cases[0].gas_cp = R_GAS, CP  # the stream rho = p = 1 at T = 1 / R
    for r in pb.bc:
        if r["type"] == 405:
            r["rel_fac"] = rel_fac
            r["T"] = 1.0 / R_GAS
    return pb


@pytest.mark.parametrize("rel_fac", [1.0, 0.5])
def test_supersonic_inflow_preserves_the_free_stream(oracle_mod, rel_fac):
    """BC 405 / 400 / 300 around a uniform stream equal to the inflow state: every cell, the inflow column
    included, keeps the stream to rounding after 40 SSP-RK3 steps."""
    pb = freestream(rel_fac=rel_fac, steps=40)
    P0 = pb.blocks[0].P[2:-2, 2:-2].copy()
    o = oracle_mod.Oracle(pb)
    rc, _ = o.run(40)
    assert rc == 0
    scale = np.abs(P0).max(axis=(0, 1))
    scale[2] = scale[1]                              # v = 0: measured against the stream speed
    d = np.abs(o.state(0)[2:-2, 2:-2] - P0).max(axis=(0, 1)) / scale
    assert d.max() < 1e-13, d
    with oracle_mod.planted("inflow_rho"):
        o = oracle_mod.Oracle(freestream(rel_fac=rel_fac, steps=5))
        o.run(5)
        assert np.abs(o.state(0)[2:-2, 2, 0] - P0[:, 0, 0]).max() > 1e-3      # the inflow column departs


@pytest.mark.parametrize("time_accurate", [False, True])
def test_cfl_ramp_scales_dt_by_iteration(oracle_mod, time_accurate):
    ramp = 100

    def dts(plant=None):
        o = oracle_mod.Oracle(freestream(ramp_iter=ramp, time_accurate=time_accurate))
        o0 = oracle_mod.Oracle(freestream(ramp_iter=0, time_accurate=time_accurate))
        out = []
        for k in range(1, 121):
            if plant:
                with oracle_mod.planted(plant):
                    rc, a = o.step()
            else:
                rc, a = o.step()
            rc0, b = o0.step()
            assert rc == 0 and rc0 == 0
            out.append((k, a[0]["dtglobal"], b[0]["dtglobal"], o.dtlocal(0), o0.dtlocal(0)))
        return out

    for k, dg, dg0, dl, dl0 in dts():
        f = k / ramp if k < ramp else 1.0
        assert abs(dg - dg0 * f) <= 1e-13 * dg0
        assert np.abs(dl - dl0 * f).max() <= 1e-13 * dl0.max()
    bad = [abs(dg - dg0 * (k / ramp if k < ramp else 1.0)) / dg0 for k, dg, dg0, _, _ in dts("ramp")]
    assert max(bad) > 1e-3


def axi_uniform(rho, u, v, p, ni=8, nj=8, r0=0.5, r1=1.5):
    node = cases.cartesian_nodes(ni, nj, 0.0, 1.0, r0, r1)
    geo = BlockGeometry(node)
    P = np.zeros((nj + 4, ni + 4, 4))
    P[:, :] = [rho, u, v, p]
    spec = cases.CaseSpec(n_rk=1, limiter="minmod", riemann="HLLC", source="axisymmetric", itermax=1, gas_R=R_GAS,
                          gas_cp=CP)
    blk = cases.Block(geo, P)
    pb = cases.Problem(blocks=[blk], bc=cases.face_records(1, ni, nj, {1: 400, 2: 400, 3: 400, 4: 400}), cases=[spec])
    cases.finish(pb, [geo])
    blk.src_a, blk.src_b = axisymmetric_geometry(geo)
    return pb, node


def axi_source_violation(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(6):
        rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        u, v = rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5)
        pb, node = axi_uniform(rho, u, v, p)
        o = oracle_mod.Oracle(pb)
        rc, _ = o.step()
        R = o.residual(0)[2:-2, 2:-2]
        r = 0.25 * (node[1:, 1:, 1] + node[1:, :-1, 1] + node[:-1, 1:, 1] + node[:-1, :-1, 1])
        sur = pb.blocks[0].geo.sur[2:-2, 2:-2]
        E = (CP - R_GAS) * p / (rho * R_GAS) + 0.5 * (u * u + v * v)
        S = -np.stack([rho * v / r, rho * u * v / r, rho * v * v / r, (rho * E + p) * v / r], -1)
        expect = -S * sur[..., None]
        sl = np.s_[1:-1, 1:-1]                       # cells whose four faces are interior faces
        worst = max(worst, float(np.abs(R[sl] - expect[sl]).max() / np.abs(expect[sl]).max()))
    return worst


def test_axisymmetric_euler_source_is_the_cylindrical_geometric_term(oracle_mod):
    assert axi_source_violation(oracle_mod, 3) < 1e-13


@pytest.mark.parametrize("plant", ["axi_sign", "axi_energy"])
def test_axisymmetric_source_check_detects_planted_mistakes(oracle_mod, plant):
    with oracle_mod.planted(plant):
        assert axi_source_violation(oracle_mod, 3) > 1e-3
