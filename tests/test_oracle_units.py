"""Oracle unit checks against closed forms and invariants (CPU).

Each check is chosen so that a plausible slip (a dropped term, a swapped
argument, a wrong sign or index) fails: textbook limiter forms written
independently (Sweby), flux consistency F(U,U) = F_phys(U), upwinding of
supersonic states, exact zero mass flux through a stationary contact (HLLC),
rotational invariance, the constant-cp Roe sound speed, and the temporal
order of the SSP-RK schemes.
"""
import math

import numpy as np
import pytest

from q2d_amd import cases

LIM = {"iord": 0, "minmod": 1, "vanleer": 2, "vanalbada": 3, "mc": 4, "superbee": 5, "LIMO3": 6}


def textbook(name, a, b):
    """Limited slope from the downstream slope a and the upstream slope b (r = b/a), Sweby forms."""
    if name == "minmod":
        return 0.5 * (np.sign(a) + np.sign(b)) * min(abs(a), abs(b))
    if name == "iord":
        return 0.0
    r = b / a
    if name == "vanleer":
        return a * (r + abs(r)) / (1 + abs(r))
    if name == "vanalbada":
        return a * (r * r + r) / (1 + r * r)
    if name == "mc":
        return a * max(0.0, min(2 * r, 0.5 * (1 + r), 2.0))
    if name == "superbee":
        return a * max(0.0, min(2 * r, 1.0), min(r, 2.0))
    if name == "LIMO3":
        t = (2 + r) / 3
        return a * max(0.0, min(t, max(-0.5 * r, 0.0, min(2 * r, t, 1.6))))
    raise KeyError(name)


@pytest.mark.parametrize("name", list(LIM))
def test_limiters_match_textbook(oracle_mod, name):
    o = oracle_mod.load()
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.normal(size=300), [1.0, -1.0, 2.0, 0.5, -3.0, 1e-3]])
    for a in vals:
        for b in vals[::7]:
            got = o.q2do_limiter(LIM[name], float(a), float(b))
            want = textbook(name, float(a), float(b))
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (name, a, b, got, want)


@pytest.mark.parametrize("name", ["minmod", "vanleer", "vanalbada", "mc", "superbee"])
def test_limiters_second_order_on_linear_data(oracle_mod, name):
    """phi(1) = 1: equal slopes pass through unchanged."""
    o = oracle_mod.load()
    for s in (0.3, -2.5, 7.0):
        assert abs(o.q2do_limiter(LIM[name], s, s) - s) <= 1e-14 * abs(s)


def test_reconstruction_exact_on_linear(oracle_mod):
    import ctypes as C
    o = oracle_mod.load()
    pl, pr = C.c_double(), C.c_double()
    dx = 0.1
    q = [1.0 + 2.0 * dx * k for k in range(4)]        # cells -1, 0, 1, 2
    for lim in (1, 2, 4, 5):
        o.q2do_reconstruct(lim, *q, dx, dx, dx, dx / 2, dx / 2, 1.0, C.byref(pl), C.byref(pr))
        face = 1.0 + 2.0 * dx * 1.5
        assert abs(pl.value - face) < 1e-14 and abs(pr.value - face) < 1e-14


def phys_flux(U, n, R, cp):
    rho, u, v, p = U
    g = cp / (cp - R)
    un = u * n[0] + v * n[1]
    H = g / (g - 1) * p / rho + 0.5 * (u * u + v * v)
    return np.array([rho * un, rho * un * u + p * n[0], rho * un * v + p * n[1], rho * un * H])


def rand_state(rng):
    return np.array([rng.uniform(0.2, 2.0), rng.normal(), rng.normal(), rng.uniform(0.2, 2.0)])


@pytest.mark.parametrize("rid", [0, 1])
def test_riemann_consistency(oracle_mod, rid):
    o = oracle_mod.Oracle.__new__(oracle_mod.Oracle)
    o.lib = oracle_mod.load()
    rng = np.random.default_rng(1)
    for _ in range(200):
        U = rand_state(rng)
        th = rng.uniform(0, 2 * math.pi)
        n = (math.cos(th), math.sin(th))
        F = o.riemann(rid, 1.0, 3.5, U, U, *n)
        np.testing.assert_allclose(F, phys_flux(U, n, 1.0, 3.5), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("rid", [0, 1])
def test_riemann_supersonic_upwind(oracle_mod, rid):
    o = oracle_mod.Oracle.__new__(oracle_mod.Oracle)
    o.lib = oracle_mod.load()
    L = np.array([1.0, 5.0, 0.3, 1.0])       # Mach ~4 to the right
    R = np.array([0.5, 4.0, -0.2, 0.7])
    F = o.riemann(rid, 1.0, 3.5, L, R, 1.0, 0.0)
    np.testing.assert_allclose(F, phys_flux(L, (1.0, 0.0), 1.0, 3.5), rtol=1e-13)
    Lm, Rm = L.copy(), R.copy()
    Lm[1], Rm[1] = -4.0, -5.0
    F = o.riemann(rid, 1.0, 3.5, Lm, Rm, 1.0, 0.0)
    np.testing.assert_allclose(F, phys_flux(Rm, (1.0, 0.0), 1.0, 3.5), rtol=1e-13)


def test_hllc_stationary_contact(oracle_mod):
    """HLLC resolves a stationary contact exactly: no mass flux, pressure-only momentum flux."""
    o = oracle_mod.Oracle.__new__(oracle_mod.Oracle)
    o.lib = oracle_mod.load()
    L = np.array([1.0, 0.0, 0.0, 0.8])
    R = np.array([0.125, 0.0, 0.0, 0.8])
    F = o.riemann(0, 1.0, 3.5, L, R, 1.0, 0.0)
    assert abs(F[0]) < 1e-15 and abs(F[1] - 0.8) < 1e-15 and abs(F[3]) < 1e-15
    F = o.riemann(1, 1.0, 3.5, L, R, 1.0, 0.0)        # HLLE smears it: nonzero mass flux
    assert abs(F[0]) > 1e-3


@pytest.mark.parametrize("rid", [0, 1])
def test_riemann_rotational_invariance(oracle_mod, rid):
    o = oracle_mod.Oracle.__new__(oracle_mod.Oracle)
    o.lib = oracle_mod.load()
    rng = np.random.default_rng(3)
    for _ in range(50):
        L, R = rand_state(rng), rand_state(rng)
        F = o.riemann(rid, 1.0, 3.5, L, R, 1.0, 0.0)
        th = rng.uniform(0, 2 * math.pi)
        c, s = math.cos(th), math.sin(th)
        rot = lambda U: np.array([U[0], c * U[1] - s * U[2], s * U[1] + c * U[2], U[3]])
        Fr = o.riemann(rid, 1.0, 3.5, rot(L), rot(R), c, s)
        np.testing.assert_allclose(Fr, rot(F), rtol=1e-11, atol=1e-12)


def test_roe_sound_speed_constant_cp(oracle_mod):
    """For constant cp, a_roe = sqrt((gamma-1)(H_roe - |u_roe|^2/2))."""
    import ctypes as C
    lib = oracle_mod.load()
    rng = np.random.default_rng(5)
    R, cp = 287.0, 1004.5
    g = cp / (cp - R)
    for _ in range(100):
        L, Rs = rand_state(rng) * [1, 300, 300, 1e5], rand_state(rng) * [1, 300, 300, 1e5]
        out = np.zeros(7)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        L = np.ascontiguousarray(L); Rs = np.ascontiguousarray(Rs)
        lib.q2do_roe_averages(R, cp, dp(L), dp(Rs), dp(out))
        h0l, h0r, rho, u, v, a, h0 = out
        sl, sr = math.sqrt(L[0]), math.sqrt(Rs[0])
        Ht = (sl * (g / (g - 1) * L[3] / L[0] + 0.5 * (L[1] ** 2 + L[2] ** 2)) +
              sr * (g / (g - 1) * Rs[3] / Rs[0] + 0.5 * (Rs[1] ** 2 + Rs[2] ** 2))) / (sl + sr)
        ut = (sl * L[1] + sr * Rs[1]) / (sl + sr)
        vt = (sl * L[2] + sr * Rs[2]) / (sl + sr)
        assert abs(h0 - Ht) <= 1e-12 * Ht
        assert abs(a - math.sqrt((g - 1) * (Ht - 0.5 * (ut * ut + vt * vt)))) <= 1e-9 * a


def source_ode_case(cfl, n_rk, steps, g0=-0.8):
    """Uniform state with a uniform area-variation source G = (g0, 0) on every cell.

    Interior face fluxes cancel exactly (identical faces, zero slopes), so each interior cell
    integrates dU/dt = (u*g0) * [rho, rho u, rho v, rho E + p] with the SSP-RK scheme under test."""
    pb = cases.uniform_inflow(64, 8, 1.2, riemann="HLLC", area_a=0.5, precision="fp64", n_rk=n_rk, steps=steps)
    for r in pb.bc:
        r["type"] = 400
    b = pb.blocks[0]
    b.src_a = np.full((b.nj, b.ni), g0)
    b.src_b = np.zeros((b.nj, b.ni))
    pb.cases[0].cfl = cfl
    return pb


def exact_source_ode(U0, g0, t, gam=1.4):
    from scipy.integrate import solve_ivp

    def f(_, U):
        rho, m, n, E = U
        u = m / rho
        p = (gam - 1.0) * (E - 0.5 * (m * m + n * n) / rho)
        return (u * g0) * np.array([rho, m, n, E + p])
    sol = solve_ivp(f, (0.0, t), U0, method="DOP853", rtol=1e-13, atol=1e-15)
    return sol.y[:, -1]


@pytest.mark.parametrize("n_rk,order", [(1, 1), (2, 2), (3, 3)])
def test_rk_temporal_order(oracle_mod, n_rk, order):
    """SSP-RK1/2/3 (Lib_RK.f90:8-27 with Compute_Cons_NoIRS, Lib_Newstate.f90:131-159): the error of a
    smooth ODE integrated with the oracle's own dt sequence falls as dt^order."""
    errs = []
    for cfl, steps in ((0.4, 12), (0.2, 24)):
        pb = source_ode_case(cfl, n_rk, steps)
        P0 = pb.blocks[0].P[6, 34].copy()
        gam = 1.4
        U0 = np.array([P0[0], P0[0] * P0[1], P0[0] * P0[2], P0[3] / (gam - 1) + 0.5 * P0[0] * (P0[1] ** 2 + P0[2] ** 2)])
        o = oracle_mod.Oracle(pb)
        t = 0.0
        for _ in range(steps):
            rc, out = o.step()
            assert rc == 0
        t = out[0]["time"]
        P = o.state(0)[6, 34]
        U = np.array([P[0], P[0] * P[1], P[0] * P[2], P[3] / (gam - 1) + 0.5 * P[0] * (P[1] ** 2 + P[2] ** 2)])
        errs.append(np.abs(U - exact_source_ode(U0, -0.8, t)).max())
    p = math.log(errs[0] / errs[1], 2)
    assert abs(p - order) < 0.3, (n_rk, errs, p)
