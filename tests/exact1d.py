"""Exact solutions the reference's 1D and axisymmetric V&V gates compare against (test helpers).

* riemann_exact: the exact solution of the Riemann problem of the 1D Euler equations for a perfect gas
  (Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4: the pressure function
  f(p) = f_L(p) + f_R(p) + (u_R - u_L), its root by Newton's method from the two-rarefaction guess,
  then the similarity solution sampled at (x - x0) / t).  Pinned by Toro's tabulated star state of the
  Sod problem (p* = 0.30313, u* = 0.92745) in tests/test_oracle_vv.py.
* noh_planar / noh_cylindrical: Noh's (1987) closed forms for the planar (m = 0) and cylindrical (m = 1)
  implosions: shock speed D = (gamma - 1) u0 / 2, post-shock rho0 ((gamma + 1) / (gamma - 1))^(m+1),
  u = 0, pre-shock rho0 (1 + u0 t / r)^m.
"""
from __future__ import annotations

import math

import numpy as np


def _fk(p, rk, pk, g):
    """Toro (4.6)-(4.7): the velocity jump across the left or right wave, and its derivative."""
    ak = math.sqrt(g * pk / rk)
    if p > pk:                                       # shock
        A = 2.0 / ((g + 1.0) * rk)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    e = (g - 1.0) / (2.0 * g)                        # rarefaction
    r = p / pk
    return 2.0 * ak / (g - 1.0) * (r ** e - 1.0), 1.0 / (rk * ak) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g=1.4, tol=1e-14):
    """p*, u* of the Riemann problem (Toro 4.5, Newton from the two-rarefaction approximation)."""
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    e = (g - 1.0) / (2.0 * g)
    p = ((al + ar - 0.5 * (g - 1.0) * (ur - ul)) / (al / pl ** e + ar / pr ** e)) ** (1.0 / e)
    p = max(p, 1e-14)
    for _ in range(200):
        fl, dl = _fk(p, rl, pl, g)
        fr, dr = _fk(p, rr, pr, g)
        pn = max(p - (fl + fr + ur - ul) / (dl + dr), 1e-14 * p)
        if abs(pn - p) <= tol * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, g)
    fr, _ = _fk(p, rr, pr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def riemann_exact(x, t, x0, rl, ul, pl, rr, ur, pr, g=1.4):
    """rho, u, p of the exact solution at the points x, time t (Toro 4.4 sampling)."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    out = np.empty((3, len(x)))
    for k, xv in enumerate(np.asarray(x, float)):
        s = (xv - x0) / t
        if s <= us:                                  # left of the contact
            if ps > pl:                              # left shock
                sl = ul - al * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                w = (rl, ul, pl) if s <= sl else (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:                                    # left rarefaction
                asl = al * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                if s <= ul - al:
                    w = (rl, ul, pl)
                elif s >= us - asl:
                    w = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) + gm / al * (ul - s)
                    w = (rl * c ** (2.0 / (g - 1.0)), 2.0 / (g + 1.0) * (al + 0.5 * (g - 1.0) * ul + s),
                         pl * c ** (2.0 * g / (g - 1.0)))
        else:                                        # right of the contact
            if ps > pr:                              # right shock
                sr = ur + ar * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                w = (rr, ur, pr) if s >= sr else (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:                                    # right rarefaction
                asr = ar * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                if s >= ur + ar:
                    w = (rr, ur, pr)
                elif s <= us + asr:
                    w = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) - gm / ar * (ur - s)
                    w = (rr * c ** (2.0 / (g - 1.0)), 2.0 / (g + 1.0) * (-ar + 0.5 * (g - 1.0) * ur + s),
                         pr * c ** (2.0 * g / (g - 1.0)))
        out[:, k] = w
    return out[0], out[1], out[2]


def noh_planar(x, t, x0=0.5, rho0=1.0, u0=1.0, g=1.4):
    """Planar Noh: two shocks at x0 -/+ D t; stagnation rho0 (g+1)/(g-1) between them, rho0 outside."""
    D = 0.5 * (g - 1.0) * u0
    inside = np.abs(np.asarray(x, float) - x0) < D * t
    return np.where(inside, rho0 * (g + 1.0) / (g - 1.0), rho0)


def noh_cylindrical(r, t, rho0=1.0, u0=1.0, g=1.4):
    """Cylindrical Noh: rho, v and p behind (r < D t) and ahead of the shock."""
    r = np.asarray(r, float)
    D = 0.5 * (g - 1.0) * u0
    inside = r < D * t
    rho = np.where(inside, rho0 * ((g + 1.0) / (g - 1.0)) ** 2, rho0 * (1.0 + u0 * t / np.maximum(r, 1e-300)))
    v = np.where(inside, 0.0, -u0)
    p = np.where(inside, rho0 * u0 * u0 * 0.5 * (g + 1.0) * (g + 1.0) / (g - 1.0), 0.0)
    return rho, v, p, D * t


def rel_l1(c, e):
    c, e = np.asarray(c, float), np.asarray(e, float)
    return float(np.abs(c - e).sum() / max(np.abs(e).sum(), 1e-300))
