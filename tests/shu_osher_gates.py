"""The reference's gates of the Shu-Osher LIMO3 case (test/1D/ShuOsher-LIMO3/check.py:19-47, 100-138),
restated for tests: total variation of the density in the smooth-wave window x in [-3, 2.3] against the
committed limiter baselines, the density bounds and the main shock near x = 2.4."""
from pathlib import Path

import numpy as np

GOLD = Path(__file__).resolve().parent / "golden" / "shu_osher"
WIN_LO, WIN_HI = -3.0, 2.3
SHOCK_X, SHOCK_TOL = 2.4, 0.4


def total_variation(x, r, lo=WIN_LO, hi=WIN_HI):
    seg = r[(x >= lo) & (x <= hi)]
    return float(np.sum(np.abs(np.diff(seg))))


def baseline(tag):
    return np.loadtxt(GOLD / f"baseline_{tag}.dat")


def limo3_gates(xc, rho):
    """The failures of check.py's gates for a LIMO3 density profile (empty: all pass)."""
    tv = {t: total_variation(*baseline(t).T) for t in ("iord", "minmod", "vanleer", "mc")}
    tv_l = total_variation(xc, rho)
    bad = []
    if not (0.7 <= rho.min() and rho.max() <= 5.0):
        bad.append("density bounds")
    if abs(xc[int(np.argmax(np.abs(np.diff(rho))))] - SHOCK_X) > SHOCK_TOL:
        bad.append("shock position")
    if tv_l < 1.5 * tv["iord"]:
        bad.append("TV above 1.5x first order")
    if tv_l < tv["vanleer"] or tv["vanleer"] < tv["minmod"]:
        bad.append("TV ordering")
    if not (0.90 * tv["mc"] <= tv_l <= 1.10 * tv["mc"]):
        bad.append("TV within 10 % of MC")
    return bad
