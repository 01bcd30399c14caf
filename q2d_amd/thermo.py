"""Property tables of FLINT's native tabulated mode, for the library's table mode (q2d_case_desc.thermo).

The reference reads one species' INPUT/thermo.dat (Tecplot POINT: Temperature, Cp, Enthalpy, Entropy,
dCp on integer kelvins from 1 K) and, for viscous cases, INPUT/transport.dat (Temperature, Viscosity,
Conductivity) through FLINT (Load_ThermoTransport.f90:15-71) and interpolates them linearly from node
int(T) (docs/theory/thermo.md:39-41).  The library takes four planes h, cp, mu, k with node k at T = k K,
k = 0 .. n-1: the files' nodes 1 .. n-1, and node 0 extrapolated linearly from nodes 1 and 2 (the files
start at 1 K; int(T) is 0 below it).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def read_point_table(path):
    """Numeric rows of a Tecplot POINT table (title, VARIABLES and ZONE lines skipped)."""
    rows = []
    for x in Path(path).read_text().splitlines():
        x = x.strip()
        if x and (x[0].isdigit() or x[0] in ".-+"):
            rows.append([float(v) for v in x.split()])
    return np.array(rows)


def planes(T, h, cp, mu=None, k=None):
    """(4, n) planes on T = 0 .. n-1 from file columns given on T = 1, 2, ... (unit spacing)."""
    T = np.asarray(T, dtype=np.float64)
    if T.size < 2 or T[0] != 1.0 or np.any(np.diff(T) != 1.0):
        raise ValueError("property tables must list T = 1, 2, 3, ... K")
    n = T.size + 1
    out = np.zeros((4, n))
    for q, col in enumerate((h, cp, mu, k)):
        if col is None:
            continue
        col = np.asarray(col, dtype=np.float64)
        if col.size < T.size:
            raise ValueError("property column shorter than its temperatures")
        out[q, 1:] = col[:T.size]
        out[q, 0] = 2.0 * out[q, 1] - out[q, 2]
    return out


def load_tables(thermo_dat, transport_dat=None):
    """The library's planes from a thermo.dat (and transport.dat, viscous cases).  Where the two files
    list different temperature ranges, the planes cover the shorter."""
    th = read_point_table(thermo_dat)
    T, cp, h = th[:, 0], th[:, 1], th[:, 2]
    mu = k = None
    if transport_dat is not None and Path(transport_dat).exists():
        tr = read_point_table(transport_dat)
        m = min(T.size, tr.shape[0])
        if np.any(tr[:m, 0] != T[:m]):
            raise ValueError("thermo.dat and transport.dat list different temperatures")
        T, cp, h, mu, k = T[:m], cp[:m], h[:m], tr[:m, 1], tr[:m, 2]
    return planes(T, h, cp, mu, k)


def constant_cp(cp, n, mu=0.0, k=0.0):
    """Planes of a calorically perfect gas written as a table (h = cp T, as the reference's Air tables)."""
    T = np.arange(1.0, n)
    return planes(T, cp * T, np.full(T.size, cp), np.full(T.size, mu), np.full(T.size, k))


def linear_cp(cp0, slope, n, mu=0.0, k=0.0):
    """A thermally perfect gas with cp(T) = cp0 + slope T: h(T) = cp0 T + slope T^2 / 2 at the nodes."""
    T = np.arange(1.0, n)
    return planes(T, cp0 * T + 0.5 * slope * T * T, cp0 + slope * T, np.full(T.size, mu), np.full(T.size, k))
