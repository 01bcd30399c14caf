"""Setup-time geometry of a structured block, restating Q2D's Setup_Metrics.

The Fortran host computes these arrays once at setup (src/lib/numerics/space/
Mod_Metrics.f90:11-148) and the explicit step only reads them.  This module
produces the same arrays, in the reference's allocation layout, for hosts that
do not run the Fortran setup (tests, benchmarks, the serving loop):

    dl   (nj+4, ni+4, 2)   dl(1-gc:ni+gc,1-gc:nj+gc)%c(2)
    M    (nj+4, ni+4, 4)   M(...)%c(2,2) column major: c11, c21, c12, c22
    sur  (nj+4, ni+4)
    fi   (nj, ni+1, 3)     dir(1)%f(0:ni,1:nj) {N(1), N(2), int_l}
    fj   (nj+1, ni, 3)     dir(2)%f(1:ni,0:nj)
    node (nj+1, ni+1, 2)

Operation order follows the Fortran (numpy evaluates elementwise in IEEE
double without contraction), so the arrays are those the reference computes.
"""
from __future__ import annotations

import numpy as np

GC = 2
# guide(6,3) — Parameters_m.f90:13-18 (faces 1..4 used)
GUIDE = {1: (1, 0), 2: (-1, 0), 3: (0, 1), 4: (0, -1)}


class MeshError(ValueError):
    pass


def metric_tensor(N1, N2, N3, N4):
    """Compute_Metric_Tensor — Lib_Metrics.f90:125-189 (vectorised).

    N1..N4 are (..., 2) arrays of the nodes (i-1,j-1), (i-1,j), (i,j-1), (i,j).
    Returns M (..., 4) as c11, c21, c12, c22, dl (..., 2) and sur (...).
    """
    A1 = ((N4 - N2) + N3) - N1
    A2 = ((N4 - N3) + N2) - N1
    A1 = 0.5 * A1
    A2 = 0.5 * A2
    det = A1[..., 0] * A2[..., 1] - A1[..., 1] * A2[..., 0]
    if np.any(np.abs(det) == 0.0):
        raise MeshError("Compute_Metric_Tensor: jacobian det=0")
    M = np.empty(det.shape + (4,))
    M[..., 0] = A2[..., 1] / det          # c(1,1) =  A(2,2)/det
    M[..., 1] = (-A1[..., 1]) / det       # c(2,1) = -A(1,2)/det
    M[..., 2] = (-A2[..., 0]) / det       # c(1,2) = -A(2,1)/det
    M[..., 3] = A1[..., 0] / det          # c(2,2) =  A(1,1)/det
    dl = np.empty(det.shape + (2,))
    dl[..., 0] = np.sqrt(A1[..., 0] * A1[..., 0] + A1[..., 1] * A1[..., 1])
    dl[..., 1] = np.sqrt(A2[..., 0] * A2[..., 0] + A2[..., 1] * A2[..., 1])
    d1 = N4 - N1
    d2 = N3 - N2
    cross = d1[..., 0] * d2[..., 1] - d1[..., 1] * d2[..., 0]
    sur = 0.5 * np.abs(cross)
    if np.any(sur <= 0.0):
        raise MeshError("Compute_Metric_Tensor: negative cell surface")
    return M, dl, sur


def _sign(a, b):
    return abs(a) if b >= 0 else -abs(a)


def norm_area(node):
    """Compute_Norm_Area — Lib_Metrics.f90:41-122. node is (nj+1, ni+1, 2)."""
    nj, ni = node.shape[0] - 1, node.shape[1] - 1
    X = lambda i, j: node[j, i]
    d1 = X(1, 1) - X(1, 0)
    snixx, sniyy = d1[1], -d1[0]
    d1 = 0.5 * (X(1, 1) + X(1, 0))
    d2 = 0.5 * (X(0, 1) + X(0, 0))
    d3 = d1 - d2
    signi = _sign(1.0, d3[0] * snixx + d3[1] * sniyy)
    d1 = X(0, 1) - X(1, 1)
    snjxx, snjyy = d1[1], -d1[0]
    d1 = 0.5 * (X(1, 1) + X(0, 1))
    d2 = 0.5 * (X(1, 0) + X(0, 0))
    d3 = d1 - d2
    signj = _sign(1.0, d3[0] * snjxx + d3[1] * snjyy)

    fi = np.empty((nj, ni + 1, 3))
    d1 = node[1:, :, :] - node[:-1, :, :]                 # node(i,j) - node(i,j-1), i=0..ni, j=1..nj
    d3x, d3y = d1[..., 1], -d1[..., 0]
    Ai = np.sqrt(d3x * d3x + d3y * d3y)
    with np.errstate(invalid="ignore", divide="ignore"):
        fi[..., 0] = np.where(Ai == 0.0, 0.0, d3x / Ai * signi)
        fi[..., 1] = np.where(Ai == 0.0, 0.0, d3y / Ai * signi)
    fi[..., 2] = Ai
    fj = np.empty((nj + 1, ni, 3))
    d1 = node[:, :-1, :] - node[:, 1:, :]                 # node(i-1,j) - node(i,j), i=1..ni, j=0..nj
    d3x, d3y = d1[..., 1], -d1[..., 0]
    Aj = np.sqrt(d3x * d3x + d3y * d3y)
    with np.errstate(invalid="ignore", divide="ignore"):
        fj[..., 0] = np.where(Aj == 0.0, 0.0, d3x / Aj * signj)
        fj[..., 1] = np.where(Aj == 0.0, 0.0, d3y / Aj * signj)
    fj[..., 2] = Aj
    return fi, fj


class BlockGeometry:
    """Geometry arrays of one block (reference layout, ghost extents)."""

    def __init__(self, node):
        node = np.ascontiguousarray(node, dtype=np.float64)
        if node.ndim != 3 or node.shape[2] != 2:
            raise MeshError("node must be (nj+1, ni+1, 2)")
        self.node = node
        self.nj, self.ni = node.shape[0] - 1, node.shape[1] - 1
        ni, nj = self.ni, self.nj
        self.dl = np.zeros((nj + 4, ni + 4, 2))
        self.M = np.zeros((nj + 4, ni + 4, 4))
        self.sur = np.zeros((nj + 4, ni + 4))
        N1 = node[:-1, :-1]
        N2 = node[1:, :-1]
        N3 = node[:-1, 1:]
        N4 = node[1:, 1:]
        M, dl, sur = metric_tensor(N1, N2, N3, N4)
        self.M[2:-2, 2:-2] = M
        self.dl[2:-2, 2:-2] = dl
        self.sur[2:-2, 2:-2] = sur
        self.fi, self.fj = norm_area(node)

    # cell (i,j) in Fortran indices -> array index
    @staticmethod
    def ix(i, j):
        return j + 1, i + 1

    def nd(self, i, j):
        return self.node[j, i]

    def set_cell(self, i, j, N1, N2, N3, N4):
        M, dl, sur = metric_tensor(np.asarray(N1), np.asarray(N2), np.asarray(N3), np.asarray(N4))
        jj, ii = self.ix(i, j)
        self.M[jj, ii] = M
        self.dl[jj, ii] = dl
        self.sur[jj, ii] = sur
        return M, dl, sur


def extrapolate_metrics(g: BlockGeometry, Im, Jm, Fm):
    """BC_Extrapolate_Metrics — Lib_Metrics.f90:285-395."""
    nd = g.nd
    if Fm == 1:
        N1 = {k: nd(Im + k - 1, Jm - 1).copy() for k in range(0, 3)}
        N2 = {k: nd(Im + k - 1, Jm).copy() for k in range(0, 3)}
        for k in (-1, -2):
            N1[k] = 3.0 * N1[k + 1] - 3.0 * N1[k + 2] + N1[k + 3]
            N2[k] = 3.0 * N2[k + 1] - 3.0 * N2[k + 2] + N2[k + 3]
            g.set_cell(Im + k, Jm, N1[k], N2[k], N1[k + 1], N2[k + 1])
    elif Fm == 2:
        N3 = {k: nd(Im + k, Jm - 1).copy() for k in range(0, -3, -1)}
        N4 = {k: nd(Im + k, Jm).copy() for k in range(0, -3, -1)}
        for k in (1, 2):
            N3[k] = 3.0 * N3[k - 1] - 3.0 * N3[k - 2] + N3[k - 3]
            N4[k] = 3.0 * N4[k - 1] - 3.0 * N4[k - 2] + N4[k - 3]
            g.set_cell(Im + k, Jm, N3[k - 1], N4[k - 1], N3[k], N4[k])
    elif Fm == 3:
        if g.nj > 2:
            N1 = {k: nd(Im - 1, Jm + k - 1).copy() for k in range(0, 3)}
            N3 = {k: nd(Im, Jm + k - 1).copy() for k in range(0, 3)}
            for k in (-1, -2):
                N1[k] = 3.0 * N1[k + 1] - 3.0 * N1[k + 2] + N1[k + 3]
                N3[k] = 3.0 * N3[k + 1] - 3.0 * N3[k + 2] + N3[k + 3]
                g.set_cell(Im, Jm + k, N1[k], N1[k + 1], N3[k], N3[k + 1])
        else:
            N1 = {k: nd(Im - 1, Jm + k - 1).copy() for k in range(0, 2)}
            N3 = {k: nd(Im, Jm + k - 1).copy() for k in range(0, 2)}
            for k in (-1, -2):
                N1[k] = 2.0 * N1[k + 1] - N1[k + 2]
                N3[k] = 2.0 * N3[k + 1] - N3[k + 2]
                g.set_cell(Im, Jm + k, N1[k], N1[k + 1], N3[k], N3[k + 1])
    elif Fm == 4:
        if g.nj > 2:
            N2 = {k: nd(Im - 1, Jm + k).copy() for k in range(0, -3, -1)}
            N4 = {k: nd(Im, Jm + k).copy() for k in range(0, -3, -1)}
            for k in (1, 2):
                N2[k] = 3.0 * N2[k - 1] - 3.0 * N2[k - 2] + N2[k - 3]
                N4[k] = 3.0 * N4[k - 1] - 3.0 * N4[k - 2] + N4[k - 3]
                g.set_cell(Im, Jm + k, N2[k - 1], N2[k], N4[k - 1], N4[k])
        else:
            N2 = {k: nd(Im - 1, Jm + k).copy() for k in range(0, -2, -1)}
            N4 = {k: nd(Im, Jm + k).copy() for k in range(0, -2, -1)}
            for k in (1, 2):
                N2[k] = 2.0 * N2[k - 1] - N2[k - 2]
                N4[k] = 2.0 * N4[k - 1] - N4[k - 2]
                g.set_cell(Im, Jm + k, N2[k - 1], N2[k], N4[k - 1], N4[k])


def symmetry_metrics(g: BlockGeometry, Im, Jm, Fm):
    """BC_Symmetry_Metrics — Lib_Metrics.f90:398-441 (mirror copy)."""
    for k in (1, 2):
        if Fm == 1:
            dst, src = (Im - k, Jm), (Im + k - 1, Jm)
        elif Fm == 2:
            dst, src = (Im + k, Jm), (Im - k + 1, Jm)
        elif Fm == 3:
            dst, src = (Im, Jm - k), (Im, Jm + k - 1)
        else:
            dst, src = (Im, Jm + k), (Im, Jm - k + 1)
        dj, di = g.ix(*dst)
        sj, si = g.ix(*src)
        g.M[dj, di] = g.M[sj, si]
        g.dl[dj, di] = g.dl[sj, si]
        g.sur[dj, di] = g.sur[sj, si]


_GUIDES = {1: (0, 0), 2: (-1, 0), 3: (0, 0), 4: (0, -1)}
_GUIDES2 = {1: (1, 0), 2: (-2, 0), 3: (0, 1), 4: (0, -2)}


def connect_metrics(gm: BlockGeometry, Im, Jm, Fm, gs: BlockGeometry, Is, Js, Fs, d11, d21, periodic=False):
    """BC_Connect_Metrics (101) and BC_Connect_Metrics_Periodic (201) — Lib_Metrics.f90:444-728."""
    N = {1: {}, 2: {}, 3: {}, 4: {}}
    nd = gm.nd
    if Fm == 1:
        N[1][0] = nd(Im - 1, Jm - 1); N[2][0] = nd(Im - 1, Jm)
    elif Fm == 2:
        N[3][0] = nd(Im, Jm - 1); N[4][0] = nd(Im, Jm)
    elif Fm == 3:
        N[1][0] = nd(Im - 1, Jm - 1); N[3][0] = nd(Im, Jm - 1)
    else:
        N[2][0] = nd(Im - 1, Jm); N[4][0] = nd(Im, Jm)
    main_slot = {1: 2, 2: 4, 3: 3, 4: 4}[Fm]
    diag_slot = {1: 1, 2: 3, 3: 1, 4: 2}[Fm]

    def donor_nodes(g):
        Is1 = Is + (GC - g) * _GUIDES[Fs][0] + (g - 1) * _GUIDES2[Fs][0]
        Js1 = Js + (GC - g) * _GUIDES[Fs][1] + (g - 1) * _GUIDES2[Fs][1]
        if Fs in (1, 2):
            II, JJ = Is1, (2 * Js1 + d11 + d21) // 2
            III, JJJ = II, JJ - d11
        else:
            II, JJ = (2 * Is1 + d11 + d21) // 2, Js1
            III, JJJ = II - d11, JJ
        return (II, JJ), (III, JJJ)

    T = np.zeros(2)
    if periodic:
        (II0, JJ0), (III0, JJJ0) = donor_nodes(0)
        T = N[main_slot][0] - gs.nd(II0, JJ0)
        T2 = N[diag_slot][0] - gs.nd(III0, JJJ0)
        if np.max(np.abs(T - T2)) > 1e-10 * (1.0 + np.max(np.abs(T))):
            raise MeshError("BC_Connect_Metrics_Periodic: non-rigid periodic seam")
    for g in (1, 2):
        (II, JJ), (III, JJJ) = donor_nodes(g)
        N[main_slot][g] = gs.nd(II, JJ) + T if periodic else gs.nd(II, JJ).copy()
        N[diag_slot][g] = gs.nd(III, JJJ) + T if periodic else gs.nd(III, JJJ).copy()
        if Fm == 1:
            gm.set_cell(Im - g, Jm, N[1][g], N[2][g], N[1][g - 1], N[2][g - 1])
        elif Fm == 2:
            gm.set_cell(Im + g, Jm, N[3][g - 1], N[4][g - 1], N[3][g], N[4][g])
        elif Fm == 3:
            gm.set_cell(Im, Jm - g, N[1][g], N[1][g - 1], N[3][g], N[3][g - 1])
        else:
            gm.set_cell(Im, Jm + g, N[2][g - 1], N[2][g], N[4][g - 1], N[4][g])


def ghost_metrics(geoms, bc_records):
    """Ghost-cell metrics of every boundary record (Mod_Metrics.f90:106-146).

    geoms: list of BlockGeometry (1-based block numbers in records);
    bc_records: iterable of dicts with b, i, j, f, type (+ bs, is, js, fs, d11, d21 for 101/201).
    """
    for r in bc_records:
        gm = geoms[r["b"] - 1]
        t = r["type"]
        if t in (101, 201):
            gs = geoms[r["bs"] - 1]
            connect_metrics(gm, r["i"], r["j"], r["f"], gs, r["is"], r["js"], r["fs"], r.get("d11", 1), r.get("d21", 0),
                            periodic=(t == 201))
        elif t == 300:
            symmetry_metrics(gm, r["i"], r["j"], r["f"])
        else:
            extrapolate_metrics(gm, r["i"], r["j"], r["f"])


def axisymmetric_geometry(g: BlockGeometry):
    """Precompute_Axisymmetric_Geometry — Lib_Source_Axisymmetry.f90:19-37: (r_ave, 1/max(r_ave,eps))."""
    y = g.node[..., 1]
    r_ave = 0.25 * (y[1:, 1:] + y[1:, :-1] + y[:-1, 1:] + y[:-1, :-1])
    inv_r = 1.0 / np.maximum(r_ave, 1e-12)
    return np.ascontiguousarray(r_ave), np.ascontiguousarray(inv_r)


def area_geometry(g: BlockGeometry, cross_section):
    """Precompute_Area_Geometry — Lib_Source_Areavariation.f90:26-50: (G_x, G_y) at cell centres.

    cross_section is the nodal A(0:ni,0:nj) as an (nj+1, ni+1) array."""
    A = np.asarray(cross_section, dtype=np.float64)
    a_ij = A[1:, 1:]; a_im = A[1:, :-1]; a_jm = A[:-1, 1:]; a_mm = A[:-1, :-1]
    dA1 = 0.5 * ((a_ij - a_im) + (a_jm - a_mm))
    dA2 = 0.5 * ((a_ij - a_jm) + (a_im - a_mm))
    M = g.M[2:-2, 2:-2]
    # dA = Matmul(dA_ij, M): dA(k) = dA_ij(1)*M(1,k) + dA_ij(2)*M(2,k)
    dAx = dA1 * M[..., 0] + dA2 * M[..., 1]
    dAy = dA1 * M[..., 2] + dA2 * M[..., 3]
    Ac = 0.25 * (a_ij + a_im + a_jm + a_mm)
    Gx = -dAx / np.maximum(Ac, 1e-12)
    Gy = -dAy / np.maximum(Ac, 1e-12)
    return np.ascontiguousarray(Gx), np.ascontiguousarray(Gy)
