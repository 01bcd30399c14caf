"""Setup-time geometry (Setup_Metrics restatement) against closed forms and invariants."""
import math

import numpy as np

from q2d_amd import cases
from q2d_amd.geometry import BlockGeometry, axisymmetric_geometry, ghost_metrics


def test_cartesian_closed_forms():
    ni, nj, dx, dy = 12, 7, 0.25, 0.5
    g = BlockGeometry(cases.cartesian_nodes(ni, nj, 0.0, ni * dx, 0.0, nj * dy))
    inner = (slice(2, -2), slice(2, -2))
    assert np.allclose(g.dl[inner][..., 0], dx, rtol=1e-15) and np.allclose(g.dl[inner][..., 1], dy, rtol=1e-15)
    assert np.allclose(g.sur[inner], dx * dy, rtol=1e-15)
    M = g.M[inner]
    assert np.allclose(M[..., 0], 1 / dx) and np.allclose(M[..., 3], 1 / dy)
    assert np.all(M[..., 1] == 0) and np.all(M[..., 2] == 0)
    assert np.allclose(g.fi[..., 0], 1.0) and np.allclose(g.fi[..., 1], 0.0) and np.allclose(g.fi[..., 2], dy)
    assert np.allclose(g.fj[..., 0], 0.0) and np.allclose(g.fj[..., 1], 1.0) and np.allclose(g.fj[..., 2], dx)


def curvilinear(ni=20, nj=14, seed=0):
    rng = np.random.default_rng(seed)
    node = cases.cartesian_nodes(ni, nj, 0.0, 1.0, 0.0, 0.7)
    x, y = node[..., 0].copy(), node[..., 1].copy()
    a, b = rng.uniform(0.02, 0.05, 2)
    node[..., 0] = x + a * np.sin(2 * math.pi * y) * np.sin(math.pi * x)
    node[..., 1] = y + b * np.sin(2 * math.pi * x)
    return node


def test_face_closure_and_area():
    """Sum over each cell's faces of N * int_l (outward) vanishes; sum of sur is the polygon area."""
    node = curvilinear()
    g = BlockGeometry(node)
    fi, fj = g.fi, g.fj
    Fi = fi[..., :2] * fi[..., 2:3]
    Fj = fj[..., :2] * fj[..., 2:3]
    net = Fi[:, 1:] - Fi[:, :-1] + Fj[1:, :] - Fj[:-1, :]
    assert np.abs(net).max() < 1e-15
    # shoelace area of the block outline
    ring = np.concatenate([node[0, :], node[1:, -1], node[-1, ::-1][1:], node[::-1, 0][1:-1]])
    area = 0.5 * abs(np.sum(ring[:, 0] * np.roll(ring[:, 1], -1) - np.roll(ring[:, 0], -1) * ring[:, 1]))
    assert abs(g.sur[2:-2, 2:-2].sum() - area) < 1e-12


def test_metric_is_inverse_transpose_jacobian():
    g = BlockGeometry(curvilinear())
    node = g.node
    N1, N2, N3, N4 = node[:-1, :-1], node[1:, :-1], node[:-1, 1:], node[1:, 1:]
    A1 = 0.5 * (((N4 - N2) + N3) - N1)
    A2 = 0.5 * (((N4 - N3) + N2) - N1)
    M = g.M[2:-2, 2:-2]
    # rows of M dotted with rows of A give the identity: sum_k M(d,k) A(e,k) = delta_de
    Mr1 = np.stack([M[..., 0], M[..., 2]], -1)
    Mr2 = np.stack([M[..., 1], M[..., 3]], -1)
    assert np.allclose((Mr1 * A1).sum(-1), 1.0) and np.allclose((Mr2 * A2).sum(-1), 1.0)
    assert np.allclose((Mr1 * A2).sum(-1), 0.0, atol=1e-12) and np.allclose((Mr2 * A1).sum(-1), 0.0, atol=1e-12)


def test_ghost_metrics_extrapolation_and_symmetry():
    ni, nj, dx = 10, 6, 0.1
    g = BlockGeometry(cases.cartesian_nodes(ni, nj, 0.0, ni * dx, 0.0, nj * dx))
    recs = cases.face_records(1, ni, nj, {1: 400, 2: 400, 3: 300, 4: 300})
    ghost_metrics([g], recs)
    # extrapolated (face 1/2) and mirrored (face 3/4) ghosts of a uniform mesh are uniform cells
    assert np.allclose(g.sur[2:-2, :2], dx * dx) and np.allclose(g.sur[2:-2, -2:], dx * dx)
    assert np.allclose(g.dl[:2, 2:-2], dx) and np.allclose(g.dl[-2:, 2:-2], dx)


def test_periodic_ghost_metrics_translate_donor():
    pb = cases.planar_shock(seed=0, n=16, steps=1)
    g1 = pb.blocks[0].geo
    # block 1 face 3 is periodic with block 3's top row: ghost cells equal the donor cells
    assert np.allclose(g1.sur[:2, 2:-2], pb.blocks[2].geo.sur[-4:-2, 2:-2][::-1])
    assert np.allclose(g1.dl[:2, 2:-2], pb.blocks[2].geo.dl[-4:-2, 2:-2][::-1])


def test_axisymmetric_radius():
    node = cases.cartesian_nodes(8, 4, 0.0, 1.0, 0.0, 0.4)
    g = BlockGeometry(node)
    r, inv = axisymmetric_geometry(g)
    assert np.allclose(r[0, :], 0.05) and np.allclose(r[:, 0], 0.05 + 0.1 * np.arange(4)) and np.allclose(inv, 1.0 / r)
