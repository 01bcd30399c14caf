"""Oracle checks of implicit residual smoothing (Lib_IRS.f90:17-177) on the CPU.

The smoothing is pinned by what the mathematics fixes, each check chosen so that a
plausible slip fails it:
  * the Fourier symbol: on cos(k1 i + a) cos(k2 j + b), two Jacobi sweeps per
    direction multiply the interior by G(k1) G(k2), G(k) = (1 + 2 beta g cos k) / (1 + 2 beta),
    g = (1 + 2 beta cos k) / (1 + 2 beta) (one sweep, or the second sweep starting
    from R* instead of R, gives g or g^2);
  * constants are invariant under zero-gradient ghosts;
  * a field varying along j only is smoothed exactly as its transpose along i
    (direction and stencil of the second pass);
  * ghost rules: a block split by a connection 101 reproduces the unsplit block
    everywhere but the two columns (rows) at the cut, in both directions (donor
    residual, ghosts refreshed between the directions from the smoothed field); a
    periodic record 201 acts as zero-gradient (`case default`, :171-172);
  * in the explicit step, beta -> 0 reproduces the step without smoothing to
    rounding (scaling -R / sur * dt before the smoothing, update with it as is).
"""
import math

import numpy as np
import pytest

from q2d_amd import cases
from q2d_amd.geometry import BlockGeometry


def grid_problem(sizes, recs, irs_beta=0.0, **spec):
    """Blocks of the given (ni, nj) on unit-spaced Cartesian meshes, one case, with records `recs`."""
    blocks, geoms = [], []
    x0 = 0.0
    for ni, nj in sizes:
        node = cases.cartesian_nodes(ni, nj, x0, x0 + ni, 0.0, float(nj))
        x0 += ni
        geo = BlockGeometry(node)
        P = np.zeros((nj + 4, ni + 4, 4))
        P[..., 0] = 1.0
        P[..., 3] = 1.0
        blocks.append(cases.Block(geo, P))
        geoms.append(geo)
    sp = cases.CaseSpec(irs_beta=irs_beta, **spec)
    pb = cases.Problem(blocks=blocks, bc=recs, cases=[sp])
    return cases.finish(pb, geoms)


def zg_records(b, ni, nj, faces=(1, 2, 3, 4), code=400):
    return [r for r in cases.face_records(b, ni, nj, {f: code for f in (1, 2, 3, 4)}) if r["f"] in faces]


def sorted_recs(recs):
    return sorted(recs, key=lambda r: (r["b"], r["f"], r["j"] if r["f"] <= 2 else r["i"]))


def smooth(oracle_mod, pb, fields, beta):
    o = oracle_mod.Oracle(pb)
    for k, f in enumerate(fields):
        R = np.zeros((pb.blocks[k].nj + 4, pb.blocks[k].ni + 4, 4))
        R[2:-2, 2:-2] = f
        o.set_residual(k, R)
    o.irs_smooth(0, beta)
    out = [o.residual(k)[2:-2, 2:-2].copy() for k in range(len(fields))]
    o.close()
    return out


def symbol(beta, k):
    g = (1.0 + 2.0 * beta * math.cos(k)) / (1.0 + 2.0 * beta)
    return (1.0 + 2.0 * beta * g * math.cos(k)) / (1.0 + 2.0 * beta)


@pytest.mark.parametrize("beta", [0.3, 0.5, 2.0])
def test_fourier_symbol(oracle_mod, beta):
    ni, nj = 24, 18
    pb = grid_problem([(ni, nj)], zg_records(1, ni, nj))
    i = np.arange(1, ni + 1)[None, :]
    j = np.arange(1, nj + 1)[:, None]
    ks = [(0.7, 1.9), (2.5, 0.4), (math.pi / 2, 3.0), (1.1, 1.1)]
    f = np.zeros((nj, ni, 4))
    for v, (k1, k2) in enumerate(ks):
        f[..., v] = np.cos(k1 * i + 0.3 * v) * np.cos(k2 * j - 0.2 * v)
    (out,) = smooth(oracle_mod, pb, [f], beta)
    for v, (k1, k2) in enumerate(ks):
        want = symbol(beta, k1) * symbol(beta, k2) * f[2:-2, 2:-2, v]    # cells 3..n-2 in both directions
        np.testing.assert_allclose(out[2:-2, 2:-2, v], want, rtol=0, atol=1e-14)
    # the symbol separates one sweep (g) and a second sweep started from R* (g^2) from two sweeps
    k = 0.7
    g = (1 + 2 * beta * math.cos(k)) / (1 + 2 * beta)
    assert abs(symbol(beta, k) - g) > 1e-3 and abs(symbol(beta, k) - g * g) > 1e-3


def test_constant_invariant(oracle_mod):
    ni, nj = 10, 7
    pb = grid_problem([(ni, nj)], zg_records(1, ni, nj))
    f = np.broadcast_to(np.array([1.5, -0.25, 3.0, 1e-3]), (nj, ni, 4)).copy()
    (out,) = smooth(oracle_mod, pb, [f], 0.5)
    np.testing.assert_allclose(out, f, rtol=4e-16, atol=0)


def test_second_direction_is_j(oracle_mod):
    """A field varying along j only, smoothed, equals the transpose of its transpose smoothed."""
    rng = np.random.default_rng(3)
    ni, nj = 9, 13
    prof = rng.normal(size=(nj, 4))
    a = grid_problem([(ni, nj)], zg_records(1, ni, nj))
    (oa,) = smooth(oracle_mod, a, [np.broadcast_to(prof[:, None, :], (nj, ni, 4)).copy()], 0.5)
    b = grid_problem([(nj, ni)], zg_records(1, nj, ni))
    (ob,) = smooth(oracle_mod, b, [np.broadcast_to(prof[None, :, :], (ni, nj, 4)).copy()], 0.5)
    np.testing.assert_allclose(oa, ob.transpose(1, 0, 2), rtol=0, atol=1e-15)
    assert np.abs(oa - np.broadcast_to(prof[:, None, :], (nj, ni, 4))).max() > 1e-2   # it does smooth


def split_records(axis, n):
    """Two n x n blocks joined by 101 along `axis` (0: i, 1: j); every other face 400."""
    recs = []
    if axis == 0:
        recs += zg_records(1, n, n, (1, 3, 4)) + zg_records(2, n, n, (2, 3, 4))
        recs += [{"b": 1, "i": n, "j": j, "f": 2, "type": 101, "bs": 2, "is": 1, "js": j, "fs": 1, "case_id": 0}
                 for j in range(1, n + 1)]
        recs += [{"b": 2, "i": 1, "j": j, "f": 1, "type": 101, "bs": 1, "is": n, "js": j, "fs": 2, "case_id": 0}
                 for j in range(1, n + 1)]
    else:
        recs += zg_records(1, n, n, (1, 2, 3)) + zg_records(2, n, n, (1, 2, 4))
        recs += [{"b": 1, "i": i, "j": n, "f": 4, "type": 101, "bs": 2, "is": i, "js": 1, "fs": 3, "case_id": 0}
                 for i in range(1, n + 1)]
        recs += [{"b": 2, "i": i, "j": 1, "f": 3, "type": 101, "bs": 1, "is": i, "js": n, "fs": 4, "case_id": 0}
                 for i in range(1, n + 1)]
    return sorted_recs(recs)


@pytest.mark.parametrize("axis", [0, 1])
def test_connection_reproduces_unsplit(oracle_mod, axis):
    rng = np.random.default_rng(11 + axis)
    n = 8
    full_shape = (n, 2 * n) if axis == 0 else (2 * n, n)          # (nj, ni)
    f = rng.normal(size=full_shape + (4,))
    whole = grid_problem([(full_shape[1], full_shape[0])], zg_records(1, full_shape[1], full_shape[0]))
    (ow,) = smooth(oracle_mod, whole, [f], 0.5)
    split = grid_problem([(n, n), (n, n)], split_records(axis, n))
    parts = [f[:, :n], f[:, n:]] if axis == 0 else [f[:n], f[n:]]
    o1, o2 = smooth(oracle_mod, split, parts, 0.5)
    joined = np.concatenate([o1, o2], axis=1 - axis)
    cut = [n - 1, n]                                              # the cells either side of the cut
    keep = np.ones(2 * n, bool)
    keep[cut] = False
    if axis == 0:
        np.testing.assert_array_equal(joined[:, keep], ow[:, keep])
        assert np.abs(joined[:, cut] - ow[:, cut]).min() > 0      # ghosts are not smoothed (s1 = R there)
    else:
        np.testing.assert_array_equal(joined[keep], ow[keep])
        assert np.abs(joined[cut] - ow[cut]).min() > 0


def test_periodic_record_is_zero_gradient(oracle_mod):
    rng = np.random.default_rng(5)
    ni, nj = 8, 6
    f = rng.normal(size=(nj, ni, 4))
    per = cases.face_records(1, ni, nj, {1: 400, 2: 400, 3: 201, 4: 201})
    for r in per:
        if r["f"] >= 3:
            r.update(bs=1, **{"is": r["i"]}, js=nj if r["f"] == 3 else 1, fs=4 if r["f"] == 3 else 3)
    (op,) = smooth(oracle_mod, grid_problem([(ni, nj)], per), [f], 0.5)
    (oz,) = smooth(oracle_mod, grid_problem([(ni, nj)], zg_records(1, ni, nj)), [f], 0.5)
    np.testing.assert_array_equal(op, oz)
    con = [dict(r, type=101) if r["f"] >= 3 else r for r in per]   # the donor rule would differ
    (oc,) = smooth(oracle_mod, grid_problem([(ni, nj)], con), [f], 0.5)
    assert np.abs(oc - oz).max() > 1e-3


@pytest.mark.parametrize("n_rk,time_accurate", [(1, True), (3, False), (2, True)])
def test_step_beta_to_zero_is_plain_step(oracle_mod, n_rk, time_accurate):
    """beta = 1e-300: 1 + 2 beta == 1 and R + beta (...) == R for every residual above 1e-280, so the
    smoothed step equals the plain one but for residuals that are exactly zero, which gain ~1e-302
    (a zero velocity becomes 1e-302 and later picks an upwind side): the states agree to rounding.
    A dropped sign, a scaling applied twice or skipped moves them by the residual itself."""
    def run(beta):
        pb = cases.riemann_planar(seed=2, ni=40, nj=6, steps=3, n_rk=n_rk)
        pb.cases[0].irs_beta = beta
        pb.cases[0].time_accurate = time_accurate
        o = oracle_mod.Oracle(pb)
        rc, outs = o.run(3)
        assert rc == 0
        st = o.state(0)
        o.close()
        return st, outs
    s0, out0 = run(0.0)
    s1, out1 = run(1e-300)
    s5, _ = run(0.5)
    np.testing.assert_allclose(s1, s0, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(out1[0]["residual"], out0[0]["residual"], rtol=1e-10, atol=1e-14)
    assert np.abs(s5 - s0).max() > 1e-6                          # smoothing changes the step
