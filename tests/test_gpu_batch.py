"""Batched uneven cases in one persistent launch, and the tile autotuner (needs the GPU).

Every case of a batch is independent: its state, dt and residual after N steps equal those of the
same case run alone, bit for bit, in both arithmetic forms; the launch shape (rows per work item,
waves per workgroup) changes the schedule, never a cell's result."""
import numpy as np
import pytest

from q2d_amd import GpuSolver, cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_case_alone_equals_case_in_batch(precision):
    pb = cases.batch_uneven(seed=5, ncases=10, steps=40, precision=precision, smax=1.2)
    g = GpuSolver(pb)
    go = g.run(40)
    for k in range(len(pb.cases)):
        alone = GpuSolver(pb.case(k))
        ao = alone.run(40)[0]
        assert np.array_equal(alone.state(0)[2:-2, 2:-2], g.state(k)[2:-2, 2:-2]), k
        assert ao["dtglobal"] == go[k]["dtglobal"] and ao["time"] == go[k]["time"] and ao["iter"] == go[k]["iter"]


def test_batch_matches_oracle_fp64(oracle_mod):
    pb = cases.batch_uneven(seed=9, ncases=6, steps=25, precision="fp64", smax=0.8)
    g = GpuSolver(pb)
    g.run(25)
    o = oracle_mod.Oracle(pb)
    rc, _ = o.run(25, stop_on_todo=False)
    assert rc == 0
    for b in range(len(pb.blocks)):
        assert np.array_equal(g.state(b)[2:-2, 2:-2], o.state(b)[2:-2, 2:-2]), b


def test_tiling_does_not_change_results():
    pb = cases.batch_uneven(seed=3, ncases=6, steps=30, precision="mixed", smax=1.0)
    ref = None
    for rows, waves in ((8, 1), (16, 4), (64, 2), (-1, 4)):     # -1: occupancy-balanced rows per launch
        g = GpuSolver(pb)
        g.set_tiling(rows, waves)
        assert g.tiling() == (rows, waves)
        g.run(30)
        st = [g.state(b)[2:-2, 2:-2] for b in range(len(pb.blocks))]
        if ref is None:
            ref = st
        else:
            for a, b in zip(st, ref):
                assert np.array_equal(a, b)


def test_autotune_picks_and_caches(tmp_path, monkeypatch):
    from q2d_amd import autotune
    monkeypatch.setenv("Q2D_TUNE_CACHE", str(tmp_path / "tiles.json"))
    pb = cases.batch_uneven(seed=4, ncases=4, steps=10, precision="mixed", smax=0.6)
    r, w, t = autotune.tune(pb, steps=2, rows=(8, 32), waves=(2, 4))
    assert (r, w) in {(8, 2), (32, 2), (8, 4), (32, 4)} and len(t) == 4
    assert (tmp_path / "tiles.json").exists()
    r2, w2, _ = autotune.tune(pb, steps=2, rows=(8, 32), waves=(2, 4))
    assert (r2, w2) == (r, w)
    g = GpuSolver(pb)
    assert g.autotune(steps=2) == (r, w) and g.tiling() == (r, w)
    g.run(10)
