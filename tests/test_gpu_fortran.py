"""The Fortran host adapter (module q2d_gpu, ISO_C_BINDING) drives the device step (needs the GPU).

q2d_fdriver reads arrays in the reference's layout, passes c_loc of them through the bind(C) types and
interfaces of q2d_amd/fortran/q2d_gpu_mod.f90, and runs the step; its states must equal the Python
host's (same library, same problem) and the oracle's, bit for bit."""
import subprocess

import numpy as np
import pytest

from q2d_amd import GpuSolver, cases
from q2d_amd.fortran_io import DRIVER, read_output, write_case_file

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built (make -C q2d_amd/fortran)")
def test_fortran_host_matches_python_host_and_oracle(oracle_mod, tmp_path):
    pb = cases.planar_shock(seed=6, n=48, steps=30)
    pb.cases[0].irs_beta = 0.5                       # every field of the scheme crosses the boundary
    pb.add(cases.axisymmetric_pipe(seed=2, ni=64, nj=24, first_cell=5e-4, reynolds=1e4, precision="fp64", steps=30))
    write_case_file(pb, tmp_path / "case.bin", 30)
    r = subprocess.run([str(DRIVER), str(tmp_path / "case.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    states, outs = read_output(tmp_path / "out.bin", pb)
    g = GpuSolver(pb)
    go = g.run(30)
    for b in range(len(pb.blocks)):
        assert np.array_equal(states[b][2:-2, 2:-2], g.state(b)[2:-2, 2:-2]), b
    for k in range(len(pb.cases)):
        assert outs[k]["iter"] == go[k]["iter"] == 30
        assert outs[k]["residual"] == go[k]["residual"]
    o = oracle_mod.Oracle(pb)
    rc, _ = o.run(30, stop_on_todo=False)
    assert rc == 0
    for b in range(len(pb.blocks)):
        assert np.array_equal(states[b][2:-2, 2:-2], o.state(b)[2:-2, 2:-2]), b
