"""The Fortran case-directory host on the device (needs the GPU): q2d_fcase runs the reference's
ShuOsher-LIMO3 case directory (tests/golden/cases/ShuOsher-LIMO3) through module q2d_gpu and writes
OUTPUT/field.tec and OUTPUT/residual-history.dat.  The written field passes the reference's LIMO3 gates
(check.py) and, with input.ini's flux limiter replaced, reproduces the committed limiter baselines; it
equals the oracle's state to the printed digits, SOLUTIONTIME is the time reached, and a restart from
the written field (newrun = false) continues the run."""
import shutil
import subprocess
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from q2d_amd import casedir

from shu_osher_gates import baseline, limo3_gates

pytestmark = pytest.mark.gpu

CASE = Path(__file__).resolve().parent / "golden" / "cases" / "ShuOsher-LIMO3"
FCASE = Path(__file__).resolve().parents[1] / "q2d_amd" / "fortran" / "build" / "q2d_fcase"
needs_host = pytest.mark.skipif(not FCASE.exists(), reason="Fortran host not built (make -C q2d_amd/fortran)")


def run_case(d, *extra):
    print(f"case dir: {d.name} {extra}", flush=True)
    r = subprocess.run([str(FCASE), str(d), *extra], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    soltime, zones, names = casedir.read_tec(d / "OUTPUT" / "field.tec")
    assert names == ["x", "y", "rho(1)", "u", "v", "p", "T", "g", "R"]
    node, cv = zones[0]
    return soltime, 0.5 * (node[0, 1:, 0] + node[0, :-1, 0]), cv


def copy_case(tmp_path, limiter=None):
    d = tmp_path / f"ShuOsher-{limiter or 'LIMO3'}"
    shutil.copytree(CASE, d)
    if limiter:
        ini = (d / "input.ini").read_text()
        ini = ini.replace("flux-limiter = LIMO3", f"flux-limiter = {limiter}")
        if limiter == "iord":
            ini = ini.replace("space-reconstruction = MUSCL", "space-reconstruction = first-order")
        (d / "input.ini").write_text(ini)
    return d


@needs_host
def test_limo3_case_directory_passes_the_reference_gates(oracle_mod, tmp_path):
    d = copy_case(tmp_path)
    soltime, xc, cv = run_case(d)
    assert limo3_gates(xc, cv[0, 0]) == []
    o = oracle_mod.Oracle(casedir.load_case(CASE))
    rc, out = o.run(100000)
    assert rc == 0 and soltime == float(f"{out[0]['time']:.15E}")
    Po = o.state(0)[2, 2:-2]
    for k in range(4):                                  # 16 significant digits in the file
        assert np.allclose(cv[k, 0], Po[:, k], rtol=2e-15, atol=1e-300), k
    hist = casedir.read_residual_history(d / "OUTPUT" / "residual-history.dat")
    assert hist.shape == (out[0]["iter"], 6) and hist[-1, 0] == out[0]["iter"]
    assert abs(hist[-1, 1] - out[0]["time"]) <= 1e-9 * out[0]["time"]


@needs_host
@pytest.mark.parametrize("lim", ["minmod", "vanleer", "mc", "iord"])
def test_case_directory_reproduces_the_limiter_baselines_on_the_device(tmp_path, lim):
    soltime, xc, cv = run_case(copy_case(tmp_path, lim))
    ref = baseline(lim)
    assert soltime >= 1.8
    assert (np.abs(cv[0, 0] - ref[:, 1]) / ref[:, 1]).max() < 1e-10


@needs_host
def test_restart_from_the_written_field(tmp_path):
    """--max-steps 300, then newrun = false: the run continues from OUTPUT/field.tec at its SOLUTIONTIME to
    the time threshold; the result matches the uninterrupted run to the rounding of the written digits."""
    full = copy_case(tmp_path, "minmod")
    t_full, _, cv_full = run_case(full)
    d = tmp_path / "restart"
    shutil.copytree(CASE, d)
    ini = (d / "input.ini").read_text().replace("flux-limiter = LIMO3", "flux-limiter = minmod")
    (d / "input.ini").write_text(ini)
    t_mid, _, _ = run_case(d, "--max-steps", "300")
    assert 0.0 < t_mid < 1.8
    (d / "input.ini").write_text(ini.replace("[Q2D-Parameters]", "[Q2D-Parameters]\nnewrun = false"))
    t_end, _, cv = run_case(d)
    assert t_end >= 1.8 and abs(t_end - t_full) <= 1e-9
    assert np.abs(cv[0, 0] - cv_full[0, 0]).max() / cv_full[0, 0].max() < 1e-9
    hist = casedir.read_residual_history(d / "OUTPUT" / "residual-history.dat")
    assert hist[0, 0] == 1 and np.all(np.diff(hist[:, 1]) > 0)


@needs_host
def test_table_mode_case_directory(oracle_mod, tmp_path):
    """q2d_fcase --tables: the case's INPUT/thermo.dat as table-mode thermodynamics.  The written field
    equals the oracle's table-mode run to the printed digits and passes the LIMO3 gates."""
    d = copy_case(tmp_path)
    soltime, xc, cv = run_case(d, "--tables")
    assert limo3_gates(xc, cv[0, 0]) == []
    pb = casedir.load_case(CASE, thermo_mode="table")
    o = oracle_mod.Oracle(pb)
    rc, out = o.run(100000)
    assert rc == 0 and soltime == float(f"{out[0]['time']:.15E}")
    Po = o.state(0)[2, 2:-2]
    for k in range(4):
        assert np.allclose(cv[k, 0], Po[:, k], rtol=2e-15, atol=1e-300), k


@needs_host
def test_mixed_table_mode_case_directory(oracle_mod, tmp_path):
    """q2d_fcase --mixed --tables: the mixed form with the case's Air table.  The written field at t = 1.8
    passes the LIMO3 gates; after 300 steps its density is within the mixed form's tolerance (rel. L1 1e-5)
    of the oracle's FP64 table-mode run.  (Over the whole run, 681 steps, the LIMO3 switches at the
    shocklets take the mixed form to 2.1e-5 rel. L1 with the table and 2.8e-5 with the closed form.)"""
    d = copy_case(tmp_path)
    soltime, xc, cv = run_case(d, "--mixed", "--tables")
    assert limo3_gates(xc, cv[0, 0]) == []
    o = oracle_mod.Oracle(casedir.load_case(CASE, thermo_mode="table"))
    rc, out = o.run(100000)
    assert rc == 0 and abs(soltime - out[0]["time"]) <= 1e-6 * out[0]["time"]
    d300 = tmp_path / "steps300"
    shutil.copytree(CASE, d300)
    _, _, cv300 = run_case(d300, "--mixed", "--tables", "--max-steps", "300")
    o300 = oracle_mod.Oracle(casedir.load_case(CASE, thermo_mode="table"))
    rc, _ = o300.run(300)
    rho_o = o300.state(0)[2, 2:-2, 0]
    assert rc == 0 and np.abs(cv300[0, 0] - rho_o).sum() / np.abs(rho_o).sum() <= 1e-5


@needs_host
def test_runtime_reread_of_input_ini(tmp_path):
    """ini-diter = 50: the host re-reads input.ini every 50 steps and passes its runtime keys to the device
    (q2d_gpu_set_runtime, whose next step runs the dt pass).  With the file unchanged the field is the one
    of the run without re-reads, to every written digit."""
    a = copy_case(tmp_path, "minmod")
    _, _, cv_a = run_case(a)
    b = tmp_path / "reread"
    shutil.copytree(a, b)
    ini = (b / "input.ini").read_text()
    (b / "input.ini").write_text(ini.replace("[Q2D-IO]", "[Q2D-IO]\nini-diter = 50") if "[Q2D-IO]" in ini
                                 else ini + "\n[Q2D-IO]\nini-diter = 50\n")
    shutil.rmtree(b / "OUTPUT", ignore_errors=True)
    _, _, cv_b = run_case(b)
    assert np.array_equal(cv_a, cv_b)
