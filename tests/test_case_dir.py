"""The case-directory contract (CPU): the Fortran host (q2d_amd/fortran/q2d_case_io.f90, q2d_fcase) and
the Python reader (q2d_amd/casedir.py) read the reference's own ShuOsher-LIMO3 case directory
(tests/golden/cases/ShuOsher-LIMO3: input.ini, INPUT/bc.txt, INPUT/ic.tec, INPUT/phase.txt,
INPUT/thermo.dat, copied from test/1D/ShuOsher-LIMO3 and test/common/Air) into the same problem, bit for
bit; the oracle run of that problem passes the reference's LIMO3 gates and, with the flux limiter of
input.ini replaced, reproduces the committed limiter baselines."""
import shutil
import subprocess
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from q2d_amd import casedir, cases
from q2d_amd.fortran_io import read_case_file

from shu_osher_gates import baseline, limo3_gates

CASE = Path(__file__).resolve().parent / "golden" / "cases" / "ShuOsher-LIMO3"
FCASE = Path(__file__).resolve().parents[1] / "q2d_amd" / "fortran" / "build" / "q2d_fcase"


def test_python_reader_reads_the_reference_case():
    pb = casedir.load_case(CASE)
    s = pb.cases[0]
    assert (s.n_rk, s.limiter, s.riemann, s.time_accurate, s.cfl, s.time_threshold) == (3, "LIMO3", "HLLC", True, 0.5, 1.8)
    assert s.gas_R == cases.AIR_R and s.gas_cp == cases.AIR_CP and s.time0 == 0.0
    assert len(pb.blocks) == 1 and (pb.blocks[0].ni, pb.blocks[0].nj) == (400, 1)
    assert [r["type"] for r in pb.bc].count(300) == 800 and [r["type"] for r in pb.bc].count(400) == 2
    gen = cases.shu_osher(400, limiter="LIMO3")            # the generator of the same case, to rounding
    assert np.array_equal(gen.blocks[0].geo.node, pb.blocks[0].geo.node)
    assert np.abs(gen.blocks[0].P - pb.blocks[0].P).max() <= 2e-15


@pytest.mark.skipif(not FCASE.exists(), reason="Fortran host not built (make -C q2d_amd/fortran)")
def test_fortran_host_builds_the_same_problem(tmp_path):
    r = subprocess.run([str(FCASE), str(CASE), "--dump", str(tmp_path / "case.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    d = read_case_file(tmp_path / "case.bin")
    pb = casedir.load_case(CASE)
    b, fb = pb.blocks[0], d["blocks"][0]
    for k, a in (("P", b.P), ("dl", b.geo.dl), ("M", b.geo.M), ("sur", b.geo.sur), ("fi", b.geo.fi), ("fj", b.geo.fj),
                 ("node", b.geo.node)):
        assert np.array_equal(a, fb[k]), k
    assert [(e.b, e.i, e.j, e.f, e.type) for e in d["bc"]] == [(q["b"], q["i"], q["j"], q["f"], q["type"]) for q in pb.bc]
    sc, s = d["cases"][0].scheme, pb.cases[0]
    assert (sc.n_rk, sc.limiter, sc.riemann, sc.time_accurate, sc.cfl, sc.gas_R, sc.gas_cp) == \
        (3, 6, 0, 1, 0.5, s.gas_R, s.gas_cp)
    assert d["cases"][0].time_threshold == 1.8 and d["cases"][0].time0 == 0.0


@pytest.mark.skipif(not FCASE.exists(), reason="Fortran host not built (make -C q2d_amd/fortran)")
def test_fortran_host_rejects_what_it_does_not_read(tmp_path):
    d = tmp_path / "case"
    shutil.copytree(CASE, d)
    (d / "input.ini").write_text((d / "input.ini").read_text().replace("equations = euler", "equations = navier-stokes"))
    r = subprocess.run([str(FCASE), str(d), "--dump", str(tmp_path / "x.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Euler" in r.stdout
    shutil.rmtree(d)
    shutil.copytree(CASE, d)
    t = (d / "INPUT" / "thermo.dat").read_text().replace("2.0 1004.500000 2009.000000", "2.0 1005.500000 2011.000000")
    (d / "INPUT" / "thermo.dat").write_text(t)
    r = subprocess.run([str(FCASE), str(d), "--dump", str(tmp_path / "x.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cp varies" in r.stdout


def test_oracle_on_the_case_directory_passes_the_reference_gates(oracle_mod):
    pb = casedir.load_case(CASE)
    o = oracle_mod.Oracle(pb)
    rc, out = o.run(100000)
    assert rc == 0 and out[0]["todo"] == 3 and out[0]["time"] >= 1.8
    rho = o.state(0)[2, 2:-2, 0]
    assert limo3_gates(baseline("mc")[:, 0], rho) == []


@pytest.mark.parametrize("lim", ["minmod", "vanleer", "mc", "iord"])
def test_case_directory_reproduces_the_limiter_baselines(oracle_mod, lim):
    pb = casedir.load_case(CASE)
    pb.cases[0] = replace(pb.cases[0], limiter="first-order" if lim == "iord" else lim)
    o = oracle_mod.Oracle(pb)
    rc, out = o.run(100000)
    assert rc == 0
    ref = baseline(lim)
    rho = o.state(0)[2, 2:-2, 0]
    assert (np.abs(rho - ref[:, 1]) / ref[:, 1]).max() < 1e-10


@pytest.mark.skipif(not FCASE.exists(), reason="Fortran host not built (make -C q2d_amd/fortran)")
def test_fortran_host_reads_the_tables(tmp_path):
    """q2d_fcase --tables builds the same planes as q2d_amd.thermo from INPUT/thermo.dat."""
    r = subprocess.run([str(FCASE), str(CASE), "--tables", "--dump", str(tmp_path / "case.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    d = read_case_file(tmp_path / "case.bin")
    pb = casedir.load_case(CASE, thermo_mode="table")
    assert d["cases"][0].thermo_n == pb.cases[0].thermo.shape[1]
    assert np.array_equal(d["tables"][0], pb.cases[0].thermo)
