"""The Fortran host checks the case file's descriptor layout before it reads anything else (CPU).

q2d_fdriver reads the C structs q2d_case_desc / q2d_bc_entry / q2d_step_out as raw bytes into its
bind(C) types; the file header carries the writer's struct sizes, and a mismatch stops the host with a
message instead of a misread (a struct field added on one side only)."""
import struct
import subprocess

import pytest

from q2d_amd import cases
from q2d_amd.fortran_io import DRIVER, HEADER, write_case_file

pytestmark = pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built (make -C q2d_amd/fortran)")


def run(path, tmp_path):
    return subprocess.run([str(DRIVER), str(path), str(tmp_path / "out.bin")], capture_output=True, text=True,
                          timeout=60)


def test_matching_layout_is_accepted(tmp_path):
    write_case_file(cases.riemann_planar(seed=1, ni=32, nj=4, steps=2), tmp_path / "case.bin", 2)
    r = run(tmp_path / "case.bin", tmp_path)
    assert "layout" not in r.stdout + r.stderr
    assert r.returncode == 0 or "no device" in r.stdout        # on a host without a GPU it stops at the device


@pytest.mark.parametrize("field", [5, 6, 7])
def test_mismatched_struct_size_is_reported(tmp_path, field):
    p = tmp_path / "case.bin"
    write_case_file(cases.riemann_planar(seed=1, ni=32, nj=4, steps=2), p, 2)
    raw = bytearray(p.read_bytes())
    hdr = list(struct.unpack_from(HEADER, raw, 0))
    hdr[field] += 8
    struct.pack_into(HEADER, raw, 0, *hdr)
    p.write_bytes(bytes(raw))
    r = run(p, tmp_path)
    assert r.returncode != 0 and "descriptor layout mismatch" in r.stdout


def test_old_magic_is_reported(tmp_path):
    p = tmp_path / "case.bin"
    write_case_file(cases.riemann_planar(seed=1, ni=32, nj=4, steps=2), p, 2)
    raw = bytearray(p.read_bytes())
    struct.pack_into("<i", raw, 0, 1362248753)
    p.write_bytes(bytes(raw))
    r = run(p, tmp_path)
    assert r.returncode != 0 and "magic mismatch" in r.stdout
