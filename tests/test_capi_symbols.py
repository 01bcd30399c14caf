"""The C-ABI library loads on a CPU host and exports every entry point include/q2d_gpu.h declares."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "q2d_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(q2d_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("q2d_gpu_create", "q2d_gpu_step", "q2d_gpu_run", "q2d_gpu_download_state", "q2d_gpu_destroy"):
        assert n in names


def test_library_exports_all_declared_symbols():
    from q2d_amd import capi
    lib = capi.load_library()
    nm = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (q2d_[a-z_0-9]+)", nm))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for n in declared():
        assert hasattr(lib, n)
    assert set(capi.SIGNATURES) == set(declared())


def test_version_string():
    from q2d_amd import capi
    assert b"gfx950" in capi.load_library().q2d_version()
