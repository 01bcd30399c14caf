"""Case files for the Fortran host adapter (q2d_amd/fortran/q2d_fdriver.f90).

The file carries a Problem in the reference's allocation layout, so the Fortran host reads each array
straight into an allocatable of the reference's shape (P(4,ni+4,nj+4), dl(2,...), ...) and passes
c_loc of it through module q2d_gpu.  The case and boundary-record descriptors are the C structs'
bytes, which the module's bind(C) types must read back field for field.
"""
from __future__ import annotations

import ctypes as C
import struct
from pathlib import Path

import numpy as np

from .capi import StepOut, to_problem_c

MAGIC = 1362248753
DRIVER = Path(__file__).resolve().parent / "fortran" / "build" / "q2d_fdriver"


def write_case_file(problem, path, nsteps):
    pc, keep = to_problem_c(problem)
    blocks, bcs, cases, _ = keep
    nb, nbc, nc = len(problem.blocks), len(problem.bc), len(problem.cases)
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", MAGIC, nb, nbc, nc, int(nsteps)))
        f.write(C.string_at(C.addressof(cases), C.sizeof(cases)))
        if nbc:
            f.write(C.string_at(C.addressof(bcs), C.sizeof(bcs[0]) * nbc))
        for b in problem.blocks:
            g = b.geo
            flags = (1 if g.node is not None else 0) | (2 if b.src_a is not None else 0) | (4 if b.src_b is not None else 0)
            f.write(struct.pack("<4i", b.ni, b.nj, b.case_id, flags))
            for a in (b.P, g.dl, g.M, g.sur, g.fi, g.fj):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            for a in (g.node, b.src_a, b.src_b):
                if a is not None:
                    f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def read_output(path, problem):
    raw = Path(path).read_bytes()
    off = 0
    states = []
    for b in problem.blocks:
        n = (b.nj + 4) * (b.ni + 4) * 4
        states.append(np.frombuffer(raw, dtype=np.float64, count=n, offset=off).reshape(b.nj + 4, b.ni + 4, 4).copy())
        off += 8 * n
    outs = []
    for _ in problem.cases:
        so = StepOut.from_buffer_copy(raw[off:off + C.sizeof(StepOut)])
        outs.append(so.as_dict())
        off += C.sizeof(StepOut)
    return states, outs
