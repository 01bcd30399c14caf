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

MAGIC = 1362248754
# the header carries the sizes of the descriptor structs written as raw bytes, so that a host built
# against another layout of include/q2d_gpu.h reports the mismatch instead of misreading the file
HEADER = "<8i"
DRIVER = Path(__file__).resolve().parent / "fortran" / "build" / "q2d_fdriver"


def write_case_file(problem, path, nsteps):
    pc, keep = to_problem_c(problem)
    blocks, bcs, cases, _ = keep
    nb, nbc, nc = len(problem.blocks), len(problem.bc), len(problem.cases)
    ptrs = [cases[k].thermo for k in range(nc)]
    for k in range(nc):                     # pointers of this process mean nothing in the file
        cases[k].thermo = None
    with open(path, "wb") as f:
        f.write(struct.pack(HEADER, MAGIC, nb, nbc, nc, int(nsteps), C.sizeof(cases[0]) if nc else 0,
                            C.sizeof(bcs[0]), C.sizeof(StepOut)))
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
        for c in problem.cases:             # table mode: the property planes of each such case
            if c.thermo is not None:
                f.write(np.ascontiguousarray(c.thermo, dtype=np.float64).tobytes())
    for k in range(nc):
        cases[k].thermo = ptrs[k]


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


def read_case_file(path):
    """A case file (as write_case_file writes it, or q2d_fcase --dump): header, case and record structs,
    and per block its arrays in the reference layout."""
    from .capi import BCEntry, CaseDesc
    raw = Path(path).read_bytes()
    hdr = struct.unpack_from(HEADER, raw, 0)
    if hdr[0] != MAGIC:
        raise ValueError("not a case file of this layout version")
    _, nb, nbc, nc, nsteps, s_case, s_bc, s_out = hdr
    if (nc and s_case != C.sizeof(CaseDesc)) or s_bc != C.sizeof(BCEntry) or s_out != C.sizeof(StepOut):
        raise ValueError("descriptor layout mismatch")
    off = struct.calcsize(HEADER)
    cs = [CaseDesc.from_buffer_copy(raw[off + k * s_case: off + (k + 1) * s_case]) for k in range(nc)]
    off += nc * s_case
    bcs = [BCEntry.from_buffer_copy(raw[off + k * s_bc: off + (k + 1) * s_bc]) for k in range(nbc)]
    off += nbc * s_bc
    blocks = []
    for _ in range(nb):
        ni, nj, cid, flags = struct.unpack_from("<4i", raw, off)
        off += 16

        def take(shape):
            nonlocal off
            n = int(np.prod(shape))
            a = np.frombuffer(raw, dtype=np.float64, count=n, offset=off).reshape(shape).copy()
            off += 8 * n
            return a
        blk = {"ni": ni, "nj": nj, "case_id": cid, "P": take((nj + 4, ni + 4, 4)), "dl": take((nj + 4, ni + 4, 2)),
               "M": take((nj + 4, ni + 4, 4)), "sur": take((nj + 4, ni + 4)), "fi": take((nj, ni + 1, 3)),
               "fj": take((nj + 1, ni, 3))}
        if flags & 1:
            blk["node"] = take((nj + 1, ni + 1, 2))
        if flags & 2:
            blk["src_a"] = take((nj, ni))
        if flags & 4:
            blk["src_b"] = take((nj, ni))
        blocks.append(blk)
    tables = []
    for c in cs:
        n = int(c.thermo_n)
        if n > 0:
            tables.append(np.frombuffer(raw, dtype=np.float64, count=4 * n, offset=off).reshape(4, n).copy())
            off += 32 * n
        else:
            tables.append(None)
    return {"nsteps": nsteps, "cases": cs, "bc": bcs, "blocks": blocks, "tables": tables}
