"""System files: a sparse system handed over offline (the layout flgpu_system_read
reads, include/flgpu.h).

femlisp's side exports a <sparse-matrix> with sparse-matrix->ccs
(src/matlisp/sparselu.lisp:235-333), i.e. compressed columns in the order of its
keys, together with the rank of every key's DOF midpoint in
sort-lexicographically order (src/mesh/meshgen.lisp:214-217) as the permutation.
``write`` and ``read`` are the host-side (numpy) writer and reader of the same
layout; ``load`` and ``save`` go through the C-ABI to and from the device.
"""
import ctypes

import numpy as np

from . import flgpu as _f
from .flgpu import F64, check, lib

MAGIC = b"FLSYS01\x00"
ROWS, COLUMNS = 0, 1


def write(path, starts, indices, values, nrows, ncols, orientation=ROWS, rhs=None, perm=None):
    """Write a system: starts/indices/values compressed by rows (orientation
    ROWS) or columns (COLUMNS); rhs (nrows values) and perm (perm[k] = canonical
    index of the file's row and column k) are optional."""
    starts = np.ascontiguousarray(starts, np.int32)
    indices = np.ascontiguousarray(indices, np.int32)
    values = np.ascontiguousarray(values, np.float64)
    nmajor = ncols if orientation == COLUMNS else nrows
    if starts.size != nmajor + 1 or indices.size != values.size or starts[-1] != values.size:
        raise ValueError("inconsistent compressed arrays")
    flags = (1 if rhs is not None else 0) | (2 if perm is not None else 0)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([orientation, flags], "<i4").tobytes())
        f.write(np.array([nrows, ncols, values.size], "<i8").tobytes())
        f.write(starts.astype("<i4").tobytes())
        f.write(indices.astype("<i4").tobytes())
        f.write(values.astype("<f8").tobytes())
        if rhs is not None:
            f.write(np.ascontiguousarray(rhs, "<f8").tobytes())
        if perm is not None:
            f.write(np.ascontiguousarray(perm, "<i4").tobytes())


def read(path):
    """The raw content of a system file as a dict (no permutation applied)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != MAGIC:
        raise ValueError("not a system file (bad magic)")
    orient, flags = np.frombuffer(data, "<i4", 2, 8)
    nrows, ncols, nnz = (int(v) for v in np.frombuffer(data, "<i8", 3, 16))
    off = 40
    nmajor = ncols if orient == COLUMNS else nrows

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(data, dtype, count, off).copy()
        off += a.nbytes
        return a
    out = {"orientation": int(orient), "nrows": nrows, "ncols": ncols,
           "starts": take("<i4", nmajor + 1), "indices": take("<i4", nnz),
           "values": take("<f8", nnz)}
    out["rhs"] = take("<f8", nrows) if flags & 1 else None
    out["perm"] = take("<i4", nrows) if flags & 2 else None
    if off != len(data):
        raise ValueError("trailing bytes in system file")
    return out


def load(path, prec=F64, ctx=None):
    """(GpuSparseMatrix, GpuVector or None) from a system file, in canonical numbering."""
    from .matlisp import GpuSparseMatrix, GpuVector, default_context
    ctx = ctx or default_context()
    A, b = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().flgpu_system_read(ctx.h, str(path).encode(), prec, ctypes.byref(A),
                                  ctypes.byref(b)))
    M = GpuSparseMatrix(A, ctx)
    return M, (GpuVector.wrap(b, ctx) if b.value else None)


def save(path, A, rhs=None):
    check(lib().flgpu_system_write(str(path).encode(), A.h, rhs.h if rhs is not None else None))
