"""Structured lattices on the unit cube: GPU assembly, interpolation, colouring.

The DOF numbering is femlisp's sort-lexicographically order of DOF midpoints
(src/mesh/meshgen.lisp:214-217); see flgpu_lattice.hip for the layout.
"""
import ctypes

import numpy as np

from .flgpu import ASM_F64, ASM_MFMA, F32, F64, check, lib, ptr
from .matlisp import GpuSparseMatrix, GpuVector, default_context


def ndofs(dim, order, n):
    return (order * n + 1) ** dim


def nnz(dim, order, n):
    return ((3 * n + 1) if order == 1 else (8 * n + 1)) ** dim


def _kappa_vec(kappa, ctx):
    if kappa is None:
        return None, 0
    if isinstance(kappa, tuple):
        return kappa
    k = np.ascontiguousarray(kappa, dtype=np.float64)
    return GpuVector.from_array(k.ravel(), F64, ctx), k.shape[0]


def assemble(dim, order, n, kappa=None, f=1.0, prec=F64, mode=ASM_F64, ctx=None, rhs=True):
    """-div(kappa grad u) = f, zero Dirichlet: returns (A, b) on the device.

    kappa: None, a numpy array of shape (nk,)*dim of per-cell values, or a
    (GpuVector, nk) pair.
    """
    ctx = ctx or default_context()
    kv, nk = _kappa_vec(kappa, ctx)
    hA, hb = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().flgpu_lattice_assemble(ctx.h, dim, order, n, kv.h if kv else None, nk, f, prec,
                                       mode, ctypes.byref(hA), ctypes.byref(hb) if rhs else None))
    A = GpuSparseMatrix(hA, ctx)
    b = GpuVector.wrap(hb, ctx) if rhs else None
    return A, b


def reassemble(A, kappa, mode=ASM_F64):
    kv, nk = _kappa_vec(kappa, A.ctx)
    check(lib().flgpu_lattice_reassemble(A.h, kv.h if kv else None, nk, mode))
    return A


def interpolation(dim, order, nc, prec=F64, ctx=None):
    """Constrained :local interpolation from nc to 2 nc cells per axis."""
    ctx = ctx or default_context()
    h = ctypes.c_void_p()
    check(lib().flgpu_lattice_interpolation(ctx.h, dim, order, nc, prec, ctypes.byref(h)))
    return GpuSparseMatrix(h, ctx)


def blackbox_interpolation(A, P):
    """Overwrite the values of the lattice interpolation P (n/2 -> n cells per
    axis, 3D Q1) with the operator-dependent weights computed from the fine
    matrix A; P's transpose follows."""
    check(lib().flgpu_lattice_blackbox_interpolation(A.h, P.h))
    return P


def colour(A, dim, order, n):
    check(lib().flgpu_lattice_colouring(A.h, dim, order, n))
    return A


def element_matrices(dim, order, n, kappa=None, prec=F64, mode=ASM_F64, first=0, count=None,
                     ctx=None):
    ctx = ctx or default_context()
    kv, nk = _kappa_vec(kappa, ctx)
    count = n ** dim - first if count is None else count
    nloc = (order + 1) ** dim
    out = np.zeros((count, nloc, nloc))
    check(lib().flgpu_lattice_element_matrices(ctx.h, dim, order, n, kv.h if kv else None, nk,
                                               prec, mode, first, count, ptr(out)))
    return out


def lattice_colours(dim, order, n):
    """Colour of every DOF: per-axis coordinate mod (order+1) digits."""
    M = order * n + 1
    m = order + 1
    idx = np.indices((M,) * dim).reshape(dim, -1)
    c = np.zeros(idx.shape[1], np.int32)
    for k in range(dim):
        c = c * m + (idx[k] % m)
    return c


def vertex_blocks(dim, order, n):
    """Keys of the vertex-centred blocks that flgpu_block_iterate sweeps (femlisp's
    find-vertex-centered-block, src/iteration/geoblock.lisp:65-91, on the lattice): one row
    per mesh vertex, in colour order (the parity of the vertex coordinates, axis 0 most
    significant) then index order; one column per slot, offsets -(order-1)..order-1 per axis
    with the last axis fastest; -1 where the point lies outside the lattice or is a Dirichlet
    DOF other than the vertex itself."""
    M, h, w1 = order * n + 1, order - 1, 2 * order - 1
    a = np.indices((n + 1,) * dim).reshape(dim, -1).T
    colour = sum((a[:, k] & 1) << (dim - 1 - k) for k in range(dim))
    a = a[np.argsort(colour, kind="stable")]
    d = np.indices((w1,) * dim).reshape(dim, -1).T - h
    c = a[:, None, :] * order + d[None, :, :]
    inside = ((c >= 0) & (c < M)).all(-1)
    bdry = ((c == 0) | (c == M - 1)).any(-1)
    centre = (d == 0).all(-1)[None, :]
    key = np.zeros(c.shape[:2], np.int64)
    for k in range(dim):
        key = key * M + c[..., k]
    return np.where(inside & (centre | ~bdry), key, -1).astype(np.int32)
