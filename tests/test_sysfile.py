"""System files (flgpu_system_read / flgpu_system_write, gnu_femlisp_amd/sysfile.py).

CPU: the numpy writer and reader agree, and malformed files are refused.
GPU: a femlisp-style export (compressed columns in an arbitrary key order plus the
sort-lexicographically permutation) of an oracle system loads as the oracle's
canonical CSR bit for bit, and the library's own writer round-trips it.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from gnu_femlisp_amd import sysfile
from oracle import pyoracle as po


def _femlisp_export(dim, order, n, seed):
    """The oracle system in a random key order, as sparse-matrix->ccs writes it."""
    rng = np.random.default_rng(seed)
    kap = np.exp(rng.normal(size=(n,) * dim))
    rp, ci, v, b = po.assemble(dim, order, n, kap)
    N = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(N, N))
    perm = rng.permutation(N).astype(np.int32)  # file key k is canonical DOF perm[k]
    F = A[perm][:, perm].tocsc()                # F[k, l] = A[perm[k], perm[l]]
    F.sort_indices()
    return (rp, ci, v, b), F, b[perm], perm


def test_numpy_round_trip(tmp_path):
    _, F, bf, perm = _femlisp_export(2, 2, 3, 1)
    path = tmp_path / "s.flsys"
    sysfile.write(path, F.indptr, F.indices, F.data, F.shape[0], F.shape[1],
                  orientation=sysfile.COLUMNS, rhs=bf, perm=perm)
    d = sysfile.read(path)
    assert d["orientation"] == sysfile.COLUMNS and d["nrows"] == F.shape[0]
    assert np.array_equal(d["starts"], F.indptr) and np.array_equal(d["indices"], F.indices)
    assert np.array_equal(d["values"], F.data) and np.array_equal(d["rhs"], bf)
    assert np.array_equal(d["perm"], perm)


def test_malformed_files_refused(tmp_path):
    bad = tmp_path / "bad.flsys"
    bad.write_bytes(b"NOTASYS\x00" + bytes(32))
    with pytest.raises(ValueError):
        sysfile.read(bad)
    _, F, bf, perm = _femlisp_export(1, 1, 4, 2)
    good = tmp_path / "good.flsys"
    sysfile.write(good, F.indptr, F.indices, F.data, F.shape[0], F.shape[1],
                  orientation=sysfile.COLUMNS)
    short = tmp_path / "short.flsys"
    short.write_bytes(good.read_bytes()[:-8])
    with pytest.raises(ValueError):
        sysfile.read(short)
    with pytest.raises(ValueError):
        sysfile.write(good, F.indptr[:-1], F.indices, F.data, F.shape[0], F.shape[1],
                      orientation=sysfile.COLUMNS)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,order,n", [(2, 1, 6), (3, 1, 5), (3, 2, 2)])
def test_femlisp_export_loads_canonically(tmp_path, dim, order, n):
    from gnu_femlisp_amd import flgpu as F_
    (rp, ci, v, b), F, bf, perm = _femlisp_export(dim, order, n, 7)
    path = tmp_path / "export.flsys"
    sysfile.write(path, F.indptr, F.indices, F.data, F.shape[0], F.shape[1],
                  orientation=sysfile.COLUMNS, rhs=bf, perm=perm)
    A, rhs = sysfile.load(path)
    grp, gci, gv = A.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci) and np.array_equal(gv, v)
    assert np.array_equal(rhs.numpy(), b)
    # the library's writer: rows compressed, canonical order, no permutation
    out = tmp_path / "again.flsys"
    sysfile.save(out, A, rhs)
    d = sysfile.read(out)
    assert d["orientation"] == sysfile.ROWS and d["perm"] is None
    assert np.array_equal(d["starts"], rp) and np.array_equal(d["indices"], ci)
    assert np.array_equal(d["values"], v) and np.array_equal(d["rhs"], b)
    # a permutation that is not one is refused
    bad = tmp_path / "bad.flsys"
    p2 = perm.copy()
    p2[0] = p2[1]
    sysfile.write(bad, F.indptr, F.indices, F.data, F.shape[0], F.shape[1],
                  orientation=sysfile.COLUMNS, perm=p2)
    with pytest.raises(F_.FlgpuError):
        sysfile.load(bad)
