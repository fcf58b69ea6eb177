"""Device vectors and sparse matrices behind femlisp's matlisp generic functions.

Mirrors the names and argument order of femlisp's FL.MATLISP interface
(src/matlisp/matrix.lisp:213-240,337-394, src/matlisp/vector.lisp:91-152,
224-326): gemm!(alpha, A, y, beta, x, job), m*, dot, norm, axpy!, copy!, scal!,
fill!, x<-0, compute-residual (src/iteration/linit.lisp:73-80).  Every call
goes through the C-ABI of libflgpu.so.
"""
import ctypes

import numpy as np

from . import flgpu as _f
from .flgpu import F32, F64, check, lib, ptr

_DEFAULT_CTX = None


class Context:
    """One device, one HIP stream (flgpu_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().flgpu_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def info(self):
        name = ctypes.create_string_buffer(128)
        cus = ctypes.c_int()
        mem = ctypes.c_long()
        check(lib().flgpu_ctx_info(self.h, name, 128, ctypes.byref(cus), ctypes.byref(mem)))
        return {"arch": name.value.decode(), "cus": cus.value, "mem": mem.value}

    def timer_start(self):
        check(lib().flgpu_ctx_timer_start(self.h))

    def timer_stop(self):
        ms = ctypes.c_double()
        check(lib().flgpu_ctx_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        check(lib().flgpu_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().flgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context():
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class GpuVector:
    """Device vector (flgpu_vec); host values are always float64."""

    def __init__(self, n, prec=F64, ctx=None, handle=None, _owner=True):
        self.ctx = ctx or default_context()
        if handle is None:
            h = ctypes.c_void_p()
            check(lib().flgpu_vec_create(self.ctx.h, int(n), prec, ctypes.byref(h)))
            handle = h
        self.h = handle
        self.n = int(n)
        self.prec = prec
        self._owner = _owner

    @classmethod
    def from_array(cls, a, prec=F64, ctx=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        v = cls(a.size, prec, ctx)
        v.upload(a)
        return v

    @classmethod
    def wrap(cls, handle, ctx, owner=True):
        n = ctypes.c_long()
        p = ctypes.c_int()
        check(lib().flgpu_vec_info(handle, ctypes.byref(n), ctypes.byref(p)))
        return cls(n.value, p.value, ctx, handle, owner)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        check(lib().flgpu_vec_upload(self.h, ptr(a)))

    def numpy(self):
        out = np.zeros(self.n)
        check(lib().flgpu_vec_download(self.h, ptr(out)))
        return out

    def __len__(self):
        return self.n

    def __del__(self):
        try:
            if self._owner and getattr(self, "h", None):
                lib().flgpu_vec_destroy(self.h)
                self.h = None
        except Exception:
            pass


class GpuSparseMatrix:
    """A <sparse-matrix> whose storage lives on the GPU (flgpu_mat)."""

    def __init__(self, handle, ctx, owner=True):
        self.h = handle
        self.ctx = ctx
        self._owner = owner
        r, c, z, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_long(), ctypes.c_int()
        check(lib().flgpu_mat_info(handle, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z),
                                   ctypes.byref(p)))
        self.nrows, self.ncols, self.nnz, self.prec = r.value, c.value, z.value, p.value

    @classmethod
    def from_csr(cls, rowptr, colidx, vals, ncols=None, prec=F64, ctx=None):
        ctx = ctx or default_context()
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        nrows = rowptr.size - 1
        ncols = nrows if ncols is None else ncols
        h = ctypes.c_void_p()
        check(lib().flgpu_csr_upload(ctx.h, nrows, ncols, colidx.size, ptr(rowptr), ptr(colidx),
                                     ptr(vals), prec, ctypes.byref(h)))
        return cls(h, ctx)

    @classmethod
    def from_scipy(cls, m, prec=F64, ctx=None):
        m = m.tocsr()
        m.sort_indices()
        return cls.from_csr(m.indptr, m.indices, m.data, m.shape[1], prec, ctx)

    def csr(self):
        rp = np.zeros(self.nrows + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        v = np.zeros(self.nnz)
        check(lib().flgpu_csr_download(self.h, ptr(rp), ptr(ci), ptr(v)))
        return rp, ci, v

    def update_values(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        assert vals.size == self.nnz
        check(lib().flgpu_csr_update_values(self.h, ptr(vals)))

    def convert(self, prec):
        h = ctypes.c_void_p()
        check(lib().flgpu_mat_convert(self.h, prec, ctypes.byref(h)))
        return GpuSparseMatrix(h, self.ctx)

    def set_colouring(self, colour):
        colour = np.ascontiguousarray(colour, dtype=np.int32)
        check(lib().flgpu_mat_set_colouring(self.h, int(colour.max()) + 1, ptr(colour)))

    def storage(self, max_form=_f.STORAGE_CELL):
        """Storage the kernels use (STORAGE_CELL / STENCIL / SELL), capped at
        max_form (richer forms are dropped for good)."""
        form = ctypes.c_int()
        check(lib().flgpu_mat_storage(self.h, int(max_form), ctypes.byref(form)))
        return form.value

    def stencil_planes(self):
        """value planes the stencil-form kernels stream per row (27, 14 when symmetric, 0 when
        the matrix is not in the stencil form)"""
        p = ctypes.c_int()
        check(lib().flgpu_mat_stencil_planes(self.h, ctypes.byref(p)))
        return p.value

    def colour_greedy(self):
        n = ctypes.c_int()
        check(lib().flgpu_mat_colour_greedy(self.h, ctypes.byref(n)))
        return n.value

    def make_domain_vector(self, prec=None):
        """make-domain-vector-for (src/matlisp/matrix.lisp:337)"""
        return GpuVector(self.ncols, self.prec if prec is None else prec, self.ctx)

    def make_image_vector(self, prec=None):
        """make-image-vector-for (src/matlisp/matrix.lisp:338)"""
        return GpuVector(self.nrows, self.prec if prec is None else prec, self.ctx)

    def __del__(self):
        try:
            if self._owner and getattr(self, "h", None):
                lib().flgpu_mat_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---- generic functions ------------------------------------------------------

def gemm(alpha, A, y, beta, x, job="nn"):
    """gemm! alpha A y beta x -> x (jobs :nn and :tn)."""
    if job == "nn":
        check(lib().flgpu_gemm_nn(alpha, A.h, y.h, beta, x.h))
    elif job == "tn":
        check(lib().flgpu_gemm_tn(alpha, A.h, y.h, beta, x.h))
    else:
        raise ValueError("job must be 'nn' or 'tn'")
    return x


def m_star(A, y):
    """m* A y: a fresh image vector holding A y."""
    x = A.make_image_vector()
    return gemm(1.0, A, y, 0.0, x)


def compute_residual(A, x, b, r):
    check(lib().flgpu_compute_residual(A.h, x.h, b.h, r.h))
    return r


def dot(x, y):
    out = ctypes.c_double()
    check(lib().flgpu_vec_dot(x.h, y.h, ctypes.byref(out)))
    return out.value


def norm(x):
    out = ctypes.c_double()
    check(lib().flgpu_vec_norm2(x.h, ctypes.byref(out)))
    return out.value


l2_norm = norm


def axpy(alpha, x, y):
    """axpy! alpha x y -> y"""
    check(lib().flgpu_vec_axpy(alpha, x.h, y.h))
    return y


def copy(x, y):
    """copy! x y -> y"""
    check(lib().flgpu_vec_copy(x.h, y.h))
    return y


def scal(alpha, x):
    check(lib().flgpu_vec_scal(alpha, x.h))
    return x


def fill(x, value):
    check(lib().flgpu_vec_fill(x.h, value))
    return x


def x_lt_0(x):
    """x<-0"""
    return fill(x, 0.0)
