"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
restatement -- never by the product package gnu_femlisp_amd.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OUTER_MG, OUTER_CG, OUTER_CG_PLAIN, OUTER_FCG = 0, 1, 2, 3
SMOOTH_JACOBI, SMOOTH_SOR, SMOOTH_SSC, SMOOTH_PSC = 0, 1, 2, 3
ORDER_INDEX, ORDER_COLOUR = 0, 1
STOP_REDUCTION, STOP_DEFNORM = 0, 1
FALLBACK_NONE, FALLBACK_LU = 0, 1


class Params(ctypes.Structure):
    _fields_ = [
        ("outer", ctypes.c_int),
        ("smoother", ctypes.c_int),
        ("omega", ctypes.c_double),
        ("ordering", ctypes.c_int),
        ("pre_steps", ctypes.c_int),
        ("post_steps", ctypes.c_int),
        ("base_level", ctypes.c_int),
        ("symmetric", ctypes.c_int),
        ("alias", ctypes.c_int),
        ("stop", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("maxsteps", ctypes.c_int),
        ("gamma", ctypes.c_int),
        ("additive", ctypes.c_int),
        ("fas", ctypes.c_int),
        ("restart_cycle", ctypes.c_int),
        ("fallback", ctypes.c_int),
    ]


def params(outer=OUTER_MG, smoother=SMOOTH_JACOBI, omega=1.0, ordering=ORDER_INDEX,
           pre_steps=1, post_steps=1, base_level=0, symmetric=0, alias=0,
           stop=STOP_REDUCTION, tol=1e-10, maxsteps=100, gamma=1, additive=0, fas=0,
           restart_cycle=0, fallback=FALLBACK_NONE):
    return Params(outer, smoother, omega, ordering, pre_steps, post_steps,
                  base_level, symmetric, alias, stop, tol, maxsteps, gamma, additive, fas,
                  restart_cycle, fallback)


def set_threads(n):
    """host threads of the oracle's parallel loops (n <= 0: all cores)"""
    lib().flo_set_threads(int(n))


def get_threads():
    return lib().flo_get_threads()


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        ci, cl, cd, vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
        L.flo_set_threads.argtypes = [ci]
        L.flo_get_threads.restype = ci
        L.flo_lattice_ndofs.restype = cl
        L.flo_lattice_ndofs.argtypes = [ci, ci, ci]
        L.flo_lattice_nnz.restype = cl
        L.flo_lattice_nnz.argtypes = [ci, ci, ci]
        L.flo_lattice_pattern.argtypes = [ci, ci, ci, vp, vp]
        L.flo_lattice_boundary.argtypes = [ci, ci, ci, vp]
        L.flo_gauss_rule.argtypes = [ci, vp, vp]
        L.flo_lagrange_1d.argtypes = [ci, cd, vp, vp]
        L.flo_element_matrix.argtypes = [ci, ci, cd, vp, cd, vp, vp]
        L.flo_assemble.argtypes = [ci, ci, ci, vp, ci, cd, vp, vp]
        L.flo_interpolation.restype = cl
        L.flo_interpolation.argtypes = [ci, ci, ci, vp, vp, vp]
        L.flo_projection.restype = cl
        L.flo_projection.argtypes = [ci, ci, ci, vp, vp, vp]
        L.flo_mg_get_projection.argtypes = [vp, ci, vp, vp, vp]
        L.flo_blackbox_interpolation.restype = ci
        L.flo_blackbox_interpolation.argtypes = [ci, vp, vp, vp, vp, vp, vp]
        L.flo_vertex_blocks.restype = cl
        L.flo_vertex_blocks.argtypes = [ci, ci, ci, vp]
        L.flo_block_iterate.argtypes = [ci, vp, vp, vp, ci, ci, ci, ci, cd, ci, vp, vp, vp]
        L.flo_spmv.argtypes = [ci, vp, vp, vp, cd, vp, cd, vp]
        L.flo_spmv_t.argtypes = [ci, ci, vp, vp, vp, cd, vp, cd, vp]
        L.flo_csr_transpose.argtypes = [ci, ci, vp, vp, vp, vp, vp, vp]
        L.flo_mg_create.restype = vp
        L.flo_mg_create.argtypes = [ci, ci, ci, ci, vp, ci, cd]
        L.flo_mg_create_ex.restype = vp
        L.flo_mg_create_ex.argtypes = [ci, ci, ci, ci, vp, ci, cd, ci]
        L.flo_mg_destroy.argtypes = [vp]
        L.flo_mg_ndofs.restype = cl
        L.flo_mg_ndofs.argtypes = [vp, ci]
        L.flo_mg_nnz.restype = cl
        L.flo_mg_nnz.argtypes = [vp, ci]
        L.flo_mg_get_matrix.argtypes = [vp, ci, vp, vp, vp]
        L.flo_mg_interp_nnz.restype = cl
        L.flo_mg_interp_nnz.argtypes = [vp, ci]
        L.flo_mg_get_interp.argtypes = [vp, ci, vp, vp, vp]
        L.flo_mg_get_rhs.argtypes = [vp, vp]
        L.flo_mg_solve.argtypes = [vp, P(Params), vp, vp, P(ci), vp]
        L.flo_mg_apply.argtypes = [vp, P(Params), vp, vp]
        L.flo_unet_nparams.restype = cl
        L.flo_unet_nparams.argtypes = [ci]
        L.flo_unet_forward.argtypes = [ci, ci, ci, vp, vp, vp, ci]
        L.flo_unet_loss_grad.restype = cd
        L.flo_unet_loss_grad.argtypes = [ci, ci, ci, vp, vp, vp, vp, ci]
        L.flo_adam.argtypes = [cl, vp, vp, vp, vp, ci, cd, cd, cd, cd]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data if a is not None else None


def ndofs(dim, order, n):
    return lib().flo_lattice_ndofs(dim, order, n)


def pattern(dim, order, n):
    N = ndofs(dim, order, n)
    nnz = lib().flo_lattice_nnz(dim, order, n)
    rp = np.zeros(N + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    lib().flo_lattice_pattern(dim, order, n, _p(rp), _p(ci))
    return rp, ci


def boundary(dim, order, n):
    b = np.zeros(ndofs(dim, order, n), np.uint8)
    lib().flo_lattice_boundary(dim, order, n, _p(b))
    return b.astype(bool)


def gauss_rule(s):
    p = np.zeros(s)
    w = np.zeros(s)
    assert lib().flo_gauss_rule(s, _p(p), _p(w)) == 0
    return p, w


def lagrange_1d(order, t):
    phi = np.zeros(order + 1)
    dphi = np.zeros(order + 1)
    lib().flo_lagrange_1d(order, t, _p(phi), _p(dphi))
    return phi, dphi


def element_matrix(dim, order, H, kappa_qp=None, f=1.0):
    nloc = (order + 1) ** dim
    K = np.zeros((nloc, nloc))
    b = np.zeros(nloc)
    kq = None if kappa_qp is None else np.ascontiguousarray(kappa_qp, np.float64)
    lib().flo_element_matrix(dim, order, H, _p(kq), f, _p(K), _p(b))
    return K, b


def assemble(dim, order, n, kappa=None, f=1.0):
    """Returns (rowptr, colidx, vals, rhs) of the Dirichlet-eliminated system."""
    rp, ci = pattern(dim, order, n)
    vals = np.zeros(len(ci))
    rhs = np.zeros(len(rp) - 1)
    nk = 0
    k = None
    if kappa is not None:
        k = np.ascontiguousarray(kappa, np.float64)
        nk = int(round(k.size ** (1.0 / dim)))
        assert nk ** dim == k.size
    rc = lib().flo_assemble(dim, order, n, _p(k), nk, f, _p(vals), _p(rhs))
    assert rc == 0, rc
    return rp, ci, vals, rhs


def interpolation(dim, order, nc):
    Nf = ndofs(dim, order, 2 * nc)
    rp = np.zeros(Nf + 1, np.int32)
    nnz = lib().flo_interpolation(dim, order, nc, _p(rp), None, None)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    lib().flo_interpolation(dim, order, nc, _p(rp), _p(ci), _p(v))
    return rp, ci, v


def projection(dim, order, nc):
    """FAS projection from 2 nc to nc cells per axis (flo_projection)."""
    Nc = ndofs(dim, order, nc)
    rp, ci, v = np.zeros(Nc + 1, np.int32), np.zeros(Nc, np.int32), np.zeros(Nc)
    lib().flo_projection(dim, order, nc, _p(rp), _p(ci), _p(v))
    return rp, ci, v


def blackbox_interpolation(nf, rp, ci, vals):
    """Operator-dependent interpolation (3D Q1) from the fine matrix with nf
    cells per axis, in the pattern of interpolation(3, 1, nf // 2)."""
    prp, pci, _ = interpolation(3, 1, nf // 2)
    pv = np.zeros(pci.size)
    rp, ci = np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    if lib().flo_blackbox_interpolation(nf, _p(rp), _p(ci), _p(vals), _p(prp), _p(pci),
                                        _p(pv)) != 0:
        raise ValueError("nf must be even")
    return prp, pci, pv


def vertex_blocks(dim, order, n):
    """Vertex-centred blocks (flo_vertex_blocks): int32 array (blocks, (2 order - 1)^dim),
    -1 for absent slots, blocks in colour order."""
    nb = lib().flo_vertex_blocks(dim, order, n, None)
    keys = np.zeros((nb, (2 * order - 1) ** dim), np.int32)
    lib().flo_vertex_blocks(dim, order, n, _p(keys))
    return keys


def block_iterate(rp, ci, vals, dim, order, n, ssc, omega, x, b, r, backwards=False):
    """One <geometric-ssc> (ssc=True) or <geometric-psc> iterate; returns new (x, r)."""
    x = np.array(x, np.float64)
    r = np.array(r, np.float64)
    b = r if b is None else np.ascontiguousarray(b, np.float64)
    rp, ci = np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    rc = lib().flo_block_iterate(len(rp) - 1, _p(rp), _p(ci), _p(vals), dim, order, n,
                                 int(ssc), omega, int(backwards), _p(x), _p(b), _p(r))
    assert rc == 0, rc
    return x, r


def spmv(rp, ci, vals, y, alpha=1.0, beta=0.0, x=None):
    n = len(rp) - 1
    x = np.zeros(n) if x is None else np.array(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    lib().flo_spmv(n, _p(rp), _p(ci), _p(vals), alpha, _p(y), beta, _p(x))
    return x


class Hierarchy:
    """Multigrid hierarchy of the oracle: re-discretised levels, or Galerkin
    levels (galerkin=True) with femlisp's interpolation (transfer="geometric")
    or the operator-dependent one (transfer="blackbox", 3D Q1)."""

    def __init__(self, dim, order, n0, levels, kappa=None, f=1.0, galerkin=False,
                 transfer="geometric"):
        self.dim, self.order, self.n0, self.levels = dim, order, n0, levels
        nk = 0
        self._kappa = None
        if kappa is not None:
            self._kappa = np.ascontiguousarray(kappa, np.float64)
            nk = int(round(self._kappa.size ** (1.0 / dim)))
        mode = 2 if transfer == "blackbox" else int(galerkin)
        self.h = lib().flo_mg_create_ex(dim, order, n0, levels, _p(self._kappa), nk, f, mode)
        if not self.h:
            raise ValueError("flo_mg_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().flo_mg_destroy(self.h)
            self.h = None

    def ndofs(self, level=-1):
        return lib().flo_mg_ndofs(self.h, level % self.levels)

    def matrix(self, level=-1):
        level %= self.levels
        N = self.ndofs(level)
        nnz = lib().flo_mg_nnz(self.h, level)
        rp = np.zeros(N + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        lib().flo_mg_get_matrix(self.h, level, _p(rp), _p(ci), _p(v))
        return rp, ci, v

    def interp(self, level):
        nnz = lib().flo_mg_interp_nnz(self.h, level)
        Nf = self.ndofs(level + 1)
        rp = np.zeros(Nf + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        lib().flo_mg_get_interp(self.h, level, _p(rp), _p(ci), _p(v))
        return rp, ci, v

    def projection(self, level):
        """the FAS projection level + 1 -> level"""
        Nc = self.ndofs(level)
        rp, ci, v = np.zeros(Nc + 1, np.int32), np.zeros(Nc, np.int32), np.zeros(Nc)
        assert lib().flo_mg_get_projection(self.h, level, _p(rp), _p(ci), _p(v)) == 0
        return rp, ci, v

    def rhs(self):
        b = np.zeros(self.ndofs())
        lib().flo_mg_get_rhs(self.h, _p(b))
        return b

    def solve(self, prm, b=None):
        b = self.rhs() if b is None else np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.ndofs())
        steps = ctypes.c_int(0)
        dn = np.zeros(prm.maxsteps + 3)
        st = lib().flo_mg_solve(self.h, ctypes.byref(prm), _p(b), _p(x),
                                ctypes.byref(steps), _p(dn))
        return st, x, steps.value, dn[: steps.value + 1].copy()

    def apply(self, prm, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.ndofs())
        assert lib().flo_mg_apply(self.h, ctypes.byref(prm), _p(b), _p(x)) == 0
        return x


# ---- learned-correction network checker (conv_oracle.c) ----------------------

def unet_nparams(cin=2):
    return lib().flo_unet_nparams(cin)


def unet_forward(theta, x, batch, n, cin=2, bf16=True):
    theta = np.ascontiguousarray(theta, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(batch * n ** 3)
    assert lib().flo_unet_forward(batch, n, cin, _p(theta), _p(x), _p(y), int(bf16)) == 0
    return y


def unet_loss_grad(theta, x, t, batch, n, cin=2, bf16=True):
    theta = np.ascontiguousarray(theta, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    g = np.zeros(unet_nparams(cin))
    loss = lib().flo_unet_loss_grad(batch, n, cin, _p(theta), _p(x), _p(t), _p(g), int(bf16))
    return loss, g


def adam(theta, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    theta, g, m, v = (np.ascontiguousarray(a, np.float64).copy() for a in (theta, g, m, v))
    lib().flo_adam(theta.size, _p(theta), _p(g), _p(m), _p(v), t, lr, b1, b2, eps)
    return theta, m, v
