"""The reference configurations at their full sizes (GPU, through the C-ABI).

- 2D Poisson, Q1, 32x32 cells, V(2,2) damped Jacobi (omega 0.6) PCG to 1e-10:
  the oracle's iteration count, defect history and solution.
- 3D Poisson, Q1, 64^3 cells, rhs U[-1,1] per interior node from a seed,
  V(2,2) multicolour GS PCG to 1e-10: the oracle's iteration count and
  solution.
- 3D variable coefficient, Q2, 48^3 cells, kappa = exp(g), g ~ N(0, 1.5^2)
  per cell, f = 1: FP32 V-cycle with BF16/FP32-accumulate MFMA assembly inside
  FP64 flexible CG with the true residual to 1e-10.  The oracle takes minutes at
  this size, so the result is checked through size-independent properties: the
  FP64 true residual, and agreement with the GPU's plain FP64 GMG-CG (whose
  parity with the oracle the smaller cases pin) to 1e-9.
- Deploy: 3D Q1, 128^3 cells, two-phase inclusions (contrast 1e4, 30 % volume fraction,
  radii 3-8 cells, SplitMix64 seed 2024): femlisp's plain FP64 GMG-CG follows the oracle's
  defect history over its first steps (the oracle needs about 0.5 s per step here), and the
  plain and the mixed black-box Jacobi paths both reach an FP64 true residual below 1e-10 and
  agree to 1e-9.
"""
import sys

import numpy as np
import pytest

from gnu_femlisp_amd import flgpu as F
from gnu_femlisp_amd import iteration as it
from gnu_femlisp_amd import lattice as lat
from gnu_femlisp_amd.fields import inclusions
from gnu_femlisp_amd.matlisp import GpuVector
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(x, y):
    d = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (d if d > 0 else 1.0)


def test_config_2d_poisson_jacobi_cg():
    say("config: 2D Q1 32^2 Poisson, V(2,2) Jacobi 0.6, PCG")
    oh = po.Hierarchy(2, 1, 1, 6)
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_JACOBI, omega=0.6, pre_steps=2,
                    post_steps=2, base_level=0, alias=0, tol=1e-10, maxsteps=100)
    st, xo, so, dno = oh.solve(prm)
    assert st == 0 and oh.ndofs() == 1089
    cs = it.GeometricCS(it.Jacobi(0.6), 2, 2, 0)
    H = it.Hierarchy.lattice(cs, 2, 1, 1, 6)
    bb = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 100)),
                  {"matrix": H.top_matrix(), "rhs": H.rhs, "hierarchy": H})
    assert bb["status"] == "success" and bb["steps"] == so
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)
    assert relerr(bb["solution"].numpy(), xo) < 1e-10


def test_config_3d_poisson_64_gs_cg():
    say("config: 3D Q1 64^3 Poisson, random rhs, V(2,2) multicolour GS, PCG")
    n, levels = 64, 7
    oh = po.Hierarchy(3, 1, 1, levels)
    N = oh.ndofs()
    assert N == 274625
    b = np.random.default_rng(2024).uniform(-1, 1, N)
    b[po.boundary(3, 1, n)] = 0
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                    ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=0,
                    symmetric=1, alias=0, tol=1e-10, maxsteps=100)
    st, xo, so, dno = oh.solve(prm, b)
    assert st == 0
    cs = it.GeometricCS(it.SOR(1.0), 2, 2, 0, symmetric=True)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, levels)
    bb = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                  failure_if=(">", "step", 100)),
                  {"matrix": H.top_matrix(), "rhs": GpuVector.from_array(b), "hierarchy": H})
    assert bb["status"] == "success" and bb["steps"] == so
    assert np.allclose(bb["defnorms"], dno, rtol=1e-8)
    assert relerr(bb["solution"].numpy(), xo) < 1e-10


def test_config_q2_48_mixed_fcg():
    say("config: 3D Q2 48^3, kappa = exp(N(0, 1.5^2)), mixed FCG")
    n0, levels = 3, 5
    n = n0 << (levels - 1)
    kap = np.exp(np.random.default_rng(48).normal(scale=1.5, size=(n, n, n)))
    A, b = lat.assemble(3, 2, n, kap, prec=F.F64)
    assert A.nrows == 912673 and A.nnz == 57066625
    cs64 = it.GeometricCS(it.SOR(1.0), 2, 2, 0, symmetric=True)
    H64 = it.Hierarchy.lattice(cs64, 3, 2, n0, levels, kappa=kap)
    plain = it.solve(it.LinearSolver(it.CG(cs64), success_if=("<", "reduction", 1e-12),
                                     failure_if=(">", "step", 300), true_residual=True),
                     {"matrix": A, "rhs": b, "hierarchy": H64})
    assert plain["status"] == "success"
    del H64
    cs32 = it.GeometricCS(it.SOR(1.0), 2, 2, 0, symmetric=True, prec=F.F32)
    H32 = it.Hierarchy.lattice(cs32, 3, 2, n0, levels, kappa=kap, mode=F.ASM_MFMA)
    mixed = it.solve(it.LinearSolver(it.FCG(cs32), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 300), true_residual=True),
                     {"matrix": A, "rhs": b, "hierarchy": H32})
    say(f"  plain {plain['steps']} steps, mixed {mixed['steps']} steps")
    assert mixed["status"] == "success"
    assert mixed["true_residual"] / np.linalg.norm(b.numpy()) < 1e-10
    assert relerr(mixed["solution"].numpy(), plain["solution"].numpy()) < 1e-9


def test_config_q2_48_block_ssc():
    """The same Q2 48^3 system with femlisp's default smoother for order 2, the vertex-centred
    block SSC (geomg.lisp:58-71), V(1,1): FP64 GMG-CG and the FP32 cycle inside FP64 FCG reach
    1e-10 and agree."""
    say("config: 3D Q2 48^3, block SSC")
    n0, levels = 3, 5
    n = n0 << (levels - 1)
    kap = np.exp(np.random.default_rng(48).normal(scale=1.5, size=(n, n, n)))
    A, b = lat.assemble(3, 2, n, kap, prec=F.F64)
    sols, steps = [], []
    for prec, itc in ((F.F64, it.CG), (F.F32, it.FCG)):
        cs = it.GeometricCS(it.GeometricSSC(1.0), 1, 1, 0, symmetric=True, prec=prec)
        H = it.Hierarchy.lattice(cs, 3, 2, n0, levels, kappa=kap,
                                 mode=F.ASM_MFMA if prec == F.F32 else F.ASM_F64)
        r = it.solve(it.LinearSolver(itc(cs), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 100), true_residual=True),
                     {"matrix": A, "rhs": b, "hierarchy": H})
        assert r["status"] == "success"
        assert r["true_residual"] / np.linalg.norm(b.numpy()) < 1e-10
        sols.append(r["solution"].numpy())
        steps.append(r["steps"])
        del H
    say(f"  ssc: FP64 CG {steps[0]} steps, FP32-in-FCG {steps[1]} steps")
    assert steps[0] <= 40 and steps[1] <= 40
    assert relerr(sols[1], sols[0]) < 1e-9


def test_config_deploy_128_inclusions():
    say("config: 3D Q1 128^3 two-phase inclusions, contrast 1e4")
    n, levels, base = 128, 8, 3
    kap = inclusions(n)
    assert abs(np.mean(kap > 1) - 0.30) < 0.01
    # (a) femlisp's plain FP64 GMG-CG, V(2,2) multicolour GS, over its first 5 steps
    oh = po.Hierarchy(3, 1, 1, levels, kappa=kap)
    assert oh.ndofs() == 2146689
    b = oh.rhs()
    prm = po.params(outer=po.OUTER_CG, smoother=po.SMOOTH_SOR, omega=1.0,
                    ordering=po.ORDER_COLOUR, pre_steps=2, post_steps=2, base_level=base,
                    symmetric=1, alias=0, tol=1e-10, maxsteps=4)
    st, xo, so, dno = oh.solve(prm, b)
    del oh
    assert st == 1 and so == 5
    cs = it.GeometricCS(it.SOR(1.0), 2, 2, base, symmetric=True)
    H = it.Hierarchy.lattice(cs, 3, 1, 1, levels, kappa=kap)
    A, bg = H.top_matrix(), H.rhs
    assert np.allclose(bg.numpy(), b, rtol=1e-13, atol=0)
    first = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 4)),
                     {"matrix": A, "rhs": bg, "hierarchy": H})
    say(f"  first steps: gpu {first['defnorms']}, oracle {dno}")
    assert first["steps"] == so
    assert np.allclose(first["defnorms"], dno, rtol=1e-8)
    assert relerr(first["solution"].numpy(), xo) < 1e-8
    # (b) plain FP64 GMG-CG and the value path (FP32 black-box Jacobi V(2,2) cycle in FP64
    # FCG) to an FP64 true residual below 1e-10
    plain = it.solve(it.LinearSolver(it.CG(cs), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 600), true_residual=True),
                     {"matrix": A, "rhs": bg, "hierarchy": H})
    bn = np.linalg.norm(b)
    assert plain["status"] == "success" and plain["true_residual"] / bn < 1e-10
    cs32 = it.GeometricCS(it.Jacobi(0.9), 2, 2, base, prec=F.F32, transfer="blackbox")
    H32 = it.Hierarchy.lattice(cs32, 3, 1, 1, levels, kappa=kap, mode=F.ASM_MFMA)
    mixed = it.solve(it.LinearSolver(it.FCG(cs32), success_if=("<", "reduction", 1e-10),
                                     failure_if=(">", "step", 600), true_residual=True),
                     {"matrix": A, "rhs": bg, "hierarchy": H32})
    say(f"  plain {plain['steps']} steps, mixed-bb-jac {mixed['steps']} steps")
    assert mixed["status"] == "success" and mixed["true_residual"] / bn < 1e-10
    assert mixed["steps"] < plain["steps"]
    assert relerr(mixed["solution"].numpy(), plain["solution"].numpy()) < 1e-9
