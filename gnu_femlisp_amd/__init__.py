"""MI355X (gfx950) implementation of femlisp's solver hot path for Lagrange FE
discretisations: sparse matrix-vector products, Jacobi / multicolour
Gauss-Seidel smoothers, the geometric-multigrid V-cycle, CG / flexible CG and
Q1/Q2 cell assembly, behind a C-ABI (include/flgpu.h, libflgpu.so).

The Python modules mirror femlisp's generic-function API:
  matlisp    gemm!, m*, dot, norm, axpy!, copy!, scal!, compute-residual
  iteration  <jacobi>, <sor>, <geometric-cs>, <cg>, <linear-solver>, solve
  lattice    GPU assembly, interpolation and colouring on box meshes
"""
from .flgpu import (ASM_F64, ASM_MFMA, F32, F64, FlgpuError, LIB_PATH, lib)  # noqa: F401
