"""flgpu_vec_dot_noalloc, the entry point that allocates nothing (its partials and result live
in a fixed device array): through the C-ABI, its result equals flgpu_vec_dot's bit for bit on
two calls in a row, the second on a shorter vector of the other precision."""
import ctypes
import sys

import numpy as np
import pytest

from gnu_femlisp_amd import flgpu as F
from gnu_femlisp_amd.flgpu import check, lib
from gnu_femlisp_amd.matlisp import GpuVector

pytestmark = pytest.mark.gpu


def _dots(x, y):
    a, b = ctypes.c_double(), ctypes.c_double()
    check(lib().flgpu_vec_dot(x.h, y.h, ctypes.byref(a)))
    check(lib().flgpu_vec_dot_noalloc(x.h, y.h, ctypes.byref(b)))
    return a.value, b.value


def test_dot_noalloc_matches_dot():
    print("dot_noalloc: 2 calls (F64 n=2146689, F32 n=1000)", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(17)
    results = []
    for n, prec in ((2146689, F.F64), (1000, F.F32)):
        xh, yh = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        x, y = GpuVector.from_array(xh, prec), GpuVector.from_array(yh, prec)
        ref, got = _dots(x, y)
        assert got == ref, (n, prec, ref, got)
        tol = 1e-12 if prec == F.F64 else 1e-5
        assert abs(got - float(xh @ yh)) <= tol * float(np.abs(xh) @ np.abs(yh))
        results.append(got)
    assert results[0] != results[1]
    # the same vectors once more, after another call used the array
    x, y = GpuVector.from_array(np.arange(7.0)), GpuVector.from_array(np.ones(7))
    ref, got = _dots(x, y)
    assert got == ref == 21.0
