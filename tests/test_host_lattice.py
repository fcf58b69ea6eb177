"""Host-side lattice helpers of the package against the oracle (no GPU needed)."""
import numpy as np
import pytest

from gnu_femlisp_amd import iteration as it
from gnu_femlisp_amd import lattice as lat
from oracle import pyoracle as po


@pytest.mark.parametrize("dim,order,n", [(1, 2, 5), (2, 1, 4), (2, 2, 3), (3, 1, 3), (3, 2, 2),
                                         (3, 2, 5)])
def test_vertex_blocks_match_oracle(dim, order, n):
    assert np.array_equal(lat.vertex_blocks(dim, order, n), po.vertex_blocks(dim, order, n))


def test_select_smoother_follows_order():
    """select-smoother (geomg.lisp:58-71): <gauss-seidel> for order 1, <geometric-ssc> above;
    a GeometricCS without a smoother passes the matching kind to the library."""
    from gnu_femlisp_amd import flgpu as F
    assert isinstance(it.select_smoother(1), it.GaussSeidel)
    assert isinstance(it.select_smoother(2), it.GeometricSSC)
    cs = it.GeometricCS()
    assert cs.params(1).smoother == F.SMOOTH_SOR and cs.params(2).smoother == F.SMOOTH_SSC
    assert it.GeometricCS(it.GeometricPSC(0.2)).params(2).smoother == F.SMOOTH_PSC
