"""The complex z filter's operator (oracle side, CPU).

filter3_CMPLX (src/filters/cf90.F90:1230-1251) forms the right-hand side of
the caller's boundary codes (only when bcn_ is present) and always solves
with penta_nn, the (0, 0) left-hand side.  Its oracle composes the pinned
pieces: A of the (0, 0) filter and B of the caller's codes (both pinned by
test_oracle_pinning.py), so the checks here are that composition and the
cases where it reduces to the plain filter.
"""
import numpy as np
import pytest

from oracle import pade_dense as pd


@pytest.mark.parametrize("n", [10, 16, 33])
def test_no_codes_is_the_bounded_filter(n):
    D = pd.filter3_complex_operator(n, False)
    assert np.allclose(D, pd.dense_operator("cf90", "filter", n, 1.0, False, 0, 0), rtol=0, atol=1e-14)
    D00 = pd.filter3_complex_operator(n, False, 0, 0)
    assert np.array_equal(D, D00)
    # bc1 without bcn is ignored, as present(bcn_) decides (cf90.F90:1242)
    assert np.array_equal(pd.filter3_complex_operator(n, False, 1, None), D)


@pytest.mark.parametrize("bc", [(1, 1), (-1, 1), (1, 0), (0, -1)])
def test_mixed_operator_composition(bc):
    n = 24
    M = pd.filter3_complex_operator(n, False, *bc)
    A_nn, _ = pd.assemble("cf90", "filter", n, 1.0, False, 0, 0)
    A_bc, B_bc = pd.assemble("cf90", "filter", n, 1.0, False, *bc)
    # A_nn M = B_bc: the caller's right-hand side with the (0, 0) left-hand side
    assert np.allclose(A_nn @ M, B_bc, rtol=0, atol=1e-12)
    # and it differs from the consistent filter of those codes (the quirk is kept)
    D_bc = np.linalg.solve(A_bc, B_bc)
    assert np.abs(M - D_bc).max() > 1e-6


def test_periodic_is_the_periodic_filter():
    n = 32
    assert np.array_equal(pd.filter3_complex_operator(n, True, 1, 1),
                          pd.dense_operator("cf90", "filter", n, 1.0, True))
