"""Pins of the oracle's FLA_LU_nopiv restatement: the exact factors of
A = L U (dyadic L, integer U with power-of-two diagonal) across the 32- and
128-column block edges and rectangular shapes, the first-zero-on-diagonal
return value, and the residual on diagonally dominant matrices."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.cases import dominant, exact_lu_nopiv


@pytest.mark.parametrize("shape", [(1, 1), (5, 5), (33, 33), (129, 129), (300, 140), (140, 300)])
def test_lu_nopiv_exact(shape):
    m, n = shape
    A, L, U = exact_lu_nopiv(n, seed=m + n, m=m)
    F, r = orc.lu_nopiv(A)
    k = min(m, n)
    assert r == orc.SUCCESS
    assert np.array_equal(np.tril(F[:, :k], -1), np.tril(L, -1))
    assert np.array_equal(np.triu(F[:k, :]), U)


@pytest.mark.parametrize("c", [0, 31, 32, 150])
def test_lu_nopiv_first_zero(c):
    A, _, _ = exact_lu_nopiv(200, seed=c, null_col=c)
    _, r = orc.lu_nopiv(A)
    assert r == c


def test_lu_nopiv_residual():
    A = dominant(300, 4)
    F, r = orc.lu_nopiv(A)
    L = np.tril(F, -1) + np.eye(300)
    U = np.triu(F)
    assert r == orc.SUCCESS
    assert np.max(np.abs(L @ U - A)) <= 1e-13 * np.max(np.abs(A))
