"""Pins of the oracle's FLA_Trinv / FLA_Ttmm / FLA_SPDinv restatements:
exact closed forms (bidiagonal inverses, integer L'L), numpy's inverse of
random well-conditioned triangles and SPD matrices, the untouched other
triangle, and the Cholesky failure index for a non-SPD input."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.inv_cases import LOWER, NONUNIT, UNIT, UPPER, bidiag_case, int_triangle, random_triangle


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("diag", [NONUNIT, UNIT])
@pytest.mark.parametrize("n", [1, 2, 7, 130])
def test_trinv_closed_form(uplo, diag, n):
    A, want = bidiag_case(n, uplo, diag, seed=n)
    assert np.array_equal(orc.trinv(A, uplo, diag), want)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", [5, 64, 200])
def test_trinv_matches_numpy(uplo, n):
    T = random_triangle(n, uplo, seed=3)
    X = orc.trinv(T, uplo, NONUNIT)
    tri = np.tril if uplo == LOWER else np.triu
    ref = np.linalg.inv(T)
    assert np.max(np.abs(tri(X) - ref)) <= 1e-13 * np.max(np.abs(ref)) * n


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", [1, 3, 50, 131])
def test_ttmm_exact(uplo, n):
    A, want = int_triangle(n, uplo, seed=n)
    assert np.array_equal(orc.ttmm(A, uplo), want)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", [1, 40, 257])
def test_spdinv_matches_numpy(uplo, n):
    A = orc.gen_spd_c2(11, n)
    X, r = orc.spdinv(A, uplo)
    assert r == orc.SUCCESS
    tri = np.tril if uplo == LOWER else np.triu
    ref = np.linalg.inv(A)
    assert np.max(np.abs(tri(X) - tri(ref))) <= 1e-12 * np.max(np.abs(ref))


def test_spdinv_reports_cholesky_failure():
    A = orc.gen_spd_c2(5, 60)
    A[33, 33] = -1.0
    _, r = orc.spdinv(A)
    _, rc = orc.chol(A)
    assert r == rc == 33
