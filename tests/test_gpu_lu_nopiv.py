"""FLA_LU_nopiv / FLA_LU_nopiv_solve on the GPU against the oracle.  Bar:
the closed-form factors of A = L U to 1e-13 relative (the GPU's A21 solve
multiplies by the 128-block inverse of U11, so it rounds where the oracle's
substitution is exact), the oracle's factors to 1e-12 on diagonally
dominant matrices, the same first-zero-on-diagonal index, and a solve with a
small backward error."""
import numpy as np
import pytest

import libflame_amd as fla
from tests.cases import dominant, exact_lu_nopiv
from tests.gpu_util import to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


def gpu_lu_nopiv(A):
    d = to_dev(A)
    r = fla.FLA_LU_nopiv(d.obj())
    return d.download(), r


@pytest.mark.parametrize("shape", [(1, 1), (37, 37), (128, 128), (129, 129), (300, 300),
                                   (700, 300), (300, 700), (1000, 1000)])
def test_lu_nopiv_closed_form(orc, shape):
    m, n = shape
    A, L, U = exact_lu_nopiv(n, seed=m * 3 + n, m=m)
    F, r = gpu_lu_nopiv(A)
    k = min(m, n)
    assert r == fla.FLA_SUCCESS
    assert np.max(np.abs(np.tril(F[:, :k], -1) - np.tril(L, -1))) <= 1e-13 * max(1.0, np.max(np.abs(L)))
    assert np.max(np.abs(np.triu(F[:k, :]) - U)) <= 1e-13 * np.max(np.abs(U))
    Fo, ro = orc.lu_nopiv(A)
    assert ro == r


@pytest.mark.parametrize("n", [64, 333, 1024, 2048])
def test_lu_nopiv_vs_oracle(orc, n):
    A = dominant(n, n)
    F, r = gpu_lu_nopiv(A)
    Fo, ro = orc.lu_nopiv(A)
    assert r == ro == fla.FLA_SUCCESS
    assert np.max(np.abs(F - Fo)) <= 1e-12 * np.max(np.abs(Fo))


@pytest.mark.parametrize("c", [0, 3, 127, 128, 250, 299])
def test_lu_nopiv_first_zero(orc, c):
    A, _, _ = exact_lu_nopiv(300, seed=5 + c, null_col=c)
    _, r = gpu_lu_nopiv(A)
    _, ro = orc.lu_nopiv(A)
    assert r == ro == c


def test_lu_nopiv_solve():
    n, nrhs = 500, 7
    A = dominant(n, 9)
    rng = np.random.default_rng(2)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dA, dB = to_dev(A), to_dev(B)
    dX = fla.DeviceMatrix(n, nrhs)
    assert fla.FLA_LU_nopiv(dA.obj()) == fla.FLA_SUCCESS
    assert fla.FLA_LU_nopiv_solve(dA.obj(), dB.obj(), dX.obj()) == fla.FLA_SUCCESS
    X = dX.download()
    res = np.max(np.abs(A @ X - B)) / (np.max(np.abs(A)) * np.max(np.abs(X)) * n)
    assert res <= 1e-15
