"""FLA_Trinv / FLA_Ttmm / FLA_SPDinv and dtrtri_ / dlauum_ / dpotri_ on the
GPU against the oracle.  Bar: the closed forms exactly (bidiagonal
inverses, integer L'L), the untouched other triangle and unit diagonal,
agreement with the oracle and numpy to 1e-12 relative on random
well-conditioned inputs, the Cholesky failure index, and LAPACK's info
conventions."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from tests.gpu_util import ip, to_dev
from tests.inv_cases import LOWER, NONUNIT, UNIT, UPPER, bidiag_case, int_triangle, random_triangle

pytestmark = pytest.mark.gpu

SIZES = [1, 37, 128, 129, 300, 513]


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


def gpu_call(fn, A, *args):
    d = to_dev(A)
    fla.check(fn(*args, A.shape[0], d.ptr, d.ld))
    fla.synchronize()
    return d.download()


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("diag", [NONUNIT, UNIT])
@pytest.mark.parametrize("n", SIZES)
def test_trinv_closed_form(orc, uplo, diag, n):
    A, want = bidiag_case(n, uplo, diag, seed=n)
    got = gpu_call(fla.lib().fla_hip_dtrtri, A, uplo, diag)
    assert np.array_equal(got, want)
    assert np.array_equal(orc.trinv(A, uplo, diag), want)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", [64, 300, 1000])
def test_trinv_vs_oracle(orc, uplo, n):
    T = random_triangle(n, uplo, seed=n)
    got = gpu_call(fla.lib().fla_hip_dtrtri, T, uplo, NONUNIT)
    ref = orc.trinv(T, uplo, NONUNIT)
    tri = np.tril if uplo == LOWER else np.triu
    assert np.max(np.abs(tri(got) - tri(ref))) <= 1e-12 * np.max(np.abs(ref))
    # the other triangle is the input's
    other = np.triu if uplo == LOWER else np.tril
    assert np.array_equal(other(got, 1 if uplo == LOWER else -1), other(T, 1 if uplo == LOWER else -1))


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", SIZES)
def test_ttmm_exact(orc, uplo, n):
    A, want = int_triangle(n, uplo, seed=n)
    got = gpu_call(fla.lib().fla_hip_dlauum, A, uplo)
    assert np.array_equal(got, want)
    assert np.array_equal(orc.ttmm(A, uplo), want)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("n", [1, 100, 300, 1000])
def test_spdinv_vs_oracle_and_numpy(orc, uplo, n):
    A = orc.gen_spd_c2(21, n)
    d = to_dev(A)
    assert fla.FLA_SPDinv(uplo, d.obj()) == fla.FLA_SUCCESS
    got = d.download()
    ref, r = orc.spdinv(A, uplo)
    assert r == orc.SUCCESS
    tri = np.tril if uplo == LOWER else np.triu
    inv = np.linalg.inv(A)
    scale = np.max(np.abs(inv))
    assert np.max(np.abs(tri(got) - tri(ref))) <= 1e-12 * scale
    assert np.max(np.abs(tri(got) - tri(inv))) <= 1e-12 * scale


def test_spdinv_cabi_failure_index(orc):
    A = orc.gen_spd_c2(5, 300)
    A[200, 200] = -1.0
    d = to_dev(A)
    info = np.zeros(1, np.int32)
    fla.check(fla.lib().fla_hip_dspdinv(LOWER, 300, d.ptr, d.ld, ip(info)))
    fla.synchronize()
    _, rc = orc.chol(A)
    assert info[0] == rc == 200


def test_flame_trinv_ttmm_objects(orc):
    n = 257
    A, want = bidiag_case(n, LOWER, NONUNIT, seed=1)
    d = to_dev(A)
    assert fla.FLA_Trinv(LOWER, NONUNIT, d.obj()) == fla.FLA_SUCCESS
    assert np.array_equal(d.download(), want)
    B, wantB = int_triangle(n, UPPER, seed=2)
    d = to_dev(B)
    assert fla.FLA_Ttmm(UPPER, d.obj()) == fla.FLA_SUCCESS
    assert np.array_equal(d.download(), wantB)


def _c(s):
    return C.c_char_p(s.encode())


def test_lapack_dtrtri_dlauum_dpotri(orc):
    L = fla.lib()
    n = 200
    # dtrtri_: closed form, and info = first zero diagonal (1-based), A untouched
    A, want = bidiag_case(n, UPPER, NONUNIT, seed=3)
    a = np.asfortranarray(A.copy())
    nn, lda, info = np.array([n], np.int32), np.array([n], np.int32), np.zeros(1, np.int32)
    L.dtrtri_(_c("U"), _c("N"), ip(nn), a.ctypes.data_as(C.POINTER(C.c_double)), ip(lda), ip(info))
    assert info[0] == 0 and np.array_equal(a, want)
    z = np.asfortranarray(A.copy())
    z[57, 57] = 0.0
    z0 = z.copy()
    L.dtrtri_(_c("U"), _c("N"), ip(nn), z.ctypes.data_as(C.POINTER(C.c_double)), ip(lda), ip(info))
    assert info[0] == 58 and np.array_equal(z, z0)
    # dpotri_ on dpotrf_'s factor = inv(A); dlauum_ = the integer closed form
    S = orc.gen_spd_c2(9, n)
    s = np.asfortranarray(S.copy())
    L.dpotrf_(_c("L"), ip(nn), s.ctypes.data_as(C.POINTER(C.c_double)), ip(lda), ip(info))
    assert info[0] == 0
    L.dpotri_(_c("L"), ip(nn), s.ctypes.data_as(C.POINTER(C.c_double)), ip(lda), ip(info))
    assert info[0] == 0
    inv = np.linalg.inv(S)
    assert np.max(np.abs(np.tril(s) - np.tril(inv))) <= 1e-12 * np.max(np.abs(inv))
    B, wantB = int_triangle(n, LOWER, seed=4)
    b = np.asfortranarray(B.copy())
    L.dlauum_(_c("L"), ip(nn), b.ctypes.data_as(C.POINTER(C.c_double)), ip(lda), ip(info))
    assert info[0] == 0 and np.array_equal(b, wantB)
