"""Level-3 kernels, solves and lapack2flame entry points on the GPU.

Exact integer/dyadic problems must come back bit for bit (every operation
order gives the exact answer); random problems are held to the dget02-style
backward error ||B - A X|| / (||A|| ||X|| n eps) <= 10."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from tests.cases import exact_lu, exact_spd, exact_tri
from tests.gpu_util import dp, gpu_gemm, gpu_syrk, gpu_trsm, ip, to_dev

pytestmark = pytest.mark.gpu

L, U = fla.FLA_LOWER_TRIANGULAR, fla.FLA_UPPER_TRIANGULAR
N, T = fla.FLA_NO_TRANSPOSE, fla.FLA_TRANSPOSE


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


@pytest.mark.parametrize("ta", [N, T])
@pytest.mark.parametrize("tb", [N, T])
@pytest.mark.parametrize("shape", [(1, 1, 1), (37, 29, 17), (128, 128, 16), (257, 130, 300),
                                   (513, 64, 128)])
def test_dgemm_exact(orc, ta, tb, shape):
    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    A = rng.integers(-5, 6, size=(m, k) if ta == N else (k, m)).astype(np.float64)
    B = rng.integers(-5, 6, size=(k, n) if tb == N else (n, k)).astype(np.float64)
    C0 = rng.integers(-9, 10, size=(m, n)).astype(np.float64)
    got = gpu_gemm(ta, tb, A, B, C0, alpha=-1.0, beta=1.0)
    want = orc.dgemm("N" if ta == N else "T", "N" if tb == N else "T", A, B, C0)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("trans", [N, T])
def test_dsyrk_exact(orc, uplo, trans):
    rng = np.random.default_rng(5)
    n, k = 300, 77
    A = rng.integers(-5, 6, size=(n, k) if trans == N else (k, n)).astype(np.float64)
    C0 = rng.integers(-9, 10, size=(n, n)).astype(np.float64)
    got = gpu_syrk(uplo, trans, A, C0, alpha=-0.5, beta=2.0)
    want = orc.dsyrk("L" if uplo == L else "U", "N" if trans == N else "T", A, C0, alpha=-0.5,
                     beta=2.0)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("side", [fla.FLA_LEFT, fla.FLA_RIGHT])
@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("trans", [N, T])
@pytest.mark.parametrize("diag", [fla.FLA_NONUNIT_DIAG, fla.FLA_UNIT_DIAG])
@pytest.mark.parametrize("dim,rhs", [(300, 1), (300, 21), (300, 40), (2000, 16)])
def test_dtrsm_all_variants_vs_oracle(orc, side, uplo, trans, diag, dim, rhs):
    """Well-conditioned triangles (the GPU solve multiplies by inverted
    128x128 diagonal blocks, so exactness of dyadic data is not preserved):
    agreement with the oracle's dtrsm_ to 1e-12 and a backward error <= 10.
    Left solves with <= 32 right-hand sides take the persistent block-row
    kernel (16 columns per launch), wider ones the MFMA block substitution."""
    m, n = (dim, rhs) if side == fla.FLA_LEFT else (rhs, dim)
    rng = np.random.default_rng(23)
    Tm = rng.uniform(-1, 1, (dim, dim)) / np.sqrt(dim)
    Tm = np.tril(Tm) if uplo == L else np.triu(Tm)
    np.fill_diagonal(Tm, rng.uniform(1.0, 2.0, dim) * rng.choice([-1.0, 1.0], dim))
    Tm = np.asfortranarray(Tm)
    B = np.asfortranarray(rng.standard_normal((m, n)))
    got = gpu_trsm(side, uplo, trans, diag, Tm, B, alpha=2.0)
    want = orc.dtrsm("L" if side == fla.FLA_LEFT else "R", "L" if uplo == L else "U",
                     "N" if trans == N else "T", "U" if diag == fla.FLA_UNIT_DIAG else "N", Tm,
                     B, alpha=2.0)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    Te = Tm.copy()
    if diag == fla.FLA_UNIT_DIAG:
        np.fill_diagonal(Te, 1.0)
    opT = Te if trans == N else Te.T
    R = (opT @ got if side == fla.FLA_LEFT else got @ opT) - 2.0 * B
    res = np.linalg.norm(R, 1) / (np.linalg.norm(Te, 1) * np.linalg.norm(got, 1) * dim * 2**-53)
    assert res <= 10.0


@pytest.mark.parametrize("side", [fla.FLA_LEFT, fla.FLA_RIGHT])
@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("trans", [N, T])
def test_dtrsm_exact_single_block(side, uplo, trans):
    """Within one 128-block with a power-of-two diagonal and a bidiagonal
    off-diagonal of +-1/2, the inverse is exact, so the solve is exact."""
    m, n = (100, 7) if side == fla.FLA_LEFT else (7, 100)
    dim = m if side == fla.FLA_LEFT else n
    rng = np.random.default_rng(3)
    Tm = np.diag(rng.choice([1.0, 2.0, -4.0], dim))
    off = rng.choice([0.5, -0.5, 0.0], dim - 1)
    if uplo == L:
        Tm += np.diag(off, -1)
    else:
        Tm += np.diag(off, 1)
    Tm = np.asfortranarray(Tm)
    X = rng.integers(-3, 4, size=(m, n)).astype(np.float64)
    opT = Tm if trans == N else Tm.T
    B = opT @ X if side == fla.FLA_LEFT else X @ opT
    got = gpu_trsm(side, uplo, trans, fla.FLA_NONUNIT_DIAG, Tm, B)
    assert np.allclose(got, X, rtol=0, atol=1e-12)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n", [64, 1000])
def test_chol_solve_backward_error(orc, upper, n):
    uplo = U if upper else L
    A = orc.gen_spd_c2(11, n)
    F, r = orc.chol(A, uplo=orc.UPPER if upper else orc.LOWER)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((n, 16))
    dF, dB, dX = to_dev(F), to_dev(B), fla.DeviceMatrix(n, 16)
    assert fla.FLA_Chol_solve(uplo, dF.obj(), dB.obj(), dX.obj()) == fla.FLA_SUCCESS
    X = dX.download()
    ref = orc.chol_solve(F, B, uplo=orc.UPPER if upper else orc.LOWER)
    res = np.linalg.norm(B - A @ X, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1) * n * 2**-53)
    assert res <= 10.0
    assert np.max(np.abs(X - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_lu_solve_backward_error(orc):
    n = 1000
    A = orc.gen_lu_c2(5, n, n)
    F, p, _ = orc.lu_piv(A)
    B = np.random.default_rng(1).standard_normal((n, 16))
    dF, dB, dX = to_dev(F), to_dev(B), fla.DeviceMatrix(n, 16)
    dpv = fla.DeviceMatrix.from_host(p.reshape(-1, 1), np.int32)
    assert fla.FLA_LU_piv_solve(dF.obj(), dpv.obj(), dB.obj(), dX.obj()) == fla.FLA_SUCCESS
    X = dX.download()
    res = np.linalg.norm(B - A @ X, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1) * n * 2**-53)
    assert res <= 10.0
    ref = orc.lu_piv_solve(F, p, B)
    assert np.max(np.abs(X - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_lapack2flame_dpotrf_dpotrs(orc):
    n, nrhs = 500, 16
    A = np.asfortranarray(orc.gen_spd_c1(3, n))
    Af = A.copy(order="F")
    info = C.c_int(7)
    fla.lib().dpotrf_(b"L", C.byref(C.c_int(n)), dp(Af), C.byref(C.c_int(n)), C.byref(info))
    assert info.value == 0
    assert orc.chol_residual(A, Af) <= 10.0
    B = np.asfortranarray(np.random.default_rng(0).standard_normal((n, nrhs)))
    X = B.copy(order="F")
    fla.lib().dpotrs_(b"L", C.byref(C.c_int(n)), C.byref(C.c_int(nrhs)), dp(Af),
                      C.byref(C.c_int(n)), dp(X), C.byref(C.c_int(n)), C.byref(info))
    assert info.value == 0
    res = np.linalg.norm(B - A @ X, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1) * n * 2**-53)
    assert res <= 10.0
    # failure: info = index + 1
    S, _ = exact_spd(200, seed=1)
    S[77, 77] = -5.0
    S = np.asfortranarray(S)
    fla.lib().dpotrf_(b"L", C.byref(C.c_int(200)), dp(S), C.byref(C.c_int(200)), C.byref(info))
    assert info.value == 78


@pytest.mark.parametrize("trans", [b"N", b"T"])
def test_lapack2flame_dgetrf_dgetrs(orc, trans):
    n, nrhs = 400, 16
    A = np.asfortranarray(orc.gen_lu_c2(9, n, n))
    Af = A.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(7)
    fla.lib().dgetrf_(C.byref(C.c_int(n)), C.byref(C.c_int(n)), dp(Af), C.byref(C.c_int(n)),
                      ip(ipiv), C.byref(info))
    assert info.value == 0
    Fo = A.copy(order="F")
    iref = np.zeros(n, dtype=np.int32)
    orc.lib().or_dgetrf(n, n, orc.dp(Fo), n, orc.ip(iref), C.byref(C.c_int()))
    assert np.array_equal(ipiv, iref)
    B = np.asfortranarray(np.random.default_rng(0).standard_normal((n, nrhs)))
    X = B.copy(order="F")
    fla.lib().dgetrs_(trans, C.byref(C.c_int(n)), C.byref(C.c_int(nrhs)), dp(Af),
                      C.byref(C.c_int(n)), ip(ipiv), dp(X), C.byref(C.c_int(n)), C.byref(info))
    assert info.value == 0
    opA = A if trans == b"N" else A.T
    res = np.linalg.norm(B - opA @ X, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1) * n * 2**-53)
    assert res <= 10.0
    # null pivot: info = index + 1
    Z, _, _, _ = exact_lu(300, seed=2, null_col=123)
    Z = np.asfortranarray(Z)
    ipz = np.zeros(300, dtype=np.int32)
    fla.lib().dgetrf_(C.byref(C.c_int(300)), C.byref(C.c_int(300)), dp(Z), C.byref(C.c_int(300)),
                      ip(ipz), C.byref(info))
    assert info.value == 124


def test_device_generators_match_host(orc):
    d = fla.DeviceMatrix(300, 200)
    fla.check(fla.lib().fla_hip_gen_uniform(42, 300, 200, d.ptr, 300))
    fla.synchronize()
    assert np.array_equal(d.download(), orc.gen_uniform(42, 300, 200))
    fla.check(fla.lib().fla_hip_gen_normal(42, 300, 200, d.ptr, 300))
    fla.synchronize()
    assert np.allclose(d.download(), orc.gen_normal(42, 300, 200), rtol=1e-13, atol=1e-13)
