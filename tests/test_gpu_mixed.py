"""Mixed-precision Cholesky solve (fla_hip_mp_chol / fla_hip_dsposv): the
reduced-precision factor, dsposv_'s refinement semantics
(src/map/common/lapacksrc/f2c/dsposv.c:192-415) and GMRES-IR.

Pins: the reduced-precision factor of an exactly representable SPD matrix is
exact (integers and small dyadics survive FP32 storage, BF16 operands and FP32
accumulation), which fixes the packed-operand layout and the update; the
solutions are compared with the FP64 oracle solve and must meet the stopping
rule's normwise backward error; the fall-back paths are forced by inputs that
overflow FP32 or break the reduced-precision factor.  The iteration counts are
not compared with the reference's dsposv_ (parity unpinned: its FP32 factor
differs from this one by construction)."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from cases import exact_spd
from gpu_util import to_dev

pytestmark = pytest.mark.gpu
L, U = fla.FLA_LOWER_TRIANGULAR, fla.FLA_UPPER_TRIANGULAR
EPS = 2.0 ** -53


def mp_chol(A, precision, uplo=L):
    n = A.shape[0]
    dA, dL = to_dev(A), fla.DeviceMatrix(n, n)
    info = C.c_int()
    st = fla.lib().fla_hip_mp_chol(uplo, n, dA.ptr, dA.ld, precision, dL.ptr, dL.ld,
                                   C.byref(info))
    return st, info.value, np.tril(dL.download())


def reflector_spd(orc, seed, n, nrefl, kappa):
    """numpy restatement of fla_hip_gen_spd_reflectors."""
    s = kappa ** (np.arange(n) / (n - 1))
    A = np.diag(s)
    for r in range(nrefl):
        v = orc.gen_normal(seed + r + 1, n, 1)[:, 0]
        w = A @ v
        tau = 2.0 / (v @ v)
        z = tau * w - 0.5 * tau * tau * (v @ w) * v
        A = A - np.outer(v, z) - np.outer(z, v)
    return A, s


def dev_reflector_spd(seed, n, nrefl, kappa):
    dA = fla.DeviceMatrix(n, n)
    fla.check(fla.lib().fla_hip_gen_spd_reflectors(seed, n, nrefl, kappa, dA.ptr, dA.ld))
    fla.synchronize()
    return dA


@pytest.fixture
def substitution_panel():
    """Exact closed forms need the backward-stable panel solve: the factors of
    these integer matrices have astronomically large inverses."""
    fla.lib().fla_hip_mp_set_panel(0)
    yield
    fla.lib().fla_hip_mp_set_panel(1)


@pytest.mark.parametrize("precision", [fla.FLA_MP_FP32, fla.FLA_MP_BF16])
@pytest.mark.parametrize("n", [40, 300, 700])
def test_mp_chol_exact(precision, n, substitution_panel):
    A, B = exact_spd(n, 5)
    st, info, Lf = mp_chol(A, precision)
    assert st == fla.FLA_HIP_OK and info == fla.FLA_SUCCESS
    assert np.array_equal(Lf, B)


@pytest.mark.parametrize("precision", [fla.FLA_MP_FP32, fla.FLA_MP_BF16])
def test_mp_chol_exact_upper(precision, substitution_panel):
    A, B = exact_spd(333, 9)
    st, info, Lf = mp_chol(A, precision, uplo=U)
    assert st == fla.FLA_HIP_OK and info == fla.FLA_SUCCESS
    assert np.array_equal(Lf, B)


@pytest.mark.parametrize("inverse", [0, 1])
def test_mp_chol_panel_modes_agree(orc, inverse):
    """Both panel solves give a factor of FP32 quality on a well-conditioned
    matrix (the inverse of a diagonal factor is then well conditioned)."""
    n = 900
    A = orc.gen_spd_c2(12, n)
    fla.lib().fla_hip_mp_set_panel(inverse)
    try:
        st, info, Lf = mp_chol(A, fla.FLA_MP_FP32)
    finally:
        fla.lib().fla_hip_mp_set_panel(1)
    assert st == fla.FLA_HIP_OK and info == fla.FLA_SUCCESS
    assert np.linalg.norm(A - Lf @ Lf.T, 1) / np.linalg.norm(A, 1) <= 8 * 2.0 ** -24


@pytest.mark.parametrize("precision,bound", [(fla.FLA_MP_FP32, 2.0 ** -24),
                                             (fla.FLA_MP_BF16, 2.0 ** -8)])
def test_mp_chol_accuracy(orc, precision, bound):
    """||A - L L'||_1 / ||A||_1 of the reduced-precision factor is of the
    order of the update format's unit roundoff (FP32 2^-24, BF16 2^-8)."""
    n = 1000
    A = orc.gen_spd_c2(3, n)
    st, info, Lf = mp_chol(A, precision)
    assert st == fla.FLA_HIP_OK and info == fla.FLA_SUCCESS
    rel = np.linalg.norm(A - Lf @ Lf.T, 1) / np.linalg.norm(A, 1)
    assert rel <= 8 * bound
    assert rel >= 2.0 ** -52  # a reduced-precision factor, not an FP64 one


def test_mp_chol_failure_index(substitution_panel):
    """Breakdown is reported at the first non-positive pivot, as FLA_Chol."""
    A, _ = exact_spd(300, 2)
    A[150, 150] = -1.0
    st, info, _ = mp_chol(A, fla.FLA_MP_FP32)
    assert st == fla.FLA_HIP_OK and info == 150


def test_mp_chol_overflow():
    A, _ = exact_spd(64, 2)
    A[10, 3] = A[3, 10] = 1e39
    st, info, _ = mp_chol(A, fla.FLA_MP_FP32)
    assert st == fla.FLA_HIP_ERANGE


def test_reflector_generator(orc):
    n, kappa = 200, 1e4
    dA = dev_reflector_spd(17, n, 64, kappa)
    A = dA.download()
    Ah, s = reflector_spd(orc, 17, n, 64, kappa)
    assert np.max(np.abs(A - Ah)) <= 1e-11 * kappa
    assert np.allclose(A, A.T, rtol=0, atol=1e-11 * kappa)
    ev = np.linalg.eigvalsh(A)
    assert np.allclose(ev, s, rtol=1e-9, atol=0)


def run_dsposv(A, B, uplo=L, **kw):
    dA, dB = to_dev(A), to_dev(B)
    dX = fla.DeviceMatrix(B.shape[0], B.shape[1])
    it, info, st = fla.dsposv(uplo, dA, dB, dX, **kw)
    return it, info, st, dX.download(), dA.download()


def backward_error(A, X, B):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf): the
    quantity dsposv_'s stopping rule bounds by eps * sqrt(n)."""
    R = B - A @ X
    an = np.max(np.sum(np.abs(A), axis=1))
    return max(np.max(np.abs(R[:, c])) / (an * np.max(np.abs(X[:, c]))) for c in range(B.shape[1]))


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("uplo", [L, U])
def test_dsposv_fp32_classic(orc, kappa, uplo):
    n = 1024
    A = dev_reflector_spd(7, n, 64, kappa).download()
    B = orc.gen_normal(99, n, 8)
    Astore = np.tril(A) if uplo == L else np.triu(A)  # only the uplo triangle is read
    it, info, st, X, Aout = run_dsposv(Astore, B, uplo)
    assert info == 0
    assert 0 <= it <= 30, it
    assert np.array_equal(Aout, Astore)  # untouched when refinement succeeded
    assert backward_error(A, X, B) <= EPS * np.sqrt(n)
    Xref = np.linalg.solve(A, B)
    fwd = np.max(np.abs(X - Xref)) / np.max(np.abs(Xref))
    assert fwd <= 50 * kappa * EPS


@pytest.mark.parametrize("refinement", [fla.FLA_IR_CLASSIC, fla.FLA_IR_GMRES])
def test_dsposv_stop_factor(orc, refinement):
    """A tighter stopping factor is honoured: backward error <= 4 eps."""
    n = 1024
    A = dev_reflector_spd(7, n, 64, 1e6).download()
    B = orc.gen_normal(98, n, 8)
    it, info, st, X, _ = run_dsposv(A, B, refinement=refinement, stop_factor=4.0)
    assert info == 0 and 0 <= it <= 30, it
    assert backward_error(A, X, B) <= 4 * EPS


@pytest.mark.parametrize("kappa", [1e2, 1e4])
def test_dsposv_bf16_gmres(orc, kappa):
    n = 1024
    A = dev_reflector_spd(8, n, 64, kappa).download()
    B = orc.gen_normal(5, n, 8)
    it, info, st, X, _ = run_dsposv(A, B, precision=fla.FLA_MP_BF16,
                                    refinement=fla.FLA_IR_GMRES)
    assert info == 0 and 0 <= it <= 30, it
    assert backward_error(A, X, B) <= EPS * np.sqrt(n)
    assert st["gmres_iters"] > 0


def test_dsposv_bf16_classic_small_kappa(orc):
    n = 768
    A = dev_reflector_spd(9, n, 64, 10.0).download()
    B = orc.gen_normal(6, n, 3)
    it, info, st, X, _ = run_dsposv(A, B, precision=fla.FLA_MP_BF16)
    assert info == 0 and 0 <= it <= 30, it
    assert backward_error(A, X, B) <= EPS * np.sqrt(n)


def test_dsposv_fallback_on_breakdown(orc):
    """SPD in FP64 but not after rounding to FP32 -> iter = -3, FP64 path,
    A overwritten by the FP64 factor (dsposv.c:318-322, 403-405)."""
    n = 200
    Q = np.linalg.qr(orc.gen_normal(3, n, n))[0]
    s = np.ones(n)
    s[-1] = 1e-12
    A = (Q * s) @ Q.T
    A = (A + A.T) / 2
    B = orc.gen_normal(4, n, 2)
    it, info, st, X, Aout = run_dsposv(A, B)
    if it >= 0:
        pytest.skip("matrix factored in FP32 after all (rounding luck)")
    assert it in (-3, -31)
    assert info == 0
    Lf = np.tril(Aout)
    assert np.linalg.norm(A - Lf @ Lf.T, 1) / (np.linalg.norm(A, 1) * n * EPS) <= 10


def test_dsposv_fallback_on_overflow(orc):
    n = 100
    A = orc.gen_spd_c2(1, n) * 1e35  # fits in FP64, overflows FP32 (max 3.4e38)
    A[0, 0] *= 1e4
    B = orc.gen_normal(2, n, 2)
    it, info, st, X, Aout = run_dsposv(A, B)
    assert it == -2 and info == 0
    assert backward_error(A, X, B) <= 10 * n * EPS


def test_dsposv_not_spd():
    A, _ = exact_spd(120, 4)
    A[60, 60] = -5.0
    B = np.ones((120, 1))
    it, info, st, X, Aout = run_dsposv(A, B)
    assert it == -3
    assert info == 61  # LAPACK convention: order of the failing leading minor


def test_dsposv_zero_rhs_count():
    A, _ = exact_spd(10, 1)
    it, info, st, X, _ = run_dsposv(A, np.zeros((10, 1)))
    assert it == 0 and info == 0 and np.all(X == 0)


def test_lapack2flame_dsposv(orc):
    """dsposv_ with host arrays: solution, iter >= 0 and A unchanged; a
    matrix that is not positive definite gives iter = -3 and info > 0."""
    n, r = 300, 4
    A = dev_reflector_spd(3, n, 16, 1e3).download()
    A = np.asfortranarray(A)
    B = np.asfortranarray(orc.gen_normal(8, n, r))
    X = np.zeros((n, r), order="F")
    A_in = A.copy(order="F")
    nn, rr, ld = C.c_int(n), C.c_int(r), C.c_int(n)
    it, info = C.c_int(), C.c_int()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    fla.lib().dsposv_(b"L", C.byref(nn), C.byref(rr), dp(A_in), C.byref(ld), dp(B), C.byref(ld),
                      dp(X), C.byref(ld), None, None, C.byref(it), C.byref(info))
    assert info.value == 0 and 0 <= it.value <= 30
    assert np.array_equal(A_in, A)
    assert backward_error(A, X, B) <= EPS * np.sqrt(n)
    Abad = A.copy(order="F")
    Abad[100, 100] = -1.0
    fla.lib().dsposv_(b"L", C.byref(nn), C.byref(rr), dp(Abad), C.byref(ld), dp(B), C.byref(ld),
                      dp(X), C.byref(ld), None, None, C.byref(it), C.byref(info))
    assert it.value == -3 and info.value > 0
