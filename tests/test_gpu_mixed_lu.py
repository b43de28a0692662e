"""Mixed-precision LU solve (fla_hip_mp_lu / fla_hip_dsgesv / dsgesv_): the
reduced-precision LU with partial pivoting and dsgesv_'s refinement
semantics (src/map/common/lapacksrc/f2c/dsgesv.c:189-407).

Pins: the reduced-precision LU of a matrix with exactly representable
factors recovers the designed P, L and U bit for bit in FP32 and BF16 (which
fixes the pivot handling, the transposed U12 operand and the update); on
well-separated pivots the pivot vector equals the FP64 oracle's; solutions
must meet the stopping rule's normwise backward error and are compared with
an FP64 solve; the fall-back paths are forced by a singular matrix and by an
entry outside the FP32 range.  The general reflector generator is compared
with a numpy restatement and its singular values with the closed form.
Iteration counts are not compared with the reference's dsgesv_ (parity
unpinned: its FP32 factor differs from this one by construction)."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from cases import exact_lu
from gpu_util import to_dev

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def mp_lu(A, precision):
    n = A.shape[0]
    dA, dF = to_dev(A), fla.DeviceMatrix(n, n)
    dp = fla.DeviceMatrix(n, 1, np.int32)
    info = C.c_int()
    st = fla.lib().fla_hip_mp_lu(n, dA.ptr, dA.ld, precision, dF.ptr, dF.ld, dp.ptr,
                                 C.byref(info))
    return st, info.value, dF.download(), dp.download().ravel().copy()


def reflector_ge(orc, seed, seed2, n, nrefl, kappa):
    """numpy restatement of fla_hip_gen_ge_reflectors."""
    s = kappa ** (np.arange(n) / (n - 1))
    A = np.diag(s)
    for r in range(nrefl):
        v = orc.gen_normal(seed + r + 1, n, 1)[:, 0]
        A = A - (2.0 / (v @ v)) * np.outer(v, A.T @ v)
        u = orc.gen_normal(seed2 + r + 1, n, 1)[:, 0]
        A = A - (2.0 / (u @ u)) * np.outer(A @ u, u)
    return A, s


def dev_reflector_ge(seed, seed2, n, nrefl, kappa):
    dA = fla.DeviceMatrix(n, n)
    fla.check(fla.lib().fla_hip_gen_ge_reflectors(seed, seed2, n, nrefl, kappa, dA.ptr, dA.ld))
    fla.synchronize()
    return dA


@pytest.mark.parametrize("precision", [fla.FLA_MP_FP32, fla.FLA_MP_BF16])
@pytest.mark.parametrize("n", [40, 300, 700])
def test_mp_lu_exact(precision, n):
    A, Lx, Ux, px = exact_lu(n, 11 + n)
    st, info, F, p = mp_lu(A, precision)
    assert st == fla.FLA_HIP_OK and info == fla.FLA_SUCCESS
    assert np.array_equal(p, px)
    assert np.array_equal(np.tril(F, -1) + np.eye(n), Lx)
    assert np.array_equal(np.triu(F), Ux)


def test_mp_lu_pivots_vs_oracle(orc):
    """Well-separated pivots: the FP32 factorization picks the FP64 oracle's
    pivots, and its factor is of FP32 quality."""
    n = 400
    A = orc.gen_lu_c2(5, n, n)
    st, info, F, p = mp_lu(A, fla.FLA_MP_FP32)
    assert st == fla.FLA_HIP_OK and info == fla.FLA_SUCCESS
    _, pref, _ = orc.lu_piv(A)
    assert np.array_equal(p, pref)
    res = orc.lu_residual(A, F, p)  # ||PA - LU|| / (n eps ||A||) in FP64 units
    assert res * n * EPS <= 32 * 2.0 ** -24


def test_mp_lu_null_pivot():
    A, _, _, _ = exact_lu(200, 3, null_col=77)
    st, info, _, _ = mp_lu(A, fla.FLA_MP_FP32)
    assert st == fla.FLA_HIP_OK and info == 77


def test_mp_lu_overflow():
    A, _, _, _ = exact_lu(64, 4)
    A[5, 9] = -1e39
    st, _, _, _ = mp_lu(A, fla.FLA_MP_FP32)
    assert st == fla.FLA_HIP_ERANGE


def test_ge_reflector_generator(orc):
    n, kappa = 180, 1e4
    A = dev_reflector_ge(17, 1017, n, 16, kappa).download()
    Ah, s = reflector_ge(orc, 17, 1017, n, 16, kappa)
    assert np.max(np.abs(A - Ah)) <= 1e-11 * kappa
    sv = np.sort(np.linalg.svd(A, compute_uv=False))
    assert np.allclose(sv, s, rtol=1e-9, atol=0)
    assert not np.allclose(A, A.T)  # independent left and right products


def run_dsgesv(A, B, **kw):
    n = A.shape[0]
    dA, dB = to_dev(A), to_dev(B)
    dX = fla.DeviceMatrix(n, B.shape[1])
    dp = fla.DeviceMatrix(n, 1, np.int32)
    it, info, st = fla.dsgesv(dA, dp, dB, dX, **kw)
    return it, info, st, dX.download(), dA.download(), dp.download().ravel().copy()


def backward_error(A, X, B):
    R = B - A @ X
    an = np.max(np.sum(np.abs(A), axis=1))
    return max(np.max(np.abs(R[:, c])) / (an * np.max(np.abs(X[:, c]))) for c in range(B.shape[1]))


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_dsgesv_fp32_classic(orc, kappa):
    n = 1024
    A = dev_reflector_ge(7, 1007, n, 32, kappa).download()
    B = orc.gen_normal(99, n, 8)
    it, info, st, X, Aout, p = run_dsgesv(A, B)
    assert info == 0
    assert 0 <= it <= 30, it
    assert np.array_equal(Aout, A)  # untouched when refinement succeeded
    assert backward_error(A, X, B) <= EPS * np.sqrt(n)
    Xref = np.linalg.solve(A, B)
    fwd = np.max(np.abs(X - Xref)) / np.max(np.abs(Xref))
    # the stopping rule's backward error eps*sqrt(n), amplified by kappa
    assert fwd <= 2 * kappa * np.sqrt(n) * EPS
    assert np.all(p >= 0) and np.all(p + np.arange(n) < n)


@pytest.mark.parametrize("precision", [fla.FLA_MP_FP32, fla.FLA_MP_BF16])
def test_dsgesv_gmres(orc, precision):
    n = 1024
    A = dev_reflector_ge(8, 1008, n, 32, 1e4).download()
    B = orc.gen_normal(5, n, 8)
    it, info, st, X, _, _ = run_dsgesv(A, B, precision=precision, refinement=fla.FLA_IR_GMRES)
    assert info == 0 and 0 <= it <= 30, it
    assert backward_error(A, X, B) <= EPS * np.sqrt(n)
    assert st["gmres_iters"] > 0


def test_dsgesv_fallback_singular(orc):
    """A null pivot of the reduced-precision LU -> iter = -3, then the FP64
    dgetrf finds the same exact zero pivot: info = k + 1 (dsgesv.c:397-402)."""
    A, _, _, _ = exact_lu(150, 6, null_col=40)
    B = orc.gen_normal(4, 150, 2)
    it, info, _, _, Aout, p = run_dsgesv(A, B)
    assert it == -3
    assert info == 41
    F, pref, _ = orc.lu_piv(A)
    assert np.array_equal(p, pref)
    assert np.array_equal(Aout, F)


def test_dsgesv_fallback_overflow(orc):
    n = 96
    A = orc.gen_lu_c2(2, n, n)
    A[3, 7] = 3e38 * 4  # beyond FLT_MAX
    B = orc.gen_normal(8, n, 3)
    it, info, _, X, Aout, p = run_dsgesv(A, B)
    assert it == -2 and info == 0
    Xref = np.linalg.solve(A, B)
    assert np.allclose(X, Xref, rtol=0, atol=1e-10 * np.max(np.abs(Xref)))
    F, pref, _ = orc.lu_piv(A)
    assert np.array_equal(p, pref)


def test_dsgesv_lapack2flame(orc):
    n, r = 300, 4
    A = np.asfortranarray(dev_reflector_ge(9, 1009, n, 8, 1e3).download())
    B = np.asfortranarray(orc.gen_normal(3, n, r))
    X = np.zeros((n, r), order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    it, info = C.c_int(), C.c_int()
    L = fla.lib()
    ip = lambda v: C.byref(C.c_int(v))
    dptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    A0 = A.copy(order="F")
    L.dsgesv_(ip(n), ip(r), dptr(A), ip(n), ipiv.ctypes.data_as(C.POINTER(C.c_int)), dptr(B),
              ip(n), dptr(X), ip(n), None, None, C.byref(it), C.byref(info))
    assert info.value == 0 and 0 <= it.value <= 30
    assert np.array_equal(A, A0)
    assert backward_error(A0, X, B) <= EPS * np.sqrt(n)
    # LAPACK pivots: 1-based absolute, row i swapped with ipiv[i] - 1 >= i
    assert np.all(ipiv >= np.arange(1, n + 1)) and np.all(ipiv <= n)
    # the same factorization through the C-ABI: the same pivots
    _, _, _, _, _, p = run_dsgesv(A0, B)
    assert np.array_equal(ipiv, p + np.arange(1, n + 1))


def test_dsgesv_zero_sizes():
    dA, dB, dX = fla.DeviceMatrix(4, 4), fla.DeviceMatrix(4, 0), fla.DeviceMatrix(4, 0)
    dp = fla.DeviceMatrix(4, 1, np.int32)
    it, info, _ = fla.dsgesv(dA, dp, dB, dX)
    assert it == 0 and info == 0
