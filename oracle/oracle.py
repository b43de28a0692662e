"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
import this module; the libflame_amd product never does.  It loads
``oracle/liboracle.so`` (the C restatement of the reference's FLA_Chol /
FLA_LU_piv control trees) and, when built, ``oracle/_ref/libflame_ref_f2c.so``
(the reference's own self-contained f2c BLAS routines).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SUCCESS = -1
LOWER, UPPER = 300, 301
NATIVE_PIVOTS, LAPACK_PIVOTS = 700, 701
NB_CHOL = (128, 32)  # FLA_Chol_cntl_init.c:25-30 (128, scaled by 0.25)
NB_LU = (128, 16)    # FLA_LU_piv_cntl_init.c:26-31 (128, scaled by 0.125)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run `make -C oracle`")
    return C.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        L = _load(os.path.join(_HERE, "liboracle.so"))
        L.or_ddot.restype = C.c_double
        L.or_ddot.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.or_idamax.restype = C.c_int
        L.or_idamax.argtypes = [C.c_int, _dp, C.c_int]
        L.or_dscal.argtypes = [C.c_int, C.c_double, _dp, C.c_int]
        L.or_dgemv.argtypes = [C.c_char, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                               _dp, C.c_int, C.c_double, _dp, C.c_int]
        L.or_dsyrk.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_double, _dp,
                               C.c_int, C.c_double, _dp, C.c_int]
        L.or_dtrsm.argtypes = [C.c_char, C.c_char, C.c_char, C.c_char, C.c_int, C.c_int,
                               C.c_double, _dp, C.c_int, _dp, C.c_int]
        L.or_dgemm.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double,
                               _dp, C.c_int, _dp, C.c_int, C.c_double, _dp, C.c_int]
        L.or_dlaswp.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip, C.c_int]
        L.or_chol.restype = C.c_int
        L.or_chol.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int]
        L.or_lu_piv.restype = C.c_int
        L.or_lu_piv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, C.c_int]
        L.or_apply_pivots_ln.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip, C.c_int]
        L.or_shift_pivots_to.argtypes = [C.c_int, C.c_int, _ip]
        L.or_chol_solve.restype = C.c_int
        L.or_chol_solve.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                    _dp, C.c_int]
        L.or_lu_piv_solve.restype = C.c_int
        L.or_lu_piv_solve.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _dp, C.c_int,
                                      _dp, C.c_int]
        L.or_dpotrf.argtypes = [C.c_char, C.c_int, _dp, C.c_int, _ip]
        L.or_dgetrf.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _ip]
        L.or_dpotrs.argtypes = [C.c_char, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _ip]
        L.or_dgetrs.argtypes = [C.c_char, C.c_int, C.c_int, _dp, C.c_int, _ip, _dp, C.c_int,
                                _ip]
        L.or_rand_u64.restype = C.c_uint64
        L.or_rand_u64.argtypes = [C.c_uint64, C.c_uint64]
        L.or_uniform_pm1.restype = C.c_double
        L.or_uniform_pm1.argtypes = [C.c_uint64, C.c_uint64]
        L.or_normal.restype = C.c_double
        L.or_normal.argtypes = [C.c_uint64, C.c_uint64]
        for f in ("or_gen_uniform", "or_gen_normal"):
            getattr(L, f).argtypes = [C.c_uint64, C.c_int, C.c_int, _dp, C.c_int]
        for f in ("or_gen_spd_c1", "or_gen_spd_c2"):
            getattr(L, f).argtypes = [C.c_uint64, C.c_int, _dp, C.c_int]
        L.or_gen_lu_c2.argtypes = [C.c_uint64, C.c_int, C.c_int, _dp, C.c_int]
        L.or_lu_nopiv.restype = C.c_int
        L.or_lu_nopiv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int]
        L.or_trinv.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int]
        L.or_ttmm.argtypes = [C.c_int, C.c_int, _dp, C.c_int]
        L.or_spdinv.restype = C.c_int
        L.or_spdinv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int]
        L.or_chol_residual.restype = C.c_double
        L.or_chol_residual.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.or_lu_residual.restype = C.c_double
        L.or_lu_residual.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _ip]
        _lib = L
    return _lib


def ref():
    """The reference's own f2c routines that compile standalone, or None."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libflame_ref_f2c.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        I = C.POINTER(C.c_int)
        R.ddot_.restype = C.c_double
        R.ddot_.argtypes = [I, _dp, I, _dp, I]
        R.idamax_.restype = C.c_int
        R.idamax_.argtypes = [I, _dp, I]
        R.dscal_.argtypes = [I, _dp, _dp, I]
        R.dlaswp_.argtypes = [I, _dp, I, I, I, I, I]
        _ref = R
    return _ref


def dp(a):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def ip(a):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


# ---- convenience wrappers (arrays are Fortran-ordered float64) -------------

def chol(A, uplo=LOWER, nb=NB_CHOL):
    A = fortran(A).copy(order="F")
    nbv = np.array(nb, dtype=np.int32)
    r = lib().or_chol(uplo, A.shape[0], dp(A), max(1, A.shape[0]), ip(nbv), len(nb))
    return A, r


def lu_nopiv(A, nb=NB_CHOL):
    """FLA_LU_nopiv (blocks 128 and 32, as fla_lu_nopiv_cntl2)."""
    A = fortran(A).copy(order="F")
    nbv = np.array(nb, dtype=np.int32)
    r = lib().or_lu_nopiv(A.shape[0], A.shape[1], dp(A), max(1, A.shape[0]), ip(nbv), len(nb))
    return A, r


def trinv(A, uplo=LOWER, diag=501):
    A = fortran(A).copy(order="F")
    lib().or_trinv(uplo, diag, A.shape[0], dp(A), max(1, A.shape[0]))
    return A


def ttmm(A, uplo=LOWER):
    A = fortran(A).copy(order="F")
    lib().or_ttmm(uplo, A.shape[0], dp(A), max(1, A.shape[0]))
    return A


def spdinv(A, uplo=LOWER, nb=NB_CHOL):
    A = fortran(A).copy(order="F")
    nbv = np.array(nb, dtype=np.int32)
    r = lib().or_spdinv(uplo, A.shape[0], dp(A), max(1, A.shape[0]), ip(nbv), len(nb))
    return A, r


def lu_piv(A, nb=NB_LU):
    A = fortran(A).copy(order="F")
    m, n = A.shape
    p = np.zeros(min(m, n), dtype=np.int32)
    nbv = np.array(nb, dtype=np.int32)
    r = lib().or_lu_piv(m, n, dp(A), max(1, m), ip(p), ip(nbv), len(nb))
    return A, p, r


def chol_solve(F, B, uplo=LOWER):
    F = fortran(F)
    B = fortran(B)
    X = np.zeros_like(B, order="F")
    lib().or_chol_solve(uplo, F.shape[0], B.shape[1], dp(F), max(1, F.shape[0]), dp(B),
                        max(1, B.shape[0]), dp(X), max(1, X.shape[0]))
    return X


def lu_piv_solve(F, p, B):
    F = fortran(F)
    B = fortran(B)
    X = np.zeros_like(B, order="F")
    p = np.ascontiguousarray(p, dtype=np.int32)
    lib().or_lu_piv_solve(F.shape[0], B.shape[1], dp(F), max(1, F.shape[0]), ip(p), dp(B),
                          max(1, B.shape[0]), dp(X), max(1, X.shape[0]))
    return X


def apply_pivots_ln(p, A):
    A = fortran(A).copy(order="F")
    p = np.ascontiguousarray(p, dtype=np.int32)
    lib().or_apply_pivots_ln(A.shape[1], dp(A), max(1, A.shape[0]), 0, len(p) - 1, ip(p), 1)
    return A


def dtrsm(side, uplo, trans, diag, Amat, B, alpha=1.0):
    Amat = fortran(Amat)
    B = fortran(B).copy(order="F")
    m, n = B.shape
    lib().or_dtrsm(side.encode(), uplo.encode(), trans.encode(), diag.encode(), m, n, alpha,
                   dp(Amat), max(1, Amat.shape[0]), dp(B), max(1, m))
    return B


def dsyrk(uplo, trans, Amat, Cm, alpha=-1.0, beta=1.0):
    Amat = fortran(Amat)
    Cm = fortran(Cm).copy(order="F")
    n = Cm.shape[0]
    k = Amat.shape[1] if trans.upper() == "N" else Amat.shape[0]
    lib().or_dsyrk(uplo.encode(), trans.encode(), n, k, alpha, dp(Amat), max(1, Amat.shape[0]),
                   beta, dp(Cm), max(1, n))
    return Cm


def dgemm(ta, tb, Amat, Bmat, Cm, alpha=-1.0, beta=1.0):
    Amat, Bmat = fortran(Amat), fortran(Bmat)
    Cm = fortran(Cm).copy(order="F")
    m, n = Cm.shape
    k = Amat.shape[1] if ta.upper() == "N" else Amat.shape[0]
    lib().or_dgemm(ta.encode(), tb.encode(), m, n, k, alpha, dp(Amat), max(1, Amat.shape[0]),
                   dp(Bmat), max(1, Bmat.shape[0]), beta, dp(Cm), max(1, m))
    return Cm


def gen_uniform(seed, m, n):
    A = np.zeros((m, n), dtype=np.float64, order="F")
    lib().or_gen_uniform(seed, m, n, dp(A), max(1, m))
    return A


def gen_normal(seed, m, n):
    A = np.zeros((m, n), dtype=np.float64, order="F")
    lib().or_gen_normal(seed, m, n, dp(A), max(1, m))
    return A


def gen_spd_c1(seed, n):
    A = np.zeros((n, n), dtype=np.float64, order="F")
    lib().or_gen_spd_c1(seed, n, dp(A), max(1, n))
    return A


def gen_spd_c2(seed, n):
    A = np.zeros((n, n), dtype=np.float64, order="F")
    lib().or_gen_spd_c2(seed, n, dp(A), max(1, n))
    return A


def gen_lu_c2(seed, m, n):
    A = np.zeros((m, n), dtype=np.float64, order="F")
    lib().or_gen_lu_c2(seed, m, n, dp(A), max(1, m))
    return A


def chol_residual(A0, F, uplo=LOWER):
    A0, F = fortran(A0), fortran(F)
    n = A0.shape[0]
    return lib().or_chol_residual(uplo, n, dp(A0), max(1, n), dp(F), max(1, n))


def lu_residual(A0, F, p):
    A0, F = fortran(A0), fortran(F)
    m, n = A0.shape
    p = np.ascontiguousarray(p, dtype=np.int32)
    return lib().or_lu_residual(m, n, dp(A0), max(1, m), dp(F), max(1, m), ip(p))


def _incpiv_lib():
    L = lib()
    if not getattr(L, "_incpiv_sigs", False):
        L.or_lu_incpiv.restype = C.c_int
        L.or_lu_incpiv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip,
                                   C.c_int, _dp, C.c_int]
        L.or_fs_incpiv.restype = C.c_int
        L.or_fs_incpiv.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip,
                                   C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int]
        L._incpiv_sigs = True
    return L


def lu_incpiv(A, b, nb_alg):
    """FLASH_LU_incpiv on tiles of b (inner block nb_alg) of the flat A.
    Returns (A_out, p, L, r): p is the flattened pivot object (m x n/b ints,
    tile (i,k) at rows i*b, column k) and L the flattened L object
    (m x (n/b)*nb_alg); their tiles that the factorization does not write
    stay zero."""
    A = fortran(A).copy(order="F")
    m, n = A.shape
    nt = max(1, n // b)
    p = np.zeros((m, nt), dtype=np.int32, order="F")
    Lm = np.zeros((m, nt * nb_alg), order="F")
    r = _incpiv_lib().or_lu_incpiv(m, n, dp(A), max(1, m), b, nb_alg, ip(p), max(1, m), dp(Lm),
                                   max(1, m))
    return A, p, Lm, r


def fs_incpiv(F, p, Lm, b, nb_alg, X):
    """FLASH_FS_incpiv on the columns of X (returns a new array)."""
    F = fortran(F)
    X = fortran(X).copy(order="F")
    p = np.asfortranarray(p, dtype=np.int32)
    Lm = fortran(Lm)
    m, n = F.shape
    _incpiv_lib().or_fs_incpiv(m, n, dp(F), max(1, m), b, nb_alg, ip(p), max(1, m), dp(Lm),
                               max(1, m), X.shape[1], dp(X), max(1, X.shape[0]))
    return X


def lu_incpiv_solve(F, p, Lm, b, nb_alg, B):
    """FLASH_LU_incpiv_solve: forward substitution, then the upper solve."""
    Y = fs_incpiv(F, p, Lm, b, nb_alg, B)
    return dtrsm("L", "U", "N", "N", F, Y)
