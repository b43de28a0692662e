"""Pin the oracle's BLAS restatements.

* Level 1 and dlaswp: bit-for-bit against the reference's own f2c sources,
  compiled standalone into oracle/_ref (src/flablas/f2c/ddot.c, idamax.c,
  dscal.c; src/map/common/lapacksrc/f2c/dlaswp.c).
* Level 2/3 (dgemv, dsyrk, dtrsm, dgemm): the reference's f2c files need
  lsame_, whose source needs the configure-generated FLA_config.h, so they are
  pinned by exact known answers on dyadic inputs (every branch: side, uplo,
  trans, diag) and by a brute-force long-double comparison.
"""
import ctypes as C

import numpy as np
import pytest

from tests.cases import exact_tri


def _ref_or_skip(orc):
    r = orc.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return r


@pytest.mark.parametrize("n", [0, 1, 4, 5, 6, 9, 10, 11, 37, 128, 1001])
@pytest.mark.parametrize("incs", [(1, 1), (3, 1), (1, 2), (-1, 1), (2, -3)])
def test_ddot_bitwise_vs_reference(orc, n, incs):
    ref = _ref_or_skip(orc)
    incx, incy = incs
    rng = np.random.default_rng(n * 7 + abs(incx))
    x = rng.standard_normal(max(1, n * abs(incx)))
    y = rng.standard_normal(max(1, n * abs(incy)))
    got = orc.lib().or_ddot(n, orc.dp(x), incx, orc.dp(y), incy)
    want = ref.ddot_(C.byref(C.c_int(n)), orc.dp(x), C.byref(C.c_int(incx)), orc.dp(y),
                     C.byref(C.c_int(incy)))
    assert np.float64(got).tobytes() == np.float64(want).tobytes()


@pytest.mark.parametrize("case", ["random", "ties", "nan_first", "nan_mid", "nan_last", "all_zero",
                                  "inf", "n1", "n0"])
@pytest.mark.parametrize("inc", [1, 2])
def test_idamax_vs_reference(orc, case, inc):
    ref = _ref_or_skip(orc)
    rng = np.random.default_rng(3)
    n = 50
    x = rng.uniform(-1, 1, n)
    if case == "ties":
        x[[7, 19, 31]] = [2.0, -2.0, 2.0]
    elif case == "nan_first":
        x[0] = np.nan
    elif case == "nan_mid":
        x[20] = np.nan
        x[5] = 3.0
    elif case == "nan_last":
        x[-1] = np.nan
    elif case == "all_zero":
        x[:] = 0.0
    elif case == "inf":
        x[[10, 40]] = [-np.inf, np.inf]
    elif case == "n1":
        n = 1
    elif case == "n0":
        n = 0
    xs = np.repeat(x, inc)  # stride inc over the same values
    nn = n if case not in ("n1", "n0") else n
    got = orc.lib().or_idamax(nn, orc.dp(xs), inc)
    want = ref.idamax_(C.byref(C.c_int(nn)), orc.dp(xs), C.byref(C.c_int(inc)))
    assert got == want


def test_dscal_bitwise_vs_reference(orc):
    ref = _ref_or_skip(orc)
    rng = np.random.default_rng(5)
    for n in (0, 1, 4, 5, 13, 100):
        x = rng.standard_normal(n + 1)
        y = x.copy()
        a = 1.0 / 3.0
        orc.lib().or_dscal(n, a, orc.dp(x), 1)
        ref.dscal_(C.byref(C.c_int(n)), C.byref(C.c_double(a)), orc.dp(y), C.byref(C.c_int(1)))
        assert x.tobytes() == y.tobytes()


@pytest.mark.parametrize("incx", [1, -1, 2, -2, 3])
@pytest.mark.parametrize("forward_only", [True, False])
def test_dlaswp_vs_reference(orc, incx, forward_only):
    """The strided forms index ipiv as dlaswp.c:130-145 does (incx > 0 from
    k1, incx < 0 from 1); ipiv may point above the row as well."""
    ref = _ref_or_skip(orc)
    rng = np.random.default_rng(11 + incx)
    m, n = 70, 45
    A = np.asfortranarray(rng.standard_normal((m, n)))
    k1, k2 = 3, 60
    ipiv = np.zeros(k2 * abs(incx) + 1, dtype=np.int32)
    for i in range(len(ipiv)):
        lo = min(i, m - 1) + 1 if forward_only else 1
        ipiv[i] = rng.integers(lo, m + 1)  # 1-based
    A1 = A.copy(order="F")
    A2 = A.copy(order="F")
    orc.lib().or_dlaswp(n, orc.dp(A1), m, k1, k2, orc.ip(ipiv), incx)
    ref.dlaswp_(C.byref(C.c_int(n)), orc.dp(A2), C.byref(C.c_int(m)), C.byref(C.c_int(k1)),
                C.byref(C.c_int(k2)), orc.ip(ipiv), C.byref(C.c_int(incx)))
    assert A1.tobytes() == A2.tobytes()


# ---- level 3: exact known answers --------------------------------------

@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_dtrsm_exact(orc, side, uplo, trans, diag):
    m, n = 37, 29
    dim = m if side == "L" else n
    T = exact_tri(dim, 17, lower=(uplo == "L"), unit=(diag == "U"))
    rng = np.random.default_rng(23)
    X = rng.integers(-3, 4, size=(m, n)).astype(np.float64)
    opT = T if trans == "N" else T.T
    alpha = 2.0
    B = (opT @ X) / alpha if side == "L" else (X @ opT) / alpha
    got = orc.dtrsm(side, uplo, trans, diag, T, B, alpha=alpha)
    assert np.array_equal(got, X)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 1.0, -2.0])
def test_dsyrk_exact(orc, uplo, trans, beta):
    rng = np.random.default_rng(31)
    n, k = 41, 19
    Am = rng.integers(-5, 6, size=(n, k) if trans == "N" else (k, n)).astype(np.float64)
    C0 = rng.integers(-9, 10, size=(n, n)).astype(np.float64)
    prod = Am @ Am.T if trans == "N" else Am.T @ Am
    want = -0.5 * prod + beta * C0
    got = orc.dsyrk(uplo, trans, Am, C0, alpha=-0.5, beta=beta)
    tri = np.tril if uplo == "L" else np.triu
    other = np.triu if uplo == "L" else np.tril
    assert np.array_equal(tri(got), tri(want))
    off = 1 if uplo == "L" else -1
    assert np.array_equal(other(got, off), other(C0, off))  # other triangle untouched


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("beta", [0.0, 1.0, 3.0])
def test_dgemm_exact(orc, ta, tb, beta):
    rng = np.random.default_rng(37)
    m, n, k = 23, 31, 17
    Am = rng.integers(-5, 6, size=(m, k) if ta == "N" else (k, m)).astype(np.float64)
    Bm = rng.integers(-5, 6, size=(k, n) if tb == "N" else (n, k)).astype(np.float64)
    C0 = rng.integers(-9, 10, size=(m, n)).astype(np.float64)
    opA = Am if ta == "N" else Am.T
    opB = Bm if tb == "N" else Bm.T
    want = -1.0 * (opA @ opB) + beta * C0
    got = orc.dgemm(ta, tb, Am, Bm, C0, alpha=-1.0, beta=beta)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgemv_exact(orc, trans):
    rng = np.random.default_rng(41)
    m, n = 29, 13
    Am = np.asfortranarray(rng.integers(-5, 6, size=(m, n)).astype(np.float64))
    lx, ly = (n, m) if trans == "N" else (m, n)
    x = rng.integers(-4, 5, size=lx).astype(np.float64)
    y = rng.integers(-4, 5, size=ly).astype(np.float64)
    want = -1.0 * ((Am if trans == "N" else Am.T) @ x) + 1.0 * y
    orc.lib().or_dgemv(trans.encode(), m, n, -1.0, orc.dp(Am), m, orc.dp(x), 1, 1.0, orc.dp(y), 1)
    assert np.array_equal(y, want)


def test_dgemm_rounding_brute_force(orc):
    """Random floats: the f2c-order product is within the k*eps bound of a
    long-double brute-force triple loop."""
    rng = np.random.default_rng(43)
    m, n, k = 20, 18, 64
    Am, Bm, C0 = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
    got = orc.dgemm("N", "N", Am, Bm, C0)
    want = C0.astype(np.longdouble) - Am.astype(np.longdouble) @ Bm.astype(np.longdouble)
    bound = (k + 2) * 2.0**-53 * (np.abs(Am) @ np.abs(Bm) + np.abs(C0))
    assert np.all(np.abs(got - want.astype(np.float64)) <= bound)
