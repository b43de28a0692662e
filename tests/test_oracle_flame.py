"""Pin the oracle's FLAME layer (FLA_Chol, FLA_LU_piv, FLA_Apply_pivots,
solves, lapack2flame) by closed forms, invariants and brute force.

The reference's FLAME layer cannot be compiled here (every file includes the
configure-generated FLA_config.h), and the reference's tests hold no golden
vectors (test/src/test_chol.c, test_lu_piv.c are residual-only), so:
  * Cholesky of B*B^T with dyadic B returns B bit for bit (uniqueness);
  * GEPP on P^{-1} L U with |l_ij| < 1 returns p, L, U bit for bit;
  * failure indices are forced at a known column;
  * residual ratios (dpot01 / dget01 form) bound random problems;
  * Apply_pivots / Shift_pivots / dlaswp are exact permutations.
"""
import numpy as np
import pytest

from tests.cases import exact_lu, exact_spd, perm_from_pivots


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 127, 128, 129, 200, 300])
@pytest.mark.parametrize("upper", [False, True])
def test_chol_exact_closed_form(orc, n, upper):
    A, F = exact_spd(n, seed=n, upper=upper)
    out, r = orc.chol(A, uplo=orc.UPPER if upper else orc.LOWER)
    assert r == orc.SUCCESS
    tri = np.triu if upper else np.tril
    assert np.array_equal(tri(out), F)
    other = np.tril if upper else np.triu
    off = -1 if upper else 1
    assert np.array_equal(other(out, off), other(A, off))  # other triangle untouched


@pytest.mark.parametrize("k", [0, 5, 31, 32, 127, 128, 150, 199])
@pytest.mark.parametrize("kind", ["negative", "nan", "zero"])
@pytest.mark.parametrize("upper", [False, True])
def test_chol_failure_index(orc, k, kind, upper):
    n = 200
    A, F = exact_spd(n, seed=7, upper=upper)
    # The pivot at step k is F[k,k]^2 exactly; shift it to -1, 0 or NaN.
    if kind == "negative":
        A[k, k] -= F[k, k] ** 2 + 1.0
    elif kind == "zero":
        A[k, k] -= F[k, k] ** 2
    else:
        A[k, k] = np.nan
    _, r = orc.chol(A, uplo=orc.UPPER if upper else orc.LOWER)
    assert r == k


def test_chol_brute_force_tiny(orc):
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 4):
        B = rng.standard_normal((n, n))
        A = B @ B.T + n * np.eye(n)
        out, r = orc.chol(A)
        assert r == orc.SUCCESS
        assert np.allclose(np.tril(out), np.linalg.cholesky(A), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n", [64, 300, 1024])
def test_chol_residual_c1(orc, n):
    A = orc.gen_spd_c1(1234, n)
    out, r = orc.chol(A)
    assert r == orc.SUCCESS
    assert orc.chol_residual(A, out) <= 10.0
    outu, ru = orc.chol(A, uplo=orc.UPPER)
    assert ru == orc.SUCCESS
    assert orc.chol_residual(A, outu, uplo=orc.UPPER) <= 10.0


def test_chol_blocking_changes_rounding_only(orc):
    """The control tree only reorders roundings: unblocked and blocked
    results agree to a few ulps of the factor norm."""
    A = orc.gen_spd_c2(9, 260)
    a, _ = orc.chol(A, nb=(128, 32))
    b, _ = orc.chol(A, nb=())
    assert np.max(np.abs(np.tril(a) - np.tril(b))) < 1e-13 * np.max(np.abs(np.tril(b)))


@pytest.mark.parametrize("shape", [(1, 1), (5, 5), (16, 16), (17, 17), (128, 128), (129, 129),
                                   (200, 200), (300, 300), (250, 90), (90, 250)])
def test_lu_exact_closed_form(orc, shape):
    m, n = shape
    A, L, U, p = exact_lu(n, seed=m * 1000 + n, m=m)
    out, pv, r = orc.lu_piv(A)
    assert r == orc.SUCCESS
    assert np.array_equal(pv, p)
    k = min(m, n)
    assert np.array_equal(np.tril(out[:, :k], -1), np.tril(L, -1))
    assert np.array_equal(np.triu(out[:k, :]), U)


@pytest.mark.parametrize("c", [0, 3, 16, 100, 128, 199])
def test_lu_null_pivot_index(orc, c):
    n = 200
    A, L, U, p = exact_lu(n, seed=5, null_col=c)
    out, pv, r = orc.lu_piv(A)
    assert r == c
    assert np.array_equal(pv, p)
    assert np.array_equal(np.triu(out), U)


def test_lu_first_max_tie_rule(orc):
    # Column 0 holds +-2 at rows 3 and 6: the first maximal |x| wins (idamax).
    A = np.asfortranarray(np.eye(8) * 0.25)
    A[3, 0], A[6, 0] = -2.0, 2.0
    _, pv, _ = orc.lu_piv(A)
    assert pv[0] == 3


@pytest.mark.parametrize("n", [64, 333, 1024])
def test_lu_residual_c2(orc, n):
    A = orc.gen_lu_c2(77, n, n)
    out, p, r = orc.lu_piv(A)
    assert r == orc.SUCCESS
    assert orc.lu_residual(A, out, p) <= 10.0
    # The pivots are a valid GEPP sequence: |L| <= 1 below the diagonal.
    assert np.max(np.abs(np.tril(out, -1))) <= 1.0


def test_lu_pivots_match_textbook_gepp(orc):
    """Right-looking textbook GEPP (numpy, same first-max rule) picks the same
    pivots as the blocked Crout restatement on the C2-style matrix."""
    n = 200
    A = orc.gen_lu_c2(3, n, n)
    W = np.array(A, dtype=np.float64)
    piv = []
    for j in range(n):
        r = j + int(np.argmax(np.abs(W[j:, j])))  # argmax = first max
        piv.append(r - j)
        W[[j, r], :] = W[[r, j], :]
        W[j + 1:, j] /= W[j, j]
        W[j + 1:, j + 1:] -= np.outer(W[j + 1:, j], W[j, j + 1:])
    _, p, _ = orc.lu_piv(A)
    assert np.array_equal(p, np.array(piv, dtype=np.int32))


def test_apply_pivots_is_permutation(orc):
    rng = np.random.default_rng(2)
    m, n = 40, 7
    A = rng.standard_normal((m, n))
    p = np.array([rng.integers(0, m - i) for i in range(25)], dtype=np.int32)
    got = orc.apply_pivots_ln(p, A)
    perm = perm_from_pivots(p, m)
    assert np.array_equal(got, A[perm, :])


def test_shift_pivots_roundtrip(orc):
    p = np.array([3, 0, 5, 1, 0], dtype=np.int32)
    q = p.copy()
    orc.lib().or_shift_pivots_to(orc.LAPACK_PIVOTS, len(q), orc.ip(q))
    assert np.array_equal(q, p + np.arange(len(p)) + 1)
    orc.lib().or_shift_pivots_to(orc.NATIVE_PIVOTS, len(q), orc.ip(q))
    assert np.array_equal(q, p)


@pytest.mark.parametrize("upper", [False, True])
def test_chol_solve_exact(orc, upper):
    n, nrhs = 150, 16
    A, F = exact_spd(n, seed=3, upper=upper)
    rng = np.random.default_rng(4)
    X = rng.integers(-3, 4, size=(n, nrhs)).astype(np.float64)
    # B = A X; the solve through the exact factor returns X when every
    # intermediate stays dyadic: use the exact factor F directly.
    B = A @ X
    got = orc.chol_solve(F, B, uplo=orc.UPPER if upper else orc.LOWER)
    assert np.allclose(got, X, rtol=0, atol=1e-9)


def test_lu_solve_backward_error(orc):
    n, nrhs = 300, 16
    A = orc.gen_lu_c2(8, n, n)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((n, nrhs))
    F, p, r = orc.lu_piv(A)
    X = orc.lu_piv_solve(F, p, B)
    res = np.linalg.norm(B - A @ X, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1) * n * 2.0**-53)
    assert res <= 10.0


def test_lapack2flame_info_and_ipiv(orc):
    import ctypes as C
    L = orc.lib()
    n = 200
    A, L0, U0, p = exact_lu(n, seed=11, null_col=57)
    A = np.asfortranarray(A)
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(0)
    L.or_dgetrf(n, n, orc.dp(A), n, orc.ip(ipiv), C.byref(info))
    assert info.value == 58
    assert np.array_equal(ipiv, p + np.arange(n) + 1)
    S, F = exact_spd(n, seed=12)
    S[90, 90] = -1.0
    S = np.asfortranarray(S)
    L.or_dpotrf(b"L", n, orc.dp(S), n, C.byref(info))
    assert info.value == 91
    L.or_dpotrf(b"X", n, orc.dp(S), n, C.byref(info))
    assert info.value == -1
    L.or_dpotrf(b"L", n, orc.dp(S), n - 1, C.byref(info))
    assert info.value == -4


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgetrs_matches_lu_solve(orc, trans):
    import ctypes as C
    n, nrhs = 120, 5
    A = orc.gen_lu_c2(21, n, n)
    F = np.asfortranarray(A.copy())
    ipiv = np.zeros(n, dtype=np.int32)
    info = C.c_int(0)
    orc.lib().or_dgetrf(n, n, orc.dp(F), n, orc.ip(ipiv), C.byref(info))
    rng = np.random.default_rng(1)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    X = B.copy(order="F")
    orc.lib().or_dgetrs(trans.encode(), n, nrhs, orc.dp(F), n, orc.ip(ipiv), orc.dp(X), n,
                        C.byref(info))
    opA = A if trans == "N" else A.T
    res = np.linalg.norm(B - opA @ X, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X, 1) * n * 2.0**-53)
    assert info.value == 0 and res <= 10.0


def test_generators_deterministic_and_distributed(orc):
    u = orc.gen_uniform(5, 400, 300)
    assert np.array_equal(u, orc.gen_uniform(5, 400, 300))
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01
    g = orc.gen_normal(5, 400, 300)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.02
    assert not np.array_equal(orc.gen_uniform(6, 10, 10), orc.gen_uniform(5, 10, 10))


# ---- pins of the oracle's error metrics and of or_dpotrs --------------------
# The residual ratios are the dpot01 / dget01 forms of the LAPACK test suite
# (netlib-test/lapack-3.4.2/TESTING/LIN/dpot01.f:88,151-156,213-217 and
# dget01.f:154-214): ||L L' - A||_1 / (N eps ||A||_1) with eps = 2^-53.  They
# are pinned against an independent numpy evaluation on factors carrying a
# known dyadic perturbation, where every product and sum is exact in binary64,
# so the two must agree to the rounding of the final division.  A dropped N,
# eps = 2^-52, the infinity norm, a one-sided triangle of the difference or a
# read of the unreferenced triangle changes the value.

_EPS = 2.0 ** -53


def _norm1(M):
    return float(np.abs(M).sum(axis=0).max())


def _perturb(F, seed, tri, count=7, scale=2.0 ** -16):
    rng = np.random.default_rng(seed)
    G = F.copy(order="F")
    n0, n1 = G.shape
    placed = 0
    while placed < count:
        i, j = int(rng.integers(0, n0)), int(rng.integers(0, n1))
        if tri == "lower" and i < j or tri == "upper" and i > j:
            continue
        G[i, j] += scale * int(rng.integers(1, 8)) * (1 if rng.random() < 0.5 else -1)
        placed += 1
    return G


@pytest.mark.parametrize("n", [5, 33, 130])
@pytest.mark.parametrize("upper", [False, True])
def test_chol_residual_metric_pinned(orc, n, upper):
    A, F = exact_spd(n, seed=100 + n, upper=upper)
    uplo = orc.UPPER if upper else orc.LOWER
    # the exact factor: the difference is exactly zero
    assert orc.chol_residual(A, F, uplo=uplo) == 0.0
    G = _perturb(F, seed=n, tri="upper" if upper else "lower")
    R = (G.T @ G if upper else G @ G.T) - A  # exact: integers plus multiples of 2^-32
    want = _norm1(R) / (n * _EPS * _norm1(A))
    # the unreferenced triangles of A0 and of the factor must not be read
    A_dirty, G_dirty = A.copy(order="F"), G.copy(order="F")
    if upper:
        A_dirty[np.tril_indices(n, -1)] = np.nan
        G_dirty[np.tril_indices(n, -1)] = np.nan
    else:
        A_dirty[np.triu_indices(n, 1)] = np.nan
        G_dirty[np.triu_indices(n, 1)] = np.nan
    got = orc.chol_residual(A_dirty, G_dirty, uplo=uplo)
    assert want > 0 and abs(got - want) <= 1e-12 * want


def test_chol_residual_zero_matrix(orc):
    # dpot01.f:151-156: 1/eps when ||A|| = 0
    Z = np.zeros((4, 4), order="F")
    assert orc.chol_residual(Z, Z) == 1.0 / _EPS


@pytest.mark.parametrize("shape", [(6, 6), (70, 70), (90, 40), (40, 90)])
def test_lu_residual_metric_pinned(orc, shape):
    m, n = shape
    k = min(m, n)
    A, L, U, p = exact_lu(n, seed=m * 7 + n, m=m)
    F = np.zeros((m, n), order="F")
    F[:, :k] += np.tril(L, -1)
    F[:k, :] += U
    assert orc.lu_residual(A, F, p) == 0.0
    G = _perturb(F, seed=m + n, tri="none")
    L2 = np.tril(G[:, :k], -1)
    L2[:k, :k] += np.eye(k)
    U2 = np.triu(G[:k, :])
    PA = A[perm_from_pivots(p, m), :]
    R = L2 @ U2 - PA
    # dget01 divides by N, the number of columns (dget01.f:209-214)
    want = _norm1(R) / (n * _EPS * _norm1(A))
    got = orc.lu_residual(A, G, p)
    assert want > 0 and abs(got - want) <= 1e-12 * want
    # the infinity norm of the same difference differs (the pin tells them apart)
    assert abs(float(np.abs(R).sum(axis=1).max()) / (n * _EPS * _norm1(A)) - want) > 1e-6 * want


def test_lu_residual_zero_matrix(orc):
    Z = np.zeros((5, 5), order="F")
    assert orc.lu_residual(Z, Z, np.zeros(5, np.int32)) == 0.0


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n", [1, 40, 161])
def test_dpotrs_exact_dyadic(orc, n, upper):
    """or_dpotrs (dpotrs.c:179-189): with the exact dyadic factor F of A and
    integer X, every intermediate of both substitutions is an integer, so
    B = A X is solved back to X bit for bit; solving the two triangles in the
    wrong order or with the wrong transpose does not return X."""
    import ctypes as C
    nrhs = 9
    A, F = exact_spd(n, seed=300 + n, upper=upper)
    rng = np.random.default_rng(n)
    X = np.asfortranarray(rng.integers(-4, 5, size=(n, nrhs)).astype(np.float64))
    B = np.asfortranarray(A @ X)
    Fd = F.copy(order="F")
    # the other triangle holds garbage: dpotrs references only uplo's
    if upper:
        Fd[np.tril_indices(n, -1)] = np.nan
    else:
        Fd[np.triu_indices(n, 1)] = np.nan
    info = C.c_int(7)
    orc.lib().or_dpotrs(b"U" if upper else b"L", n, nrhs, orc.dp(Fd), n, orc.dp(B), n,
                        C.byref(info))
    assert info.value == 0
    assert np.array_equal(B, X)
    if n > 1:
        # The swapped order of the two substitutions (a plausible mistake)
        # solves F'F X = B instead of F F' X = B and does not return X.
        ul = "U" if upper else "L"
        first, second = ("N", "T") if upper else ("T", "N")
        Bw = orc.dtrsm("L", ul, first, "N", F, A @ X)
        Bw = orc.dtrsm("L", ul, second, "N", F, Bw)
        assert not np.array_equal(Bw, X)


def test_dpotrs_argument_checks(orc):
    import ctypes as C
    a = np.eye(4, order="F")
    b = np.ones((4, 2), order="F")
    info = C.c_int(0)
    L = orc.lib()
    for args, want in [((b"X", 4, 2, 4, 4), -1), ((b"L", -1, 2, 4, 4), -2), ((b"L", 4, -1, 4, 4), -3),
                       ((b"L", 4, 2, 3, 4), -5), ((b"L", 4, 2, 4, 3), -7)]:
        u, n, nrhs, lda, ldb = args
        L.or_dpotrs(u, n, nrhs, orc.dp(a), lda, orc.dp(b), ldb, C.byref(info))
        assert info.value == want
    assert np.array_equal(b, np.ones((4, 2)))


def test_openmp_oracle_build_is_bit_identical(orc):
    """oracle/liboracle_omp.so (the multi-core CPU baseline) splits the
    level-3 routines' independent columns / rows over threads without changing
    any element's sequence of operations: its FLA_Chol and FLA_LU_piv equal the
    serial oracle's bit for bit, at sizes where the threaded paths run."""
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                        "liboracle_omp.so")
    L = C.CDLL(path)
    dp_, ip_ = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.or_chol.argtypes = [C.c_int, C.c_int, dp_, C.c_int, ip_, C.c_int]
    L.or_lu_piv.argtypes = [C.c_int, C.c_int, dp_, C.c_int, ip_, ip_, C.c_int]
    L.or_dtrsm.argtypes = [C.c_char, C.c_char, C.c_char, C.c_char, C.c_int, C.c_int, C.c_double,
                           dp_, C.c_int, dp_, C.c_int]
    n = 700
    for uplo in (orc.LOWER, orc.UPPER):
        A = orc.gen_spd_c2(5, n)
        ref, r = orc.chol(A, uplo=uplo)
        got = np.asfortranarray(A.copy())
        nb = np.array(orc.NB_CHOL, np.int32)
        assert L.or_chol(uplo, n, orc.dp(got), n, orc.ip(nb), 2) == r == orc.SUCCESS
        assert got.tobytes() == ref.tobytes()
    A = orc.gen_lu_c2(6, n, 650)
    ref, p, r = orc.lu_piv(A)
    got = np.asfortranarray(A.copy())
    q = np.zeros(650, np.int32)
    nb = np.array(orc.NB_LU, np.int32)
    assert L.or_lu_piv(n, 650, orc.dp(got), n, orc.ip(q), orc.ip(nb), 2) == r
    assert got.tobytes() == ref.tobytes() and np.array_equal(p, q)
    rng = np.random.default_rng(1)
    T = np.asfortranarray(np.tril(rng.uniform(-1, 1, (300, 300))) + 4 * np.eye(300))
    for side, dims in (("L", (300, 90)), ("R", (500, 300))):
        for tr in ("N", "T"):
            B = np.asfortranarray(rng.standard_normal(dims))
            want = orc.dtrsm(side, "L", tr, "N", T, B, alpha=0.5)
            got = B.copy(order="F")
            L.or_dtrsm(side.encode(), b"L", tr.encode(), b"N", dims[0], dims[1], 0.5, orc.dp(T),
                       300, orc.dp(got), dims[0])
            assert got.tobytes() == want.tobytes(), (side, tr)
