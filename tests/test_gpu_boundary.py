"""The LAPACK boundary and device utilities on the GPU, against the oracle.

* dlaswp_ (host ABI), fla_hip_dlaswp and FLA_Apply_pivots: the device
  interchange kernels resolve each touched row's final source in parallel
  and move only touched rows; pinned bit for bit against the oracle's
  f2c-order dlaswp / FLA_Apply_pivots_ln restatement (itself pinned to the
  reference's dlaswp.c in tests/test_oracle_blas.py), forward and reverse
  order, strided pivots, k1 > 1, more interchanges than one LDS chunk, and a
  target outside the matrix (skipped).
* dposv_ / dgesv_ (dposv.c:120, dgesv.c:113): the factor then the solve,
  info codes, B untouched after a failed factorization; exact on dyadic
  closed forms, oracle agreement on random problems.
* dgemm / dsyrk with alpha == 0: C := beta*C without reading A
  (src/flablas/f2c/dgemm.c:194-214, dsyrk.c quick return), so NaN in A does
  not reach C.
* FLA_Chol at the benchmark's outer blocks (nb = 512 and 1024: 128-column
  sub-blocks, K = nb trailing updates) against the oracle.
* Triangular solves with ill-conditioned diagonal blocks: backward error of
  dgetrs_ / dpotrs_ at the oracle's bound.
* Device norms and copies (dlange / dlansy / dlacpy / dlaset) against numpy.
"""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from tests.cases import exact_lu, exact_spd, perm_from_pivots
from tests.gpu_util import dp, gpu_chol, gpu_gemm, gpu_syrk, ip, to_dev

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
Lo, Up = fla.FLA_LOWER_TRIANGULAR, fla.FLA_UPPER_TRIANGULAR
N, T = fla.FLA_NO_TRANSPOSE, fla.FLA_TRANSPOSE


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


def _i(v):
    return np.array([v], dtype=np.int32)


# ---- row interchanges ------------------------------------------------------

def _oracle_dlaswp(orc, A, k1, k2, ipiv, incx):
    out = np.asfortranarray(A.copy())
    orc.lib().or_dlaswp(out.shape[1], orc.dp(out), out.shape[0], k1, k2,
                        orc.ip(np.ascontiguousarray(ipiv, np.int32)), incx)
    return out


@pytest.mark.parametrize("incx", [1, -1, 2, -2])
@pytest.mark.parametrize("m,ncols,k1,k2", [(50, 7, 1, 50), (300, 33, 5, 260),
                                           (3000, 16, 1, 2500), (40, 1, 40, 40)])
def test_dlaswp_lapack_abi_bit_exact(orc, incx, m, ncols, k1, k2):
    rng = np.random.default_rng(m + ncols + k1 + incx)
    A = np.asfortranarray(rng.standard_normal((m, ncols)))
    npiv = k1 + (k2 - k1) * abs(incx) + 1 if incx > 0 else 1 + (k2 - 1) * abs(incx) + 1
    ipiv = rng.integers(1, m + 1, size=npiv).astype(np.int32)  # any row, either direction
    want = _oracle_dlaswp(orc, A, k1, k2, ipiv, incx)
    got = np.asfortranarray(A.copy())
    fla.lib().dlaswp_(ip(_i(ncols)), dp(got), ip(_i(m)), ip(_i(k1)), ip(_i(k2)), ip(ipiv),
                      ip(_i(incx)))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("incp", [1, -1, 3])
@pytest.mark.parametrize("m,K", [(20000, 20000), (5000, 1500), (777, 1)])
def test_device_dlaswp_vs_oracle(orc, incp, m, K):
    """FLAME-native pivots on the device, more interchanges than one LDS chunk
    of the resolve kernel (1024), targets below k2 and beyond it."""
    rng = np.random.default_rng(K + incp)
    ncols = 16
    A = np.asfortranarray(rng.standard_normal((m, ncols)))
    k1 = 3 if K < m - 3 else 0
    k2 = k1 + K - 1
    # row i's pivot lives at p[i * |incp|] (FLA_Apply_pivots_ln_opt_var1.c:313-330);
    # relative offsets keep i + p in [0, m)
    p = np.zeros(k2 * abs(incp) + 1, dtype=np.int32)
    for i in range(k1, k2 + 1):
        p[i * abs(incp)] = rng.integers(-i, m - i)
    # the oracle's dlaswp takes LAPACK (1-based absolute) pivots indexed like
    # dlaswp.c: build the equivalent sequence for |incx| = 1
    seq = list(range(k1, k2 + 1)) if incp > 0 else list(range(k2, k1 - 1, -1))
    want = np.asfortranarray(A.copy())
    for i in seq:
        t = i + int(p[i * abs(incp)])
        want[[i, t], :] = want[[t, i], :]
    dA = to_dev(A)
    dP = fla.DeviceMatrix.from_host(p.reshape(-1, 1), np.int32)
    fla.check(fla.lib().fla_hip_dlaswp(ncols, dA.ptr, dA.ld, k1, k2, dP.ptr, incp))
    fla.synchronize()
    assert np.array_equal(dA.download(), want)
    if abs(incp) == 1:
        ipiv = np.zeros(k2 + 1, dtype=np.int32)
        for i in range(k1, k2 + 1):
            ipiv[i] = i + p[i] + 1
        ref = _oracle_dlaswp(orc, A, k1 + 1, k2 + 1, ipiv, incp)
        assert np.array_equal(ref, want)


def test_apply_pivots_touched_rows_only_and_vs_oracle(orc):
    m, ncols = 4000, 9
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.standard_normal((m, ncols)))
    p = np.array([rng.integers(0, m - i) for i in range(600)], dtype=np.int32)
    want = orc.apply_pivots_ln(p, A)
    dA = to_dev(A)
    dp_ = fla.DeviceMatrix.from_host(p.reshape(-1, 1), np.int32)
    assert fla.FLA_Apply_pivots(fla.FLA_LEFT, fla.FLA_NO_TRANSPOSE, dp_.obj(), dA.obj()) == \
        fla.FLA_SUCCESS
    got = dA.download()
    assert np.array_equal(got, want)
    assert np.array_equal(got, A[perm_from_pivots(p, m), :])


def test_apply_pivots_skips_target_outside_matrix():
    m, ncols = 100, 4
    A = np.asfortranarray(np.arange(m * ncols, dtype=np.float64).reshape(m, ncols))
    p = np.array([3, 500, 0, 2, -10], dtype=np.int32)  # rows 1 and 4 point outside
    want = A.copy()
    for i, pi in enumerate(p):
        t = i + int(pi)
        if 0 <= t < m:
            want[[i, t], :] = want[[t, i], :]
    dA = to_dev(A)
    dP = fla.DeviceMatrix.from_host(p.reshape(-1, 1), np.int32)
    fla.check(fla.lib().fla_hip_dlaswp_rows(m, ncols, dA.ptr, dA.ld, 0, len(p) - 1, dP.ptr, 1))
    fla.synchronize()
    assert np.array_equal(dA.download(), want)


# ---- dposv_ / dgesv_ ----------------------------------------------------------

@pytest.mark.parametrize("uplo", [b"L", b"U"])
@pytest.mark.parametrize("n", [1, 129, 700])
def test_dposv_exact_and_info(orc, uplo, n):
    A, F = exact_spd(n, seed=50 + n, upper=(uplo == b"U"))
    rng = np.random.default_rng(n)
    nrhs = 5
    X = rng.integers(-3, 4, size=(n, nrhs)).astype(np.float64)
    B = np.asfortranarray(A @ X)
    a, b, info = np.asfortranarray(A.copy()), B.copy(order="F"), _i(7)
    fla.lib().dposv_(C.c_char_p(uplo), ip(_i(n)), ip(_i(nrhs)), dp(a), ip(_i(n)), dp(b),
                     ip(_i(n)), ip(info))
    assert info[0] == 0
    tri = np.triu if uplo == b"U" else np.tril
    assert np.array_equal(tri(a), F)
    # the oracle's dpotrs on the same factor (pinned bit for bit on dyadic data)
    ref, oinfo = B.copy(order="F"), C.c_int()
    orc.lib().or_dpotrs(uplo, n, nrhs, orc.dp(np.asfortranarray(a)), n, orc.dp(ref), n,
                        C.byref(oinfo))
    assert np.array_equal(ref, X)
    assert np.allclose(b, X, rtol=0, atol=1e-9 * max(1.0, np.abs(X).max()))


def test_dposv_not_spd_leaves_b():
    n, k = 300, 211
    A, F = exact_spd(n, seed=4)
    A[k, k] -= F[k, k] ** 2 + 1.0
    B = np.asfortranarray(np.ones((n, 3)))
    a, b, info = np.asfortranarray(A.copy()), B.copy(order="F"), _i(0)
    fla.lib().dposv_(C.c_char_p(b"L"), ip(_i(n)), ip(_i(3)), dp(a), ip(_i(n)), dp(b), ip(_i(n)),
                     ip(info))
    assert info[0] == k + 1
    assert np.array_equal(b, B)
    assert np.array_equal(np.tril(a)[:k, :k], F[:k, :k])


@pytest.mark.parametrize("n", [1, 130, 600])
def test_dgesv_exact_pivots_and_solution(orc, n):
    A, L, U, p = exact_lu(n, seed=60 + n)
    rng = np.random.default_rng(n)
    nrhs = 4
    X = rng.integers(-3, 4, size=(n, nrhs)).astype(np.float64)
    B = np.asfortranarray(A @ X)
    a, b, ipiv, info = np.asfortranarray(A.copy()), B.copy(order="F"), np.zeros(n, np.int32), _i(7)
    fla.lib().dgesv_(ip(_i(n)), ip(_i(nrhs)), dp(a), ip(_i(n)), ip(ipiv), dp(b), ip(_i(n)),
                     ip(info))
    assert info[0] == 0
    assert np.array_equal(ipiv, p + np.arange(n) + 1)
    assert np.array_equal(np.triu(a), U)
    assert np.array_equal(np.tril(a, -1), np.tril(L, -1))
    ref, oinfo = B.copy(order="F"), C.c_int()
    orc.lib().or_dgetrs(b"N", n, nrhs, orc.dp(np.asfortranarray(a)), n, orc.ip(ipiv), orc.dp(ref),
                        n, C.byref(oinfo))
    assert np.allclose(b, ref, rtol=0, atol=1e-9 * max(1.0, np.abs(ref).max()))
    assert np.allclose(b, X, rtol=0, atol=1e-9 * max(1.0, np.abs(X).max()))


def test_dgesv_null_pivot_info_and_random_vs_oracle(orc):
    n, c = 200, 77
    A, L, U, p = exact_lu(n, seed=3, null_col=c)
    B = np.asfortranarray(np.ones((n, 2)))
    a, b, ipiv, info = np.asfortranarray(A.copy()), B.copy(order="F"), np.zeros(n, np.int32), _i(0)
    fla.lib().dgesv_(ip(_i(n)), ip(_i(2)), dp(a), ip(_i(n)), ip(ipiv), dp(b), ip(_i(n)), ip(info))
    assert info[0] == c + 1
    assert np.array_equal(b, B)
    assert np.array_equal(ipiv, p + np.arange(n) + 1)
    # random well-separated-pivot problem: same pivots and solution as the oracle
    n = 900
    A = orc.gen_lu_c2(12, n, n)
    B = np.asfortranarray(np.random.default_rng(1).standard_normal((n, 6)))
    a, b, ipiv = np.asfortranarray(A.copy()), B.copy(order="F"), np.zeros(n, np.int32)
    fla.lib().dgesv_(ip(_i(n)), ip(_i(6)), dp(a), ip(_i(n)), ip(ipiv), dp(b), ip(_i(n)), ip(info))
    ra, rb, ripiv, rinfo = np.asfortranarray(A.copy()), B.copy(order="F"), np.zeros(n, np.int32), \
        C.c_int()
    orc.lib().or_dgetrf(n, n, orc.dp(ra), n, orc.ip(ripiv), C.byref(rinfo))
    orc.lib().or_dgetrs(b"N", n, 6, orc.dp(ra), n, orc.ip(ripiv), orc.dp(rb), n, C.byref(rinfo))
    assert info[0] == 0 and rinfo.value == 0
    assert np.array_equal(ipiv, ripiv)
    assert np.max(np.abs(b - rb)) <= 1e-9 * np.max(np.abs(rb))


# ---- alpha == 0 quick return -----------------------------------------------

@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_dgemm_dsyrk_alpha_zero_do_not_read_operands(orc, beta):
    rng = np.random.default_rng(2)
    m, n, k = 130, 70, 40
    A = rng.standard_normal((m, k))
    A[5, 3] = np.nan
    B = rng.standard_normal((k, n))
    B[0, 0] = np.inf
    C0 = rng.standard_normal((m, n))
    got = gpu_gemm(N, N, A, B, C0, alpha=0.0, beta=beta)
    want = orc.dgemm("N", "N", A, B, C0, alpha=0.0, beta=beta)
    assert np.array_equal(got, want) and np.all(np.isfinite(got))
    assert np.array_equal(got, beta * C0)
    S = rng.standard_normal((m, k))
    S[7, 7] = np.nan
    C1 = rng.standard_normal((m, m))
    for uplo, u in ((Lo, "L"), (Up, "U")):
        got = gpu_syrk(uplo, N, S, C1, alpha=0.0, beta=beta)
        want = orc.dsyrk(u, "N", S, C1, alpha=0.0, beta=beta)
        assert np.array_equal(got, want) and np.all(np.isfinite(got))


def test_dgemm_beta_zero_overwrites_nan_c(orc):
    rng = np.random.default_rng(3)
    A, B = rng.integers(-3, 4, (64, 16)).astype(float), rng.integers(-3, 4, (16, 50)).astype(float)
    C0 = np.full((64, 50), np.nan)
    got = gpu_gemm(N, N, A, B, C0, alpha=1.0, beta=0.0)
    assert np.array_equal(got, A @ B)


# ---- FLA_Chol at the benchmark's outer blocks -----------------------------------

@pytest.mark.parametrize("nb", [512, 1024])
def test_chol_large_outer_block_vs_oracle(orc, nb):
    """The n = 32768 benchmark runs with nb = 1024 (eight 128-column
    sub-blocks per outer panel, K = 1024 trailing updates): here the same
    outer blocks at n = 3000, against the oracle (agreement of the factors to
    a small multiple of eps, equal residual bound) and bit for bit on the
    exact closed form."""
    fla.set_blocksizes(chol_nb=nb)
    try:
        n = 3000
        A = orc.gen_spd_c2(31, n)
        ref, rr = orc.chol(A)
        out, r = gpu_chol(A)
        assert r == rr == fla.FLA_SUCCESS
        Lg, Lr = np.tril(out), np.tril(ref)
        assert np.max(np.abs(Lg - Lr)) <= 1e-12 * np.max(np.abs(Lr))
        assert orc.chol_residual(A, out) <= 10.0
        E, F = exact_spd(n, seed=nb)
        outE, rE = gpu_chol(E)
        assert rE == fla.FLA_SUCCESS and np.array_equal(np.tril(outE), F)
        # a failure inside the third outer block, at a sub-block edge
        k = 2 * nb + 128 if 2 * nb + 128 < n else nb + 256
        E2 = E.copy(order="F")
        E2[k, k] -= F[k, k] ** 2 + 1.0
        _, rk = gpu_chol(E2)
        _, rko = orc.chol(E2)
        assert rk == rko == k
    finally:
        fla.set_blocksizes(chol_nb=-1)


# ---- solves with ill-conditioned diagonal blocks ------------------------------

def _graded_upper(n, seed, decades):
    """U = D (I + E): the diagonal D graded over `decades` orders of magnitude
    inside every 128-block and a small strictly upper E, so each diagonal
    block's condition number is about 10^decades (far above what a solve by
    inverted blocks tolerates) while the solution stays finite (the U of an
    ill-conditioned LU, up to row scaling, looks like this)."""
    rng = np.random.default_rng(seed)
    d = 10.0 ** (-decades * ((np.arange(n) % 128) / 127.0)) * rng.choice([-1.0, 1.0], n)
    E = np.triu(rng.uniform(-1.0, 1.0, (n, n)), 1) * (0.5 / np.sqrt(n))
    return np.asfortranarray(d[:, None] * (np.eye(n) + E))


@pytest.mark.parametrize("decades", [4, 8])
def test_getrs_potrs_backward_error_ill_conditioned_blocks(orc, decades):
    n, nrhs = 1000, 16
    U = _graded_upper(n, 7, decades)
    Lunit = np.tril(np.random.default_rng(8).uniform(-1, 1, (n, n)), -1) / 4
    F = np.asfortranarray(np.triu(U) + Lunit)
    ipiv = (np.arange(n) + 1).astype(np.int32)
    rng = np.random.default_rng(9)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    b, info = B.copy(order="F"), _i(0)
    fla.lib().dgetrs_(C.c_char_p(b"N"), ip(_i(n)), ip(_i(nrhs)), dp(F), ip(_i(n)), ip(ipiv),
                      dp(b), ip(_i(n)), ip(info))
    assert info[0] == 0
    Lm = Lunit + np.eye(n)
    Um = np.triu(U)
    # backward error of each triangular stage, dget02 form, as the oracle's
    # substitution achieves (dtrsm_ order): y = L^-1 b, x = U^-1 y
    y = np.linalg.solve(Lm, B)
    ref = B.copy(order="F")
    orc.lib().or_dgetrs(b"N", n, nrhs, orc.dp(F), n, orc.ip(ipiv), orc.dp(ref), n,
                        C.byref(C.c_int()))

    def bwd(Tm, x, rhs):
        return np.max(np.abs(rhs - Tm @ x)) / (np.abs(Tm).sum(1).max() * np.max(np.abs(x)) * n * EPS)

    assert bwd(Um, ref, y) <= 10.0  # the oracle meets the bound
    assert bwd(Um, b, y) <= 10.0
    # Cholesky solve with a graded factor
    Lc = np.asfortranarray(np.tril(U.T))
    b2 = B.copy(order="F")
    fla.lib().dpotrs_(C.c_char_p(b"L"), ip(_i(n)), ip(_i(nrhs)), dp(Lc), ip(_i(n)), dp(b2),
                      ip(_i(n)), ip(info))
    Am = Lc @ Lc.T
    assert info[0] == 0
    r = np.max(np.abs(B - Am @ b2)) / (np.abs(Lc).sum(1).max() ** 2 * np.max(np.abs(b2)) * n * EPS)
    assert r <= 10.0


# ---- device norms and copies ----------------------------------------------------

def test_device_norms_and_copies():
    rng = np.random.default_rng(4)
    m, n = 301, 190
    A = rng.standard_normal((m, n))
    dA = to_dev(A)
    v = C.c_double()
    L = fla.lib()
    for code, want in ((ord("1"), np.abs(A).sum(0).max()), (ord("I"), np.abs(A).sum(1).max()),
                       (ord("M"), np.abs(A).max()), (ord("F"), np.linalg.norm(A))):
        fla.check(L.fla_hip_dlange(code, m, n, dA.ptr, dA.ld, C.byref(v)))
        assert abs(v.value - want) <= 1e-13 * want, chr(code)
    S = rng.standard_normal((n, n))
    for uplo, tri in ((Lo, np.tril), (Up, np.triu)):
        Sym = tri(S) + tri(S, -1 if tri is np.tril else 1).T
        garbage = S.copy()
        garbage[tri(np.ones((n, n))) == 0] = np.nan  # the other triangle is never read
        dS = to_dev(garbage)
        for code, want in ((ord("1"), np.abs(Sym).sum(0).max()),
                           (ord("I"), np.abs(Sym).sum(1).max()),
                           (ord("M"), np.abs(Sym).max()), (ord("F"), np.linalg.norm(Sym))):
            fla.check(L.fla_hip_dlansy(code, uplo, n, dS.ptr, dS.ld, C.byref(v)))
            assert abs(v.value - want) <= 1e-13 * want, (chr(code), uplo)
    # NaN propagates
    An = A.copy()
    An[3, 4] = np.nan
    dN = to_dev(An)
    fla.check(L.fla_hip_dlange(ord("M"), m, n, dN.ptr, dN.ld, C.byref(v)))
    assert np.isnan(v.value)
    # dlacpy / dlaset
    dB = to_dev(np.full((m, n), 7.0))
    fla.check(L.fla_hip_dlacpy(Lo, m, n, dA.ptr, dA.ld, dB.ptr, dB.ld))
    fla.synchronize()
    want = np.full((m, n), 7.0)
    want[np.tril_indices(m, 0, n)] = A[np.tril_indices(m, 0, n)]
    assert np.array_equal(dB.download(), want)
    fla.check(L.fla_hip_dlaset(Up, m, n, -1.0, 3.0, dB.ptr, dB.ld))
    fla.synchronize()
    want[np.triu_indices(m, 1, n)] = -1.0
    want[np.arange(n), np.arange(n)] = 3.0
    assert np.array_equal(dB.download(), want)
