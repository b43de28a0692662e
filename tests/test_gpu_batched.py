"""Batched factorizations and solves of uneven sizes in one persistent launch
(fla_hip_dposv_batched / fla_hip_sposv_batched / fla_hip_dgesv_batched).

Pins: exactly representable closed forms (Cholesky of B*B^T is B; partial
pivoting recovers the designed P, L, U; integer solutions of integer systems)
must come out bit-exact per matrix in FP64 and FP32; general inputs are
compared with the oracle's FLA_Chol / FLA_LU_piv per matrix (pivots identical
on well-separated pivots) and checked by residuals; a failing matrix reports
its column without disturbing the others; graph replay and tuned
configurations give the same answers as the plain launch."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from cases import exact_lu, exact_spd

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53

SIZES = [1, 2, 31, 32, 33, 64, 100, 127, 128, 129, 200, 255, 256, 300, 384, 511, 512, 17, 5, 96]


def _ptrs(mats):
    arr = (C.c_void_p * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = m.ptr.value if m is not None else None
    return arr


def _ints(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


def run_batch(kind, As, Bs=None, dtype=np.float64):
    """Uploads host matrices, runs the batched entry point, returns
    (factors, pivots, solutions, info)."""
    cnt = len(As)
    dA = [fla.DeviceMatrix.from_host(np.asfortranarray(a), dtype) for a in As]
    n = [a.shape[0] for a in As]
    lda = [d.ld for d in dA]
    dB = [fla.DeviceMatrix.from_host(np.asfortranarray(b), dtype) for b in Bs] if Bs else None
    info = fla.DeviceMatrix(cnt, 1, np.int32)
    L = fla.lib()
    if kind == "lu":
        piv = [fla.DeviceMatrix(max(k, 1), 1, np.int32) for k in n]
        f = L.fla_hip_dgesv_batched
        f.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_int), C.c_void_p,
                      C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p]
        st = f(cnt, _ints(n), _ptrs(dA), _ints(lda), _ptrs(piv),
               _ptrs(dB) if dB else None, _ints([d.ld for d in dB]) if dB else None,
               _ints([b.shape[1] for b in Bs]) if Bs else None, info.ptr)
    else:
        f = L.fla_hip_dposv_batched if dtype == np.float64 else L.fla_hip_sposv_batched
        f.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_int), C.c_void_p,
                      C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p]
        st = f(cnt, _ints(n), _ptrs(dA), _ints(lda), _ptrs(dB) if dB else None,
               _ints([d.ld for d in dB]) if dB else None,
               _ints([b.shape[1] for b in Bs]) if Bs else None, info.ptr)
    fla.check(st, "batched")
    fla.synchronize()
    F = [d.download() for d in dA]
    X = [d.download() for d in dB] if dB else None
    P = [p.download().ravel()[:k].copy() for p, k in zip(piv, n)] if kind == "lu" else None
    return F, P, X, info.download().ravel().copy()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_chol_exact_with_solve(dtype):
    rng = np.random.default_rng(1)
    As, Ls, Bs, Xs = [], [], [], []
    for i, n in enumerate(SIZES):
        A, Bf = exact_spd(n, 100 + i)
        r = int(rng.integers(1, 9))
        X = rng.integers(-2, 3, size=(n, r)).astype(np.float64)
        As.append(A)
        Ls.append(Bf)
        Xs.append(X)
        Bs.append(A @ X)
    F, _, X, info = run_batch("chol", As, Bs, dtype)
    assert np.all(info == fla.FLA_SUCCESS)
    for i in range(len(SIZES)):
        assert np.array_equal(np.tril(F[i]), Ls[i].astype(dtype)), (i, SIZES[i])
        assert np.array_equal(X[i], Xs[i].astype(dtype)), (i, SIZES[i])


@pytest.mark.parametrize("kind", ["chol", "lu"])
def test_batched_solve_many_rhs_exact(kind):
    """More right-hand sides than one wave's 8 per pass: the forward sweep
    spreads column groups over the waves, the backward sweep runs in groups
    of 8; integer solutions of the exact problems come back bit for bit."""
    rng = np.random.default_rng(11)
    sizes, nrhs = [40, 300, 17, 512, 129], [9, 33, 100, 256, 64]
    As, Bs, Xs = [], [], []
    for i, (n, r) in enumerate(zip(sizes, nrhs)):
        A = exact_spd(n, 400 + i)[0] if kind == "chol" else exact_lu(n, 500 + i)[0]
        X = rng.integers(-2, 3, size=(n, r)).astype(np.float64)
        As.append(A)
        Xs.append(X)
        Bs.append(A @ X)
    _, _, X, info = run_batch(kind, As, Bs)
    assert np.all(info == fla.FLA_SUCCESS)
    for i in range(len(sizes)):
        if kind == "chol":
            assert np.array_equal(X[i], Xs[i]), (i, sizes[i], nrhs[i])
        else:  # L has entries k/4: the products are exact, U's solve divides by powers of 2
            assert np.allclose(X[i], Xs[i], rtol=0, atol=1e-9), (i, sizes[i], nrhs[i])


def test_batched_chol_vs_oracle(orc):
    As = [orc.gen_spd_c2(7 + i, n) for i, n in enumerate(SIZES)]
    F, _, _, info = run_batch("chol", As)
    assert np.all(info == fla.FLA_SUCCESS)
    for i, A in enumerate(As):
        ref, r = orc.chol(A)
        Lf = np.tril(F[i])
        n = A.shape[0]
        assert np.max(np.abs(Lf - np.tril(ref))) <= 1e-12 * np.max(np.abs(ref))
        res = np.linalg.norm(A - Lf @ Lf.T, 1) / (n * EPS * np.linalg.norm(A, 1))
        assert res <= 10.0


def test_batched_chol_failure_is_per_matrix():
    As = []
    for i, n in enumerate([50, 300, 64, 200]):
        A, _ = exact_spd(n, i)
        As.append(A)
    As[1][123, 123] = -7.0
    As[3][0, 0] = 0.0
    F, _, _, info = run_batch("chol", As)
    assert list(info) == [fla.FLA_SUCCESS, 123, fla.FLA_SUCCESS, 0]


def test_batched_lu_exact_with_solve():
    rng = np.random.default_rng(2)
    As, Ls, Us, ps, Bs, Xs = [], [], [], [], [], []
    for i, n in enumerate(SIZES):
        A, L, U, p = exact_lu(n, 200 + i)
        r = int(rng.integers(1, 9))
        X = rng.integers(-2, 3, size=(n, r)).astype(np.float64)
        As.append(A)
        Ls.append(L)
        Us.append(U)
        ps.append(p)
        Xs.append(X)
        Bs.append(A @ X)
    F, P, X, info = run_batch("lu", As, Bs)
    assert np.all(info == fla.FLA_SUCCESS)
    for i in range(len(SIZES)):
        assert np.array_equal(P[i], ps[i]), i
        assert np.array_equal(np.tril(F[i], -1) + np.eye(SIZES[i]), Ls[i]), i
        assert np.array_equal(np.triu(F[i]), Us[i]), i
        assert np.allclose(X[i], Xs[i], rtol=0, atol=1e-9 * max(1.0, np.max(np.abs(Xs[i])))), i


def test_batched_lu_vs_oracle(orc):
    As = [orc.gen_lu_c2(3 + i, n, n) for i, n in enumerate(SIZES)]
    F, P, _, info = run_batch("lu", As)
    assert np.all(info == fla.FLA_SUCCESS)
    for i, A in enumerate(As):
        ref, pref, r = orc.lu_piv(A)
        assert np.array_equal(P[i], pref), i
        res = orc.lu_residual(A, F[i], P[i])
        assert res <= 10.0, (i, res)


@pytest.mark.parametrize("col,rows", [(0, [7]), (20, [40, 300]), (20, [511]), (37, [0, 2, 500])])
def test_batched_lu_nan_rule(orc, col, rows):
    """A column holding NaNs takes idamax_'s rule (the scan restarts after
    the last NaN): its pivot and those of the columns before it equal the
    oracle's.  (The NaN then spreads through the trailing matrix.)"""
    n = 512
    A = np.asfortranarray(orc.gen_lu_c2(5, n, n))
    for r in rows:
        A[r, col] = np.nan
    _, P, _, info = run_batch("lu", [A, orc.gen_lu_c2(6, 100, 100)])
    _, pref, _ = orc.lu_piv(A)
    assert np.array_equal(P[0][:col + 1], pref[:col + 1])
    assert info[1] == fla.FLA_SUCCESS


def test_batched_lu_null_pivot():
    A, L, U, p = exact_lu(100, 9, null_col=40)
    F, P, _, info = run_batch("lu", [A, exact_lu(60, 1)[0]])
    assert list(info) == [40, fla.FLA_SUCCESS]
    assert np.array_equal(P[0], p)


def test_batched_graph_replay_and_tuning(orc):
    sizes = [300, 17, 512, 64, 128, 200, 33, 450]
    n_arr, r_arr = _ints(sizes), _ints([3] * len(sizes))
    nb, wg, ms = C.c_int(), C.c_int(), C.c_double()
    fla.check(fla.lib().fla_hip_batched_tune(0, len(sizes), n_arr, r_arr, C.byref(nb),
                                             C.byref(wg), C.byref(ms)))
    assert nb.value in (16, 32) and wg.value in (1, 2, 4) and ms.value > 0
    for rep in range(3):  # the second and third calls replay the cached graph
        As = [orc.gen_spd_c2(50 * rep + i, n) for i, n in enumerate(sizes)]
        Bs = [orc.gen_normal(7 * rep + i, n, 3) for i, n in enumerate(sizes)]
        F, _, X, info = run_batch("chol", As, Bs)
        assert np.all(info == fla.FLA_SUCCESS)
        for i, A in enumerate(As):
            R = Bs[i] - A @ X[i]
            n = A.shape[0]
            err = np.linalg.norm(R, 1) / (np.linalg.norm(A, 1) * np.linalg.norm(X[i], 1) * n * EPS)
            assert err <= 10.0, (rep, i, err)


def test_batched_rejects_oversize():
    A = np.eye(513)
    with pytest.raises(RuntimeError):
        run_batch("chol", [A])
