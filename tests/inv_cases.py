"""Closed-form cases for the triangular / SPD inverses (exact in FP64).

T = D Lu with Lu unit lower bidiagonal (sub-diagonal -1/2) and D a diagonal of
powers of two: inv(T)(i, j) = 2^-(i-j) / d_j for i >= j.  Every product and
sum on the way to it involves powers of two and exact zeros, so any
evaluation order gives it exactly.  The other triangle (and, for a unit
diagonal, the diagonal) holds garbage that must come back untouched."""
import numpy as np

LOWER, UPPER = 300, 301
UNIT, NONUNIT = 500, 501


def bidiag_case(n, uplo, diag, seed=0):
    rng = np.random.default_rng(seed)
    d = 2.0 ** ((np.arange(n) % 5) - 2) if diag == NONUNIT else np.ones(n)
    T = np.diag(d)
    for i in range(n - 1):
        T[i + 1, i] = -0.5 * d[i + 1]
    i, j = np.indices((n, n))
    X = np.where(i >= j, 0.5 ** np.maximum(i - j, 0) / d[np.minimum(j, n - 1)], 0.0)
    garbage = rng.uniform(-9, 9, size=(n, n))
    A = np.where(i >= j, T, garbage)
    want = np.where(i >= j, X, garbage)
    if diag == UNIT:  # stored diagonal is not part of the matrix
        A[np.diag_indices(n)] = 7.0
        want[np.diag_indices(n)] = 7.0
    if uplo == UPPER:
        A, want = A.T.copy(), want.T.copy()
    return np.asfortranarray(A), np.asfortranarray(want)


def int_triangle(n, uplo, seed=0):
    """Small-integer triangle (exact products) plus garbage in the other one;
    returns (A, exact L'L or U U' on the triangle with the garbage kept)."""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.integers(-3, 4, size=(n, n))).astype(np.int64)
    P = L.T @ L  # exact in int64
    i, j = np.indices((n, n))
    garbage = rng.uniform(-9, 9, size=(n, n))
    A = np.where(i >= j, L.astype(np.float64), garbage)
    want = np.where(i >= j, P.astype(np.float64), garbage)
    if uplo == UPPER:
        A, want = A.T.copy(), want.T.copy()
    return np.asfortranarray(A), np.asfortranarray(want)


def random_triangle(n, uplo, seed=0):
    rng = np.random.default_rng(seed)
    T = np.tril(rng.uniform(-1, 1, size=(n, n)) / np.sqrt(max(n, 1)))
    T[np.diag_indices(n)] = 1.5 + rng.uniform(0, 1, size=n)
    return np.asfortranarray(T if uplo == LOWER else T.T)
