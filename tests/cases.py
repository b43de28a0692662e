"""Problem builders whose exact answers are known in closed form.

Every value below is a dyadic rational with few significant bits, so every
intermediate quantity of a Cholesky or LU factorization of these matrices is
exactly representable in binary64: any correct algorithm, in any operation
order, returns the known factors bit for bit.  A dropped term, a wrong sign,
a transposed operand or a misplaced pivot changes the result.
"""
import numpy as np


def exact_spd(n, seed, upper=False):
    """A = B*B^T with B lower triangular, diag(B) in {1, 2, 4}, off-diagonal
    entries small integers; returns (A, B).  Cholesky(A) == B exactly."""
    rng = np.random.default_rng(seed)
    B = np.tril(rng.integers(-3, 4, size=(n, n))).astype(np.float64)
    np.fill_diagonal(B, rng.choice([1.0, 2.0, 4.0], size=n))
    A = B @ B.T  # integers far below 2^53: exact
    if upper:
        return np.asfortranarray(A), np.asfortranarray(B.T)
    return np.asfortranarray(A), np.asfortranarray(B)


def exact_lu(n, seed, m=None, null_col=None):
    """A = P^{-1} * L * U with unit-lower L (|l_ij| in {0, 1/4, 1/2, 3/4}),
    upper U with integer entries and diagonal in {+-1, +-2, +-4}, and random
    relative pivots p.  Partial pivoting recovers p, L and U exactly.
    If null_col is set, U[null_col, null_col] = 0 (first null pivot there)."""
    m = n if m is None else m
    k = min(m, n)
    rng = np.random.default_rng(seed)
    L = np.tril(rng.integers(-3, 4, size=(m, k)).astype(np.float64) / 4.0, -1)
    for i in range(k):
        L[i, i] = 1.0
    U = np.triu(rng.integers(-4, 5, size=(k, n)).astype(np.float64))
    d = rng.choice([1.0, 2.0, 4.0], size=k) * rng.choice([-1.0, 1.0], size=k)
    for i in range(k):
        U[i, i] = d[i]
    if null_col is not None:
        U[null_col, null_col] = 0.0
        # keep the rest of the null column's reduced entries zero too
        L[null_col + 1:, null_col] = 0.0
    p = np.array([rng.integers(0, m - i) for i in range(k)], dtype=np.int32)
    if null_col is not None:
        p[null_col] = 0  # an all-zero candidate column: idamax picks the first
    A = L @ U
    # A := P^{-1} (L U): undo the interchanges in reverse order
    for i in reversed(range(k)):
        j = i + p[i]
        A[[i, j], :] = A[[j, i], :]
    return np.asfortranarray(A), np.asfortranarray(L), np.asfortranarray(U), p


def exact_lu_nopiv(n, seed, m=None, null_col=None):
    """A = L U with the factors of exact_lu and no interchanges: LU without
    pivoting recovers L and U (exactly under forward substitution; the GPU's
    inverse-based A21 solve rounds).  With null_col, U[null_col, null_col] =
    0: the first exact zero on the diagonal is there."""
    m = n if m is None else m
    k = min(m, n)
    rng = np.random.default_rng(seed)
    L = np.tril(rng.integers(-3, 4, size=(m, k)).astype(np.float64) / 4.0, -1)
    for i in range(k):
        L[i, i] = 1.0
    U = np.triu(rng.integers(-4, 5, size=(k, n)).astype(np.float64))
    d = rng.choice([1.0, 2.0, 4.0], size=k) * rng.choice([-1.0, 1.0], size=k)
    for i in range(k):
        U[i, i] = d[i]
    if null_col is not None:
        U[null_col, null_col] = 0.0
    return np.asfortranarray(L @ U), np.asfortranarray(L), np.asfortranarray(U)


def dominant(n, seed, m=None):
    """Random U(-1,1) matrix with n added to the diagonal: LU without
    pivoting is stable on it."""
    m = n if m is None else m
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, size=(m, n))
    for i in range(min(m, n)):
        A[i, i] += n
    return np.asfortranarray(A)


def perm_from_pivots(p, m):
    """Row permutation (as an index array) produced by the interchanges
    i <-> i + p[i], applied in order i = 0, 1, ...: (P A)[r] = A[perm[r]]."""
    perm = np.arange(m)
    for i, pi in enumerate(p):
        j = i + int(pi)
        perm[[i, j]] = perm[[j, i]]
    return perm


def exact_tri(n, seed, lower=True, unit=False):
    """Triangular matrix with small-integer off-diagonal entries and a
    power-of-two diagonal (or unit), so that solves with integer right-hand
    sides built as T*X stay exact."""
    rng = np.random.default_rng(seed)
    T = rng.integers(-2, 3, size=(n, n)).astype(np.float64)
    T = np.tril(T, -1) if lower else np.triu(T, 1)
    dg = np.ones(n) if unit else rng.choice([1.0, 2.0, 4.0], size=n) * rng.choice([-1.0, 1.0], size=n)
    T += np.diag(dg)
    return np.asfortranarray(T)
