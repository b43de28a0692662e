"""Pins of the oracle's LU with incremental pivoting (FLASH_LU_incpiv,
src/lapack/dec/lu/incpiv/zaux/flamec/flash/FLASH_LU_incpiv_var2.c and the
FLA_SA_* leaves) and its forward substitution (FLASH_FS_incpiv).

The reference ships no fixtures for this path (its own test,
src/lapack/dec/lu/incpiv/front/flamec/test/flash/time_LU_incpiv.c, checks the
residual of a solve), so the restatement is pinned by what the mathematics
fixes:
* one tile (b = n): no SA_LU ever runs, so the result is the tile's
  unblocked Crout LU bit for bit;
* the solve: x = A^-1 r to a backward error of a few eps (a dropped term, a
  wrong swap row or a transposed operand in SA_LU / SA_FS breaks it);
* |det A| = prod |U_jj| (every transformation is a unit-lower or a
  permutation);
* no interchange when the tiles below the diagonal are zero (p = 0 there,
  those tiles stay zero, and the diagonal tiles are their own Crout LUs);
* interchanges at most steps when the diagonal tiles are tiny;
* nb_alg that does not divide b (the reference's own test: b = 128,
  nb_alg = 20);
* m or n not a multiple of b is refused.
"""
import numpy as np
import pytest

from oracle import oracle as orc

EPS = 2.0 ** -53


def _bwd(A, X, R):
    return np.max(np.abs(A @ X - R)) / (np.abs(A).sum(1).max() * np.abs(X).max() * EPS)


def test_single_tile_equals_crout_lu():
    n = 96
    A = orc.gen_lu_c2(3, n, n)
    out, p, L, r = orc.lu_incpiv(A, n, 16)
    ref, pref, rref = orc.lu_piv(A, nb=[])
    assert r == -1
    assert np.array_equal(out, ref)
    assert np.array_equal(p[:, 0], pref)


@pytest.mark.parametrize("n,b,nb_alg", [(256, 64, 16), (384, 128, 20), (200, 40, 40),
                                        (192, 64, 64)])
def test_solve_residual_and_determinant(n, b, nb_alg):
    rng = np.random.default_rng(n + b)
    A = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(n, n)))
    F, p, Lm, r = orc.lu_incpiv(A, b, nb_alg)
    assert r == -1
    R = np.asfortranarray(rng.standard_normal((n, 3)))
    X = orc.lu_incpiv_solve(F, p, Lm, b, nb_alg, R)
    assert _bwd(A, X, R) <= 50.0 * n ** 0.5
    sign, logdet = np.linalg.slogdet(A)
    assert abs(np.sum(np.log(np.abs(np.diag(F)))) - logdet) <= 1e-10 * max(1.0, abs(logdet))
    # several pivots really moved rows between tiles
    nt = n // b
    below = [p[i * b:(i + 1) * b, k] for k in range(nt) for i in range(k + 1, nt)]
    assert any(np.any(t != 0) for t in below)


def test_block_upper_triangular_never_interchanges():
    n, b = 192, 64
    rng = np.random.default_rng(5)
    A = np.triu(rng.uniform(-1.0, 1.0, size=(n, n)), 0)
    for k in range(n // b):  # full diagonal tiles, zero tiles below them
        A[k * b:(k + 1) * b, k * b:(k + 1) * b] = rng.uniform(-1.0, 1.0, size=(b, b))
    A = np.asfortranarray(A)
    F, p, Lm, r = orc.lu_incpiv(A, b, 16)
    nt = n // b
    for k in range(nt):
        tile = A[k * b:(k + 1) * b, k * b:(k + 1) * b]
        ref, pref, _ = orc.lu_piv(tile, nb=[])
        assert np.array_equal(F[k * b:(k + 1) * b, k * b:(k + 1) * b], ref)
        assert np.array_equal(p[k * b:(k + 1) * b, k], pref)
        for i in range(k + 1, nt):
            assert not np.any(p[i * b:(i + 1) * b, k])
            assert not np.any(F[i * b:(i + 1) * b, k * b:(k + 1) * b])


def test_tiny_diagonal_tile_interchanges():
    n, b = 128, 32
    rng = np.random.default_rng(9)
    A = rng.uniform(-1.0, 1.0, size=(n, n))
    for k in range(n // b):
        A[k * b:(k + 1) * b, k * b:(k + 1) * b] *= 1e-6
    A = np.asfortranarray(A)
    F, p, Lm, r = orc.lu_incpiv(A, b, 8)
    # the first SA_LU of the first step takes most pivots from the tile below
    t = p[b:2 * b, 0]
    assert t[0] != 0 and np.count_nonzero(t) >= b // 2
    R = np.asfortranarray(rng.standard_normal((n, 2)))
    X = orc.lu_incpiv_solve(F, p, Lm, b, 8, R)
    assert _bwd(A, X, R) <= 50.0 * n ** 0.5


def test_ragged_sizes_refused():
    A = np.asfortranarray(np.eye(100))
    assert orc.lu_incpiv(A, 64, 16)[3] == -2
