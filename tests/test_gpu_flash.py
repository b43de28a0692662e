"""FLASH (hierarchical storage) objects and front-ends on the GPU.

* Storage: FLASH_Obj_create lays the leaves out in the reference's order
  (FLASH_Obj_create_hierarchy, src/base/flamec/hierarchy/main/FLASH_Obj.c:
  271-404: block-column-major at every level, each leaf column-major with
  leading dimension = its row count, ragged last blocks): the raw device
  allocation is compared with the layout built independently in numpy, for
  depth 1 and 2 and ragged edges.
* Copies: FLASH_Copy_flat_to_hier / hier_to_flat at scalar offsets, and the
  flatten / hierarchify round trip (int and double).
* FLASH_Chol / FLASH_LU_piv / FLASH_LU_nopiv / the solves / FLASH_Apply_pivots
  on hierarchical objects: the exact closed forms bit for bit, the oracle's
  restatement at the FLASH blocking (var3 / var5 over blocks of b with the
  unblocked leaf: or_chol / or_lu_piv with nb = {b}) for pivots and factors,
  and the reference's hierarchical return values: a Cholesky failure at
  scalar column k of a depth-1 object with blocks of b returns k // b + k % b
  (FLA_Chol_l_blk_var3.c:42-43 counts blocks above the failing leaf),
  FLASH_LU_piv the scalar index of the first zero on U's diagonal.
* Objects whose leaves are attached to a flat device matrix
  (FLASH_Obj_attach_buffer): the factorization lands in that matrix.
"""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from tests.cases import exact_lu, exact_spd

pytestmark = pytest.mark.gpu
Lo, Up = fla.FLA_LOWER_TRIANGULAR, fla.FLA_UPPER_TRIANGULAR


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


def _leaf_layout(A, bs):
    """The reference's leaf order for a depth-len(bs) hierarchy of A, as one
    flat array (block sizes bs outermost first, in storage blocks)."""
    A = np.asarray(A)
    elem = [1] * len(bs)
    elem[-1] = bs[-1]
    for i in range(len(bs) - 1, 0, -1):
        elem[i - 1] = elem[i] * bs[i - 1]
    out = []

    def rec(M, level):
        if level == len(bs):
            out.append(np.asfortranarray(M).ravel(order="F"))
            return
        e = elem[level]
        m, n = M.shape
        for j in range(0, n, e):
            for i in range(0, m, e):
                rec(M[i:i + e, j:j + e], level + 1)

    rec(A, 0)
    return np.concatenate(out) if out else np.zeros(0, A.dtype)


def _raw(H, count, dtype=np.float64):
    addr = fla.lib().FLASH_Obj_extract_buffer(H.o)
    out = np.empty(count, dtype)
    fla.check(fla.lib().fla_hip_memcpy_d2h(out.ctypes.data, C.c_void_p(addr), out.nbytes))
    return out


@pytest.mark.parametrize("m,n,bs", [(300, 200, [128]), (256, 256, [64]), (1000, 700, [4, 64]),
                                    (129, 1, [32]), (90, 333, [3, 16])])
def test_storage_order_matches_reference_layout(m, n, bs):
    rng = np.random.default_rng(m * n)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    H = fla.FlashObj.from_host(A, bs)
    assert fla.lib().FLASH_Obj_depth(H.o) == len(bs)
    assert fla.lib().FLASH_Obj_scalar_length(H.o) == m
    assert fla.lib().FLASH_Obj_scalar_width(H.o) == n
    assert np.array_equal(_raw(H, m * n), _leaf_layout(A, bs))
    assert np.array_equal(H.download(), A)


def test_int_vector_roundtrip_and_partial_copies():
    p = np.arange(1000, dtype=np.int32)[::-1].copy()
    P = fla.FlashObj.from_host(p, [128], b_n=[1])
    assert np.array_equal(P.download().ravel(), p)
    assert np.array_equal(_raw(P, 1000, np.int32), p)  # blocks of a column vector stack
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.standard_normal((500, 400)))
    H = fla.FlashObj.from_host(A, [96])
    # FLASH_Copy_hier_to_flat / flat_to_hier on a scalar sub-region at (i, j)
    i, j, mm, nn = 37, 150, 300, 101
    F = fla.DeviceMatrix(mm, nn)
    Fo = F.obj()  # keep the wrapper alive across the call
    fla.check_fla(fla.lib().FLASH_Copy_hier_to_flat(i, j, H.o, Fo.o), "hier_to_flat")
    assert np.array_equal(F.download(), A[i:i + mm, j:j + nn])
    Z = np.asfortranarray(rng.standard_normal((mm, nn)))
    F.upload(Z)
    fla.check_fla(fla.lib().FLASH_Copy_flat_to_hier(Fo.o, i, j, H.o), "flat_to_hier")
    want = A.copy(order="F")
    want[i:i + mm, j:j + nn] = Z
    assert np.array_equal(H.download(), want)


@pytest.mark.parametrize("uplo", [Lo, Up])
@pytest.mark.parametrize("n,bs", [(640, [128]), (700, [192]), (1000, [2, 128])])
def test_flash_chol_exact_and_vs_oracle(orc, uplo, n, bs):
    A, F = exact_spd(n, seed=n + len(bs), upper=(uplo == Up))
    H = fla.FlashObj.from_host(A, bs)
    r = fla.lib().FLASH_Chol(uplo, H.o)
    assert r == fla.FLA_SUCCESS
    out = H.download()
    tri, other, off = (np.triu, np.tril, -1) if uplo == Up else (np.tril, np.triu, 1)
    assert np.array_equal(tri(out), F)
    assert np.array_equal(other(out, off), other(A, off)), "other triangle untouched"
    # random SPD: agreement with the oracle at the FLASH blocking (var3 over b)
    B = orc.gen_spd_c2(5, n)
    # the FLASH blocking: var3 over the blocks of every level, unblocked leaf
    ref, rr = orc.chol(B, uplo=uplo, nb=[int(np.prod(bs[i:])) for i in range(len(bs))])
    H2 = fla.FlashObj.from_host(B, bs)
    assert fla.lib().FLASH_Chol(uplo, H2.o) == rr == fla.FLA_SUCCESS
    got = H2.download()
    assert np.max(np.abs(tri(got) - tri(ref))) <= 1e-12 * np.max(np.abs(tri(ref)))
    assert orc.chol_residual(B, got, uplo=uplo) <= 10.0


@pytest.mark.parametrize("k,b", [(300, 128), (128, 128), (77, 64), (0, 32)])
def test_flash_chol_failure_returns_reference_index(orc, k, b):
    n = 400
    A, F = exact_spd(n, seed=k + b)
    A[k, k] -= F[k, k] ** 2 + 1.0
    _, rflat = orc.chol(A, nb=[b])
    assert rflat == k
    H = fla.FlashObj.from_host(A, [b])
    assert fla.lib().FLASH_Chol(Lo, H.o) == k // b + k % b
    # a flat object behaves as a single leaf
    d = fla.DeviceMatrix.from_host(A)
    do = d.obj()
    assert fla.lib().FLASH_Chol(Lo, do.o) == k


@pytest.mark.parametrize("n,b", [(512, 128), (700, 96)])
def test_flash_lu_piv_pivots_vs_oracle_and_solve(orc, n, b):
    M = orc.gen_lu_c2(11, n, n)
    ref, pref, rref = orc.lu_piv(M, nb=[b])
    H = fla.FlashObj.from_host(M, [b])
    P = fla.FlashObj(n, 1, [b], fla.FLA_INT, b_n=[1])
    r = fla.lib().FLASH_LU_piv(H.o, P.o)
    assert r == rref == fla.FLA_SUCCESS
    piv = P.download().ravel()
    assert np.array_equal(piv, pref)
    out = H.download()
    assert np.max(np.abs(out - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert orc.lu_residual(M, out, piv) <= 10.0
    # FLASH_LU_piv_solve on hierarchical B / X
    rng = np.random.default_rng(n)
    Bh = np.asfortranarray(rng.standard_normal((n, 5)))
    HB = fla.FlashObj.from_host(Bh, [b], b_n=[b])
    HX = fla.FlashObj(n, 5, [b], b_n=[b])
    assert fla.lib().FLASH_LU_piv_solve(H.o, P.o, HB.o, HX.o) == fla.FLA_SUCCESS
    X = HX.download()
    xref = orc.lu_piv_solve(out, piv, Bh)
    assert np.max(np.abs(X - xref)) <= 1e-9 * np.max(np.abs(xref))


def test_flash_lu_piv_exact_and_null_pivot():
    n, b = 300, 64
    A, L, U, p = exact_lu(n, seed=8)
    H = fla.FlashObj.from_host(A, [b])
    Pv = fla.FlashObj(n, 1, [b], fla.FLA_INT, b_n=[1])
    assert fla.lib().FLASH_LU_piv(H.o, Pv.o) == fla.FLA_SUCCESS
    out = H.download()
    assert np.array_equal(Pv.download().ravel(), p)
    assert np.array_equal(np.triu(out), U)
    assert np.array_equal(np.tril(out, -1), L - np.eye(n))
    An, *_ = exact_lu(n, seed=9, null_col=131)
    Hn = fla.FlashObj.from_host(An, [b])
    Pn = fla.FlashObj(n, 1, [b], fla.FLA_INT, b_n=[1])
    assert fla.lib().FLASH_LU_piv(Hn.o, Pn.o) == 131


def test_flash_chol_solve_and_apply_pivots(orc):
    n, b = 600, 128
    B0 = orc.gen_spd_c2(2, n)
    H = fla.FlashObj.from_host(B0, [b])
    assert fla.lib().FLASH_Chol(Lo, H.o) == fla.FLA_SUCCESS
    rng = np.random.default_rng(1)
    R = np.asfortranarray(rng.standard_normal((n, 7)))
    HB = fla.FlashObj.from_host(R, [b], b_n=[b])
    HX = fla.FlashObj(n, 7, [b], b_n=[b])
    assert fla.lib().FLASH_Chol_solve(Lo, H.o, HB.o, HX.o) == fla.FLA_SUCCESS
    X = HX.download()
    assert np.max(np.abs(B0 @ X - R)) / (np.abs(B0).sum(0).max() * np.abs(X).max()) <= 1e-14
    # FLASH_Apply_pivots == FLA_Apply_pivots on the flat copy
    p = rng.integers(0, 40, size=n).astype(np.int32)
    p = np.minimum(p, n - 1 - np.arange(n)).astype(np.int32)
    A = np.asfortranarray(rng.standard_normal((n, 50)))
    HA = fla.FlashObj.from_host(A, [b], b_n=[b])
    HP = fla.FlashObj.from_host(p, [b], b_n=[1])
    assert fla.lib().FLASH_Apply_pivots(fla.FLA_LEFT, fla.FLA_NO_TRANSPOSE, HP.o,
                                        HA.o) == fla.FLA_SUCCESS
    assert np.array_equal(HA.download(), orc.apply_pivots_ln(p, A))


def test_flash_lu_nopiv_exact():
    n, b = 384, 128
    rng = np.random.default_rng(4)
    L = np.tril(rng.integers(-2, 3, size=(n, n)).astype(np.float64) / 4.0, -1) + np.eye(n)
    U = np.triu(rng.integers(-3, 4, size=(n, n)).astype(np.float64))
    U[np.diag_indices(n)] = rng.choice([1.0, 2.0, 4.0, -2.0], size=n)
    A = np.asfortranarray(L @ U)
    H = fla.FlashObj.from_host(A, [b])
    assert fla.lib().FLASH_LU_nopiv(H.o) == fla.FLA_SUCCESS
    out = H.download()
    assert np.array_equal(np.triu(out), U) and np.array_equal(np.tril(out, -1), L - np.eye(n))


def test_attached_leaves_write_the_flat_matrix():
    n, b = 512, 128
    A, F = exact_spd(n, seed=21)
    d = fla.DeviceMatrix.from_host(A)
    H = fla.FlashObj(n, n, [b], buffer=False)
    fla.check_fla(fla.lib().FLASH_Obj_attach_buffer(d.ptr, 1, n, C.byref(H.o)), "attach")
    assert fla.lib().FLASH_Chol(Lo, H.o) == fla.FLA_SUCCESS
    assert np.array_equal(np.tril(d.download()), F)
