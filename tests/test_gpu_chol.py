"""FLA_Chol on the GPU against the oracle (parity tests, FP64).

Bar: exact closed forms are bit-for-bit; random SPD problems meet
||A - L L^T||_1 / (n eps ||A||_1) <= 10 (the bound the task states) and the
failure index equals the oracle's."""
import numpy as np
import pytest

import libflame_amd as fla
from tests.cases import exact_spd
from tests.gpu_util import gpu_chol

pytestmark = pytest.mark.gpu

RESID_MAX = 10.0


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


@pytest.mark.parametrize("n", [1, 2, 3, 31, 127, 128, 129, 255, 256, 257, 300, 513, 1000])
@pytest.mark.parametrize("upper", [False, True])
def test_chol_exact_closed_form(n, upper):
    A, F = exact_spd(n, seed=n + 1, upper=upper)
    out, r = gpu_chol(A, fla.FLA_UPPER_TRIANGULAR if upper else fla.FLA_LOWER_TRIANGULAR)
    assert r == fla.FLA_SUCCESS
    tri = np.triu if upper else np.tril
    assert np.array_equal(tri(out), F)
    other, off = (np.tril, -1) if upper else (np.triu, 1)
    assert np.array_equal(other(out, off), other(A, off)), "other triangle must be untouched"


@pytest.mark.parametrize("nb", [128, 256, 384])
def test_chol_c1_residual_vs_oracle(orc, nb):
    """C1: n=1024, A = B B^T + n I with B U[-1,1] (same bytes as the oracle)."""
    fla.set_blocksizes(chol_nb=nb)
    try:
        A = orc.gen_spd_c1(2024, 1024)
        ref, rr = orc.chol(A)
        out, r = gpu_chol(A)
        assert r == rr == fla.FLA_SUCCESS
        res = orc.chol_residual(A, out)
        assert res <= RESID_MAX, res
        # both are backward stable: the factors agree to a small multiple of eps*cond
        L, R = np.tril(out), np.tril(ref)
        assert np.max(np.abs(L - R)) <= 1e-12 * np.max(np.abs(R))
    finally:
        fla.set_blocksizes(chol_nb=-1)


@pytest.mark.parametrize("n", [700, 2048])
def test_chol_c2_style_residual(orc, n):
    A = orc.gen_spd_c2(7, n)
    out, r = gpu_chol(A)
    assert r == fla.FLA_SUCCESS
    assert orc.chol_residual(A, out) <= RESID_MAX


@pytest.mark.parametrize("k", [0, 5, 127, 128, 129, 255, 256, 400, 599])
@pytest.mark.parametrize("kind", ["negative", "zero", "nan"])
def test_chol_failure_index_matches_oracle(orc, k, kind):
    n = 600
    A, F = exact_spd(n, seed=9)
    if kind == "negative":
        A[k, k] -= F[k, k] ** 2 + 1.0
    elif kind == "zero":
        A[k, k] -= F[k, k] ** 2
    else:
        A[k, k] = np.nan
    _, r_ref = orc.chol(A)
    _, r = gpu_chol(A)
    assert r == r_ref == k


def test_chol_upper_matches_lower_transposed(orc):
    A = orc.gen_spd_c1(5, 777)
    lo, r1 = gpu_chol(A, fla.FLA_LOWER_TRIANGULAR)
    up, r2 = gpu_chol(A, fla.FLA_UPPER_TRIANGULAR)
    assert r1 == r2 == fla.FLA_SUCCESS
    assert np.array_equal(np.triu(up), np.tril(lo).T)


def test_chol_views_and_leading_dimension(orc):
    """A 300x300 SPD block inside a 400x350 buffer (lda = 400), via FLA_Part_2x2."""
    import ctypes as C
    big = np.asfortranarray(np.random.default_rng(0).standard_normal((400, 350)))
    A, F = exact_spd(300, seed=4)
    big[50:350, 20:320] = A
    d = fla.DeviceMatrix.from_host(big)
    whole = d.obj()
    tl, tr_, bl, br = fla.FLA_Obj(), fla.FLA_Obj(), fla.FLA_Obj(), fla.FLA_Obj()
    fla.lib().FLA_Part_2x2(whole.o, C.byref(tl), C.byref(tr_), C.byref(bl), C.byref(br), 50, 20,
                           fla.FLA_TL)
    sub = fla.FLA_Obj()
    a, b, c2 = fla.FLA_Obj(), fla.FLA_Obj(), fla.FLA_Obj()
    fla.lib().FLA_Part_2x2(br, C.byref(sub), C.byref(a), C.byref(b), C.byref(c2), 300, 300,
                           fla.FLA_TL)
    r = fla.lib().FLA_Chol(fla.FLA_LOWER_TRIANGULAR, sub)
    assert r == fla.FLA_SUCCESS
    out = d.download()
    assert np.array_equal(np.tril(out[50:350, 20:320]), F)
    mask = np.ones_like(big, dtype=bool)
    mask[50:350, 20:320] = np.tril(np.ones((300, 300), dtype=bool))
    assert np.array_equal(out[~mask], big[~mask])


def test_chol_host_buffer_staging(orc):
    """An FLA_Obj attached to host memory is staged through the GPU."""
    import ctypes as C
    A, F = exact_spd(200, seed=3)
    host = np.asfortranarray(A.copy())
    o = fla.Obj.attach(host.ctypes.data, 200, 200, fla.FLA_DOUBLE, 200)
    assert fla.lib().FLA_Obj_is_device(o.o) == 0
    assert fla.FLA_Chol(fla.FLA_LOWER_TRIANGULAR, o) == fla.FLA_SUCCESS
    assert np.array_equal(np.tril(host), F)
