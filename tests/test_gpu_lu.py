"""FLA_LU_piv / FLA_Apply_pivots / FLA_Shift_pivots_to on the GPU against the
oracle.  Bar: identical pivot vectors (exact closed forms and the C2-style
well-separated matrices), bit-exact factors where the answer is exact, and
||P A - L U||_1 / (n eps ||A||_1) <= 10 otherwise."""
import numpy as np
import pytest

import libflame_amd as fla
from tests.cases import exact_lu, perm_from_pivots
from tests.gpu_util import gpu_lu, to_dev

pytestmark = pytest.mark.gpu

RESID_MAX = 10.0


@pytest.fixture(scope="module", autouse=True, params=[256, 128, 64])
def _init(request):
    """Every test runs with 256-column blocks of two panels (the default),
    with one 128-column panel per step, and with 64-column panels."""
    fla.FLA_Init()
    fla.set_blocksizes(lu_nb=request.param)
    yield
    fla.set_blocksizes(lu_nb=256)


@pytest.mark.parametrize("shape", [(1, 1), (5, 5), (16, 16), (127, 127), (128, 128), (129, 129),
                                   (256, 256), (257, 257), (300, 300), (600, 600), (700, 300),
                                   (300, 700), (1000, 1000)])
def test_lu_exact_closed_form(shape):
    m, n = shape
    A, L, U, p = exact_lu(n, seed=m * 7 + n, m=m)
    out, pv, r = gpu_lu(A)
    assert r == fla.FLA_SUCCESS
    assert np.array_equal(pv, p)
    k = min(m, n)
    assert np.array_equal(np.tril(out[:, :k], -1), np.tril(L, -1))
    assert np.array_equal(np.triu(out[:k, :]), U)


@pytest.mark.parametrize("c", [0, 3, 127, 128, 250, 299])
def test_lu_null_pivot_index(orc, c):
    A, L, U, p = exact_lu(300, seed=5, null_col=c)
    out, pv, r = gpu_lu(A)
    _, pref, rref = orc.lu_piv(A)
    assert r == rref == c
    assert np.array_equal(pv, pref)
    assert np.array_equal(np.triu(out), U)


@pytest.mark.parametrize("n", [64, 333, 1024, 2048])
def test_lu_c2_pivots_identical_and_residual(orc, n):
    A = orc.gen_lu_c2(77, n, n)
    ref, pref, rref = orc.lu_piv(A)
    out, pv, r = gpu_lu(A)
    assert r == rref == fla.FLA_SUCCESS
    assert np.array_equal(pv, pref), "pivot vectors differ"
    assert orc.lu_residual(A, out, pv) <= RESID_MAX
    assert np.max(np.abs(out - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref)))


def test_lu_first_max_tie_rule(orc):
    A = np.asfortranarray(np.eye(300) * 0.25)
    A[3, 0], A[200, 0] = -2.0, 2.0   # tie across two workgroups' rows: first wins
    A[290, 1] = 2.0
    A[280, 1] = -2.0
    _, pv, _ = gpu_lu(A)
    _, pref, _ = orc.lu_piv(A)
    assert pv[0] == 3
    assert np.array_equal(pv, pref)


@pytest.mark.parametrize("where", ["first", "mid", "last", "two"])
def test_lu_nan_pivot_follows_idamax(orc, where):
    """idamax semantics with NaN: the scan restarts after the last NaN."""
    n = 520
    A = np.asfortranarray(orc.gen_lu_c2(3, n, n))
    col = {"first": [0], "mid": [260], "last": [n - 1], "two": [10, 300]}[where]
    for r in col:
        A[r, 0] = np.nan
    _, pv, _ = gpu_lu(A)
    _, pref, _ = orc.lu_piv(A)
    assert pv[0] == pref[0]


def test_apply_pivots_and_shift(orc):
    rng = np.random.default_rng(2)
    m, ncols = 700, 9
    X = rng.standard_normal((m, ncols))
    p = np.array([rng.integers(0, m - i) for i in range(m)], dtype=np.int32)
    dX = to_dev(X)
    dp = fla.DeviceMatrix.from_host(p.reshape(-1, 1), np.int32)
    assert fla.FLA_Apply_pivots(fla.FLA_LEFT, fla.FLA_NO_TRANSPOSE, dp.obj(), dX.obj()) == \
        fla.FLA_SUCCESS
    assert np.array_equal(dX.download(), X[perm_from_pivots(p, m), :])
    assert np.array_equal(dX.download(), orc.apply_pivots_ln(p, X))
    assert fla.FLA_Shift_pivots_to(fla.FLA_LAPACK_PIVOTS, dp.obj()) == fla.FLA_SUCCESS
    assert np.array_equal(dp.download().ravel(), p + np.arange(m) + 1)
    assert fla.FLA_Shift_pivots_to(fla.FLA_NATIVE_PIVOTS, dp.obj()) == fla.FLA_SUCCESS
    assert np.array_equal(dp.download().ravel(), p)


@pytest.mark.parametrize("m", [16641, 33000])
def test_lu_tall_panel_many_workgroups(orc, m):
    """A panel taller than 64 x 256 rows: more than 64 workgroups, so every
    header poll and the pivot selection run over several rounds of the
    wave's lanes.  Exact closed form (pivots, L, U)."""
    n = 192
    A, L, U, p = exact_lu(n, seed=m, m=m)
    out, pv, r = gpu_lu(A)
    assert r == fla.FLA_SUCCESS
    assert np.array_equal(pv, p)
    assert np.array_equal(np.tril(out[:, :n], -1), np.tril(L, -1))
    assert np.array_equal(np.triu(out[:n, :]), U)


@pytest.mark.parametrize("rows", [[16700], [5, 16900], [16999]])
def test_lu_tall_panel_nan_rule(orc, rows):
    """The NaN restart rule across more than 64 workgroups: pivots of the
    first columns identical to the oracle's."""
    m, n = 17000, 130
    A = np.asfortranarray(orc.gen_lu_c2(9, m, n))
    for r in rows:
        A[r, 0] = np.nan
    _, pv, _ = gpu_lu(A)
    _, pref, _ = orc.lu_piv(A)
    assert pv[0] == pref[0]
    assert pv[0] >= max(rows)  # the scan restarts after the last NaN
