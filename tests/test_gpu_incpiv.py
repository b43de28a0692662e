"""FLASH_LU_incpiv / FLASH_FS_incpiv / FLASH_LU_incpiv_solve on the GPU,
against the oracle's restatement of the reference's tile algorithm
(src/lapack/dec/lu/incpiv/zaux/flamec/flash/FLASH_LU_incpiv_var2.c and the
FLA_SA_* leaves, pinned in tests/test_oracle_incpiv.py).

Objects come from FLASH_LU_incpiv_create_hier_matrices as in the
reference's own test (time_LU_incpiv.c: b = 128, nb_alg = 20).  Checked:
the pivot tiles (exact), the factored A and the L tiles (to 1e-10 of their
scale), the forward substitution and the solve against the oracle's and by
backward error."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


def _create(A, b, nb_alg):
    n = A.shape[0]
    F = fla.DeviceMatrix.from_host(A)
    Fo = F.obj()  # keep the wrapper alive across the call
    Ao, po, Lo = fla.FLA_Obj(), fla.FLA_Obj(), fla.FLA_Obj()
    bf = (C.c_ulong * 1)(b)
    fla.check_fla(fla.lib().FLASH_LU_incpiv_create_hier_matrices(
        Fo.o, 1, bf, nb_alg, C.byref(Ao), C.byref(po), C.byref(Lo)), "create")
    m = A.shape[0]
    nt = A.shape[1] // b
    return (fla.FlashObj.wrap(Ao, m, A.shape[1]), fla.FlashObj.wrap(po, m, nt, fla.FLA_INT),
            fla.FlashObj.wrap(Lo, m, nt * nb_alg))


def _upper_tiles_zero(M, b, width):
    """M with the tiles (i, k), i < k (or i <= k), of a tile grid set to 0."""
    out = M.copy()
    nt = M.shape[1] // width
    for k in range(nt):
        out[: (k + 1) * b, k * width:(k + 1) * width] = 0
    return out


CASES = [(512, 128, 20, "uniform"), (384, 128, 20, "uniform"), (512, 128, 32, "uniform"),
         (384, 64, 16, "uniform"), (256, 128, 128, "uniform"), (384, 128, 32, "tiny_diag"),
         (256, 256, 64, "uniform")]


@pytest.mark.parametrize("n,b,nb_alg,kind", CASES)
def test_lu_incpiv_vs_oracle_and_solve(orc, n, b, nb_alg, kind):
    print(f"case n={n} b={b} nb_alg={nb_alg} {kind}", flush=True)
    rng = np.random.default_rng(n + b + nb_alg)
    A = rng.uniform(-1.0, 1.0, size=(n, n))
    if kind == "tiny_diag":
        for k in range(n // b):
            A[k * b:(k + 1) * b, k * b:(k + 1) * b] *= 1e-6
    A = np.asfortranarray(A)
    F, pr, Lr, r = orc.lu_incpiv(A, b, nb_alg)
    assert r == -1
    HA, Hp, HL = _create(A, b, nb_alg)
    assert fla.lib().FLASH_LU_incpiv(HA.o, Hp.o, HL.o) == fla.FLA_SUCCESS
    nt = n // b
    pg = Hp.download()
    for k in range(nt):
        for i in range(k, nt):
            assert np.array_equal(pg[i * b:(i + 1) * b, k], pr[i * b:(i + 1) * b, k]), (i, k)
    Fg = HA.download()
    worst = {(i, j): float(np.max(np.abs(Fg[i * b:(i + 1) * b, j * b:(j + 1) * b] -
                                         F[i * b:(i + 1) * b, j * b:(j + 1) * b])))
             for i in range(nt) for j in range(nt)}
    bad = {t: v for t, v in worst.items() if v > 1e-10 * np.max(np.abs(F))}
    if bad:
        (i, j), _ = max(bad.items(), key=lambda kv: kv[1])
        d = np.abs(Fg[i * b:(i + 1) * b, j * b:(j + 1) * b] - F[i * b:(i + 1) * b, j * b:(j + 1) * b])
        rows = np.flatnonzero(d.max(1) > 1e-10)
        cols = np.flatnonzero(d.max(0) > 1e-10)
        print(f"tiles off: {sorted(bad)}; worst tile {(i, j)} rows {rows[:12]} cols {cols[:12]}",
              flush=True)
    assert not bad, sorted(bad)
    Lg = HL.download()
    for k in range(nt):
        for i in range(k + 1, nt):
            t = (slice(i * b, (i + 1) * b), slice(k * nb_alg, (k + 1) * nb_alg))
            assert np.max(np.abs(Lg[t] - Lr[t])) <= 1e-10 * max(1.0, np.max(np.abs(Lr[t])))
    # the solve on hierarchical right-hand sides
    R = np.asfortranarray(rng.standard_normal((n, 3)))
    HB = fla.FlashObj.from_host(R, [b], b_n=[b])
    HX = fla.FlashObj(n, 3, [b], b_n=[b])
    assert fla.lib().FLASH_LU_incpiv_solve(HA.o, Hp.o, HL.o, HB.o, HX.o) == fla.FLA_SUCCESS
    X = HX.download()
    xr = orc.lu_incpiv_solve(F, pr, Lr, b, nb_alg, R)
    assert np.max(np.abs(X - xr)) <= 1e-8 * np.max(np.abs(xr))
    bwd = np.max(np.abs(A @ X - R)) / (np.abs(A).sum(1).max() * np.abs(X).max() * EPS)
    assert bwd <= 50.0 * n ** 0.5


def test_fs_incpiv_single_column_matches_oracle(orc):
    n, b, nb_alg = 384, 128, 20
    rng = np.random.default_rng(1)
    A = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(n, n)))
    HA, Hp, HL = _create(A, b, nb_alg)
    assert fla.lib().FLASH_LU_incpiv(HA.o, Hp.o, HL.o) == fla.FLA_SUCCESS
    r = rng.standard_normal((n, 1))
    Hb = fla.FlashObj.from_host(r, [b], b_n=[1])
    assert fla.lib().FLASH_FS_incpiv(HA.o, Hp.o, HL.o, Hb.o) == fla.FLA_SUCCESS
    y = Hb.download()
    yr = orc.fs_incpiv(HA.download(), Hp.download(), HL.download(), b, nb_alg, r)
    assert np.max(np.abs(y - yr)) <= 1e-10 * np.max(np.abs(yr))


def test_default_inner_block():
    """b_alg = 0 picks max(b / 4, 1) (FLASH_LU_incpiv_determine_alg_blocksize)."""
    A = np.asfortranarray(np.eye(256))
    HA, Hp, HL = _create(A, 128, 0)
    assert fla.lib().FLASH_Obj_scalar_width_tl(HL.o) == 32
    assert fla.lib().FLASH_LU_incpiv(HA.o, Hp.o, HL.o) == fla.FLA_SUCCESS
    assert np.array_equal(HA.download(), A)
