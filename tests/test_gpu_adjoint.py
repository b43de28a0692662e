"""Adjoints of the Cholesky factorization and of the Cholesky solve
(fla_hip_dchol_adjoint, fla_hip_dchol_solve_adjoint, FLA_Chol_adjoint,
FLA_Chol_solve_adjoint) and the Gaussian-process log-marginal-likelihood
gradient built on them.

Pins: (1) an FP64 CPU reference gradient assembled from the oracle's
restatements of dtrsm_ / dgemm_ (the BLAS behind FLA_Trsm / FLA_Gemm);
(2) central finite differences (h = 1e-6 relative) of scalar functions of
the factor and of the solution along random symmetric directions;
(3) for the GP, the closed form Kbar = (alpha alpha^T - K^-1)/2 and central
differences of the log marginal likelihood in (l, s)."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla
from gpu_util import to_dev

pytestmark = pytest.mark.gpu
L_ = fla.FLA_LOWER_TRIANGULAR


def phi(M):
    P = np.tril(M)
    P[np.diag_indices_from(P)] *= 0.5
    return P


def ref_adjoint(orc, L, Lbar):
    """Abar from the oracle's dgemm_ / dtrsm_ (FP64 CPU)."""
    n = L.shape[0]
    Lc, Lb = np.tril(L), np.tril(Lbar)
    M = orc.dgemm("T", "N", Lc, Lb, np.zeros((n, n)), alpha=1.0, beta=0.0)
    Y = orc.dtrsm("L", "L", "T", "N", Lc, phi(M))
    G = orc.dtrsm("R", "L", "N", "N", Lc, Y)
    return 0.5 * (G + G.T)


def gpu_adjoint(L, Lbar):
    n = L.shape[0]
    dL, dLb, dA = to_dev(L), to_dev(Lbar), fla.DeviceMatrix(n, n)
    fla.check(fla.lib().fla_hip_dchol_adjoint(n, dL.ptr, dL.ld, dLb.ptr, dLb.ld, dA.ptr, dA.ld))
    fla.synchronize()
    return dA.download()


def sym_dir(rng, n):
    E = rng.standard_normal((n, n))
    return (E + E.T) / 2


@pytest.mark.parametrize("n", [1, 37, 200, 513])
def test_chol_adjoint_vs_cpu_reference(orc, n):
    A = orc.gen_spd_c2(4, n)
    L, _ = orc.chol(A)
    L = np.tril(L)
    rng = np.random.default_rng(n)
    Lbar = np.tril(rng.standard_normal((n, n)))
    # garbage in the unreferenced upper parts
    Lg = L + np.triu(np.full((n, n), 7.0), 1)
    Lbg = Lbar + np.triu(np.full((n, n), -3.0), 1)
    got = gpu_adjoint(Lg, Lbg)
    want = ref_adjoint(orc, L, Lbar)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
    assert np.array_equal(got, got.T)


def test_chol_adjoint_finite_differences(orc):
    n = 120
    A = orc.gen_spd_c2(8, n)
    L = np.tril(orc.chol(A)[0])
    rng = np.random.default_rng(1)
    W = np.tril(rng.standard_normal((n, n)))  # f(A) = <W, chol(A)>, Lbar = W
    Abar = gpu_adjoint(L, W)
    for trial in range(3):
        E = sym_dir(rng, n)
        h = 1e-6
        fp = np.sum(W * np.tril(orc.chol(A + h * E)[0]))
        fm = np.sum(W * np.tril(orc.chol(A - h * E)[0]))
        fd = (fp - fm) / (2 * h)
        an = np.sum(Abar * E)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (trial, fd, an)


def test_chol_solve_adjoint_finite_differences(orc):
    n, r = 90, 3
    A = orc.gen_spd_c2(9, n)
    L = np.tril(orc.chol(A)[0])
    rng = np.random.default_rng(2)
    B = rng.standard_normal((n, r))
    V = rng.standard_normal((n, r))  # f = <V, A^-1 B>
    X = np.linalg.solve(A, B)
    dL, dX, dV = to_dev(L), to_dev(X), to_dev(V)
    dBb, dA = fla.DeviceMatrix(n, r), to_dev(np.zeros((n, n)))
    fla.check(fla.lib().fla_hip_dchol_solve_adjoint(n, r, dL.ptr, dL.ld, dX.ptr, dX.ld, dV.ptr,
                                                    dV.ld, dBb.ptr, dBb.ld, dA.ptr, dA.ld))
    fla.synchronize()
    Abar, Bbar = dA.download(), dBb.download()
    assert np.allclose(Bbar, np.linalg.solve(A, V), rtol=0, atol=1e-12 * np.max(np.abs(Bbar)))
    for trial in range(3):
        E = sym_dir(rng, n)
        h = 1e-6
        fd = (np.sum(V * np.linalg.solve(A + h * E, B)) -
              np.sum(V * np.linalg.solve(A - h * E, B))) / (2 * h)
        an = np.sum(Abar * E)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (trial, fd, an)


def test_flame_api_adjoint(orc):
    n = 64
    A = orc.gen_spd_c2(3, n)
    L = np.tril(orc.chol(A)[0])
    Lbar = np.tril(np.random.default_rng(5).standard_normal((n, n)))
    dL, dLb, dA = to_dev(L), to_dev(Lbar), fla.DeviceMatrix(n, n)
    assert fla.FLA_Chol_adjoint(L_, dL.obj(), dLb.obj(), dA.obj()) == fla.FLA_SUCCESS
    want = ref_adjoint(orc, L, Lbar)
    assert np.max(np.abs(dA.download() - want)) <= 1e-12 * np.max(np.abs(want))


def gp_lml_numpy(x, y, l, s, noise):
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    K = s * s * np.exp(-d2 / (2 * l * l)) + noise * np.eye(len(y))
    Lc = np.linalg.cholesky(K)
    alpha = np.linalg.solve(K, y)
    return -0.5 * y @ alpha - np.sum(np.log(np.diag(Lc))) - 0.5 * len(y) * np.log(2 * np.pi), K, d2


def test_gp_lml_gradient(orc):
    n = 300
    x = np.asfortranarray(orc.gen_uniform(21, n, 2) * 0.5 + 0.5)
    y = orc.gen_normal(22, n, 1)[:, 0]
    l, s, noise = 0.2, 1.0, 1e-3
    dx, dy = to_dev(x), to_dev(y.reshape(-1, 1))
    lml, g, info = C.c_double(), (C.c_double * 2)(), C.c_int()
    fla.check(fla.lib().fla_hip_gp_rbf_lml_grad(n, 2, dx.ptr, dx.ld, dy.ptr, l, s, noise,
                                                C.byref(lml), g, C.byref(info)))
    assert info.value == fla.FLA_SUCCESS
    ref, K, d2 = gp_lml_numpy(x, y, l, s, noise)
    assert abs(lml.value - ref) <= 1e-9 * abs(ref)
    # closed form gradient
    alpha = np.linalg.solve(K, y)
    Kbar = 0.5 * (np.outer(alpha, alpha) - np.linalg.inv(K))
    E = np.exp(-d2 / (2 * l * l))
    gl = np.sum(Kbar * s * s * E * d2 / l ** 3)
    gs = np.sum(Kbar * 2 * s * E)
    assert abs(g[0] - gl) <= 1e-7 * max(1.0, abs(gl)), (g[0], gl)
    assert abs(g[1] - gs) <= 1e-7 * max(1.0, abs(gs)), (g[1], gs)
    # central differences
    h = 1e-6
    fdl = (gp_lml_numpy(x, y, l + h, s, noise)[0] - gp_lml_numpy(x, y, l - h, s, noise)[0]) / (2 * h)
    fds = (gp_lml_numpy(x, y, l, s + h, noise)[0] - gp_lml_numpy(x, y, l, s - h, noise)[0]) / (2 * h)
    assert abs(g[0] - fdl) <= 1e-4 * max(1.0, abs(fdl))
    assert abs(g[1] - fds) <= 1e-4 * max(1.0, abs(fds))
