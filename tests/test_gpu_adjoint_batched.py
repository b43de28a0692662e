"""Batched FP32 Cholesky adjoint (fla_hip_schol_adjoint_batched) and the
batched GP log-marginal-likelihood gradient (fla_hip_sgp_rbf_lml_grad_batched).

Pins: (1) per problem, the FP64 CPU reference adjoint assembled from the
oracle's dgemm_ / dtrsm_ restatements, on FP32-rounded factors, to FP32
accuracy; (2) for the GP, the closed form Kbar = (alpha alpha^T - K^-1)/2 and
central differences (h = 1e-6) of the FP64 log marginal likelihood in (l, s)
at the same FP32-rounded points, to the accuracy FP32 arithmetic allows at
the problem's condition number; (3) a problem whose K is not positive
definite reports its failing column and NaN outputs without disturbing the
others."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla

pytestmark = pytest.mark.gpu


def phi(M):
    P = np.tril(M)
    P[np.diag_indices_from(P)] *= 0.5
    return P


def ref_adjoint(orc, L, Lbar):
    n = L.shape[0]
    Lc, Lb = np.tril(L), np.tril(Lbar)
    M = orc.dgemm("T", "N", Lc, Lb, np.zeros((n, n)), alpha=1.0, beta=0.0)
    Y = orc.dtrsm("L", "L", "T", "N", Lc, phi(M))
    G = orc.dtrsm("R", "L", "N", "N", Lc, Y)
    return 0.5 * (G + G.T)


def dev_f32(a):
    return fla.DeviceMatrix.from_host(np.asfortranarray(a, dtype=np.float32), np.float32)


@pytest.mark.parametrize("n", [1, 37, 100, 256])
def test_schol_adjoint_batched_vs_fp64_reference(orc, n):
    count = 6
    rng = np.random.default_rng(n)
    Ls, Lbs = [], []
    for i in range(count):
        A = orc.gen_spd_c2(40 + i, n)  # well conditioned: B B^T / n + I
        L = np.tril(np.linalg.cholesky(A)).astype(np.float32)
        Lb = np.tril(rng.standard_normal((n, n))).astype(np.float32)
        # garbage above the diagonal must not be read
        Ls.append(L + np.triu(np.full((n, n), 9.0, np.float32), 1))
        Lbs.append(Lb + np.triu(np.full((n, n), -5.0, np.float32), 1))
    Lall = np.concatenate([np.asfortranarray(m).ravel(order="F") for m in Ls])
    Lball = np.concatenate([np.asfortranarray(m).ravel(order="F") for m in Lbs])
    dL, dLb = dev_f32(Lall.reshape(-1, 1)), dev_f32(Lball.reshape(-1, 1))
    dA = fla.DeviceMatrix(count * n * n, 1, np.float32)
    fla.check(fla.lib().fla_hip_schol_adjoint_batched(count, n, dL.ptr, dLb.ptr, dA.ptr))
    fla.synchronize()
    got = dA.download().ravel()
    for i in range(count):
        Ai = got[i * n * n:(i + 1) * n * n].reshape((n, n), order="F")
        want = ref_adjoint(orc, np.tril(Ls[i]).astype(np.float64), np.tril(Lbs[i]).astype(np.float64))
        scale = max(1.0, np.max(np.abs(want)))
        assert np.max(np.abs(Ai - want)) <= 2e-4 * scale, (i, np.max(np.abs(Ai - want)), scale)
        assert np.array_equal(Ai, Ai.T)


def gp_lml64(x, y, l, s, noise):
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    K = s * s * np.exp(-d2 / (2 * l * l)) + noise * np.eye(len(y))
    Lc = np.linalg.cholesky(K)
    alpha = np.linalg.solve(K, y)
    return -0.5 * y @ alpha - np.sum(np.log(np.diag(Lc))) - 0.5 * len(y) * np.log(2 * np.pi), K, d2


def run_gp(xs, ys, ls, s, noise):
    count, n, d = len(xs), xs[0].shape[0], xs[0].shape[1]
    X = np.concatenate([np.asfortranarray(x, dtype=np.float32).ravel(order="F") for x in xs])
    Y = np.concatenate([np.asarray(y, dtype=np.float32) for y in ys])
    dX, dY = dev_f32(X.reshape(-1, 1)), dev_f32(Y.reshape(-1, 1))
    dl = dev_f32(np.asarray(ls, dtype=np.float32).reshape(-1, 1))
    dlml = fla.DeviceMatrix(count, 1, np.float32)
    dg = fla.DeviceMatrix(2 * count, 1, np.float32)
    dinfo = fla.DeviceMatrix(count, 1, np.int32)
    fla.check(fla.lib().fla_hip_sgp_rbf_lml_grad_batched(count, n, d, dX.ptr, dY.ptr, dl.ptr, s,
                                                         noise, dlml.ptr, dg.ptr, dinfo.ptr))
    fla.synchronize()
    return (dlml.download().ravel().copy(), dg.download().ravel().reshape(count, 2).copy(),
            dinfo.download().ravel().copy())


def test_gp_batched_gradient(orc):
    count, n, s, noise = 24, 256, 1.0, 1e-3
    rng = np.random.default_rng(7)
    xs = [(orc.gen_uniform(100 + i, n, 2) * 0.5 + 0.5).astype(np.float32) for i in range(count)]
    ys = [orc.gen_normal(200 + i, n, 1)[:, 0].astype(np.float32) for i in range(count)]
    ls = rng.uniform(0.05, 0.5, size=count).astype(np.float32)
    lml, g, info = run_gp(xs, ys, ls, s, noise)
    assert np.all(info == fla.FLA_SUCCESS)
    for i in range(count):
        x, y, l = xs[i].astype(np.float64), ys[i].astype(np.float64), float(ls[i])
        ref, K, d2 = gp_lml64(x, y, l, s, noise)
        kappa = np.linalg.cond(K)
        tol = 64 * kappa * 2.0 ** -24  # FP32 factor and adjoint at this condition number
        assert abs(lml[i] - ref) <= tol * max(1.0, abs(ref)), (i, lml[i], ref, kappa)
        alpha = np.linalg.solve(K, y)
        Kbar = 0.5 * (np.outer(alpha, alpha) - np.linalg.inv(K))
        E = np.exp(-d2 / (2 * l * l))
        gl = np.sum(Kbar * s * s * E * d2 / l ** 3)
        gs = np.sum(Kbar * 2 * s * E)
        gscale = np.sum(np.abs(Kbar) * s * s * E * d2 / l ** 3) + abs(gl)
        sscale = np.sum(np.abs(Kbar) * 2 * s * E) + abs(gs)
        assert abs(g[i, 0] - gl) <= tol * gscale, (i, g[i, 0], gl, kappa)
        assert abs(g[i, 1] - gs) <= tol * sscale, (i, g[i, 1], gs, kappa)
        if i < 3:  # central differences of the FP64 likelihood
            h = 1e-6
            fdl = (gp_lml64(x, y, l + h, s, noise)[0] - gp_lml64(x, y, l - h, s, noise)[0]) / (2 * h)
            assert abs(fdl - gl) <= 1e-4 * max(1.0, abs(gl))


def test_gp_batched_failure_is_per_problem(orc):
    """A NaN point makes K_1 fail the Cholesky (NaN pivot) at or before its
    column; its outputs are NaN and the other problems are unaffected."""
    count, n = 3, 64
    xs = [(orc.gen_uniform(300 + i, n, 2)).astype(np.float32) for i in range(count)]
    ys = [orc.gen_normal(400 + i, n, 1)[:, 0].astype(np.float32) for i in range(count)]
    xs[1][5, 0] = np.nan
    lml, g, info = run_gp(xs, ys, [0.2, 0.2, 0.2], 1.0, 1e-3)
    assert 0 <= info[1] <= 5
    assert np.isnan(lml[1]) and np.all(np.isnan(g[1]))
    ok_lml, ok_g, ok_info = run_gp([xs[0], xs[2]], [ys[0], ys[2]], [0.2, 0.2], 1.0, 1e-3)
    assert info[0] == info[2] == fla.FLA_SUCCESS
    assert np.array_equal(lml[[0, 2]], ok_lml) and np.array_equal(g[[0, 2]], ok_g)
