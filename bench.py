#!/usr/bin/env python3
"""Benchmark: FP64 blocked Cholesky (FLA_Chol, lower) at n = 32768 on one MI355X.

Workload (the headline configuration): A = B*B'/n + I with B i.i.d. N(0,1),
n = 32768, generated on the GPU; one step = one FLA_Chol(FLA_LOWER_TRIANGULAR, A)
on a fresh copy of A already resident in HBM (the copy is outside the timed
region).  value = n^3/3 / seconds (libflame's flop count, test/src/test_chol.c:233).
Beside it, in the same run: FLA_LU_piv at n = 16384 (2n^3/3 flops) on the
well-separated-pivot matrix, dpotrs/dgetrs-style solves with 16 right-hand
sides, a randomized residual check of the factor, the trailing-update kernel's
roofline from HIP events, and the CPU oracle timed on a bounded sample.

  python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import ctypes as C
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
EPS = 2.0 ** -53
# Mixed-precision solves stop at a normwise backward error of 16 eps (the
# target), tighter than dsposv_'s sqrt(n) eps = 128 eps at n = 16384.
STOP_FACTOR = 16.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_spd(fla, n, seed):
    A = fla.DeviceMatrix(n, n)
    B = fla.DeviceMatrix(n, n)
    fla.check(fla.lib().fla_hip_gen_normal(seed, n, n, B.ptr, n), "gen_normal")
    fla.check(fla.lib().fla_hip_gram_lower(n, n, B.ptr, n, 1.0 / n, 1.0, A.ptr, n), "gram")
    fla.synchronize()
    B.buf.free()
    return A


def make_lu(fla, n, seed):
    A = fla.DeviceMatrix(n, n)
    fla.check(fla.lib().fla_hip_gen_uniform(seed, n, n, A.ptr, n), "gen_uniform")
    fla.check(fla.lib().fla_hip_scale_rows_logspace(n, n, A.ptr, n), "scale_rows")
    fla.synchronize()
    return A


def profiled_traffic():
    """HBM bytes per trailing-update launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (None if absent)."""
    path = os.path.join(ROOT, "profiles", "chol_n32768_update_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def copy(fla, dst, src):
    fla.check(fla.lib().fla_hip_memcpy_d2d(dst.ptr, src.ptr, src.ld * src.n * 8), "copy")


def chol_probe_residual(fla, np, A0, L, n, nprobe=4):
    """max_j ||A x_j - L (L' x_j)||_inf / (n eps ||A||_est ||x_j||_inf), x_j N(0,1)."""
    X = fla.DeviceMatrix.from_host(np.random.default_rng(0).standard_normal((n, nprobe)))
    Y1, Y2, Y3 = (fla.DeviceMatrix(n, nprobe) for _ in range(3))
    Lib = fla.lib()
    fla.check(Lib.fla_hip_dmatmul_check(1, fla.FLA_LOWER_TRIANGULAR, fla.FLA_NO_TRANSPOSE,
                                        fla.FLA_NONUNIT_DIAG, n, nprobe, A0.ptr, n, X.ptr, n, 0.0,
                                        Y1.ptr, n))
    fla.check(Lib.fla_hip_dmatmul_check(2, fla.FLA_LOWER_TRIANGULAR, fla.FLA_TRANSPOSE,
                                        fla.FLA_NONUNIT_DIAG, n, nprobe, L.ptr, n, X.ptr, n, 0.0,
                                        Y2.ptr, n))
    fla.check(Lib.fla_hip_dmatmul_check(2, fla.FLA_LOWER_TRIANGULAR, fla.FLA_NO_TRANSPOSE,
                                        fla.FLA_NONUNIT_DIAG, n, nprobe, L.ptr, n, Y2.ptr, n, 0.0,
                                        Y3.ptr, n))
    x, y1, y3 = X.download(), Y1.download(), Y3.download()
    anorm = max(np.max(np.abs(y1[:, j])) / np.max(np.abs(x[:, j])) for j in range(nprobe))
    return float(max(np.max(np.abs(y1[:, j] - y3[:, j])) / (n * EPS * anorm * np.max(np.abs(x[:, j])))
                     for j in range(nprobe)))


def _oracle_omp():
    """The oracle's OpenMP build (oracle/liboracle_omp.so): libflame's control
    trees with dgemm/dsyrk/dtrsm split over threads by independent columns or
    rows, bit-identical to the serial oracle."""
    import ctypes as C
    path = os.path.join(ROOT, "oracle", "liboracle_omp.so")
    L = C.CDLL(path)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.or_threads.restype = C.c_int
    L.or_chol.restype = C.c_int
    L.or_chol.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, C.c_int]
    L.or_lu_piv.restype = C.c_int
    L.or_lu_piv.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int]
    return L


def cpu_baseline(fla, n_chol, n_lu):
    """The oracle's FLA_Chol and FLA_LU_piv (libflame's control trees: var3
    128/32 + opt_var2, var5 128/16 + Crout opt_var4; f2c BLAS loop order) on
    every host core the process may use (the OpenMP build, identical results
    to the serial oracle), on bounded samples of the same matrix families
    generated on the GPU and copied to the host."""
    import ctypes as C

    import numpy as np
    L = _oracle_omp()
    cores = int(L.or_threads())
    dA = make_spd(fla, n_chol, 11)
    A = dA.download()
    A = np.asfortranarray(np.tril(A) + np.tril(A, -1).T)
    del dA
    nb = np.array([128, 32], np.int32)
    t0 = time.perf_counter()
    r = L.or_chol(300, n_chol, A.ctypes.data_as(C.POINTER(C.c_double)), n_chol,
                  nb.ctypes.data_as(C.POINTER(C.c_int)), 2)
    dt = time.perf_counter() - t0
    assert r == -1
    log(f"cpu baseline: FLA_Chol n={n_chol} {dt:.1f} s on {cores} threads")
    del A
    dL = make_lu(fla, n_lu, 7)
    M = np.asfortranarray(dL.download())
    del dL
    p = np.zeros(n_lu, np.int32)
    nbl = np.array([128, 16], np.int32)
    t1 = time.perf_counter()
    r = L.or_lu_piv(n_lu, n_lu, M.ctypes.data_as(C.POINTER(C.c_double)), n_lu,
                    p.ctypes.data_as(C.POINTER(C.c_int)), nbl.ctypes.data_as(C.POINTER(C.c_int)), 2)
    dtl = time.perf_counter() - t1
    assert r == -1
    log(f"cpu baseline: FLA_LU_piv n={n_lu} {dtl:.1f} s on {cores} threads")
    return {"value": round(n_chol ** 3 / 3.0 / dt / 1e12, 6), "unit": "TFLOP/s", "cores": cores,
            "kind": "port",
            "lu_value": round(2.0 * n_lu ** 3 / 3.0 / dtl / 1e12, 6),
            "sample": f"oracle FLA_Chol lower (var3 128/32 + opt_var2, f2c BLAS order) n={n_chol}, "
                      f"A=B*B'/n+I: {dt:.1f} s; oracle FLA_LU_piv (var5 128/16 + Crout "
                      f"opt_var4) n={n_lu}, U[-1,1] rows scaled 10^(-i/n): {dtl:.1f} s (lu_value); "
                      f"OpenMP build on {cores} threads, results identical to the serial oracle"}


def lu_moved_bytes(np, piv, n, W=128):
    """HBM bytes of the LU's row interchanges from its pivot vector: per
    128-column panel, every row the panel's interchanges move is read and
    written once in each of the n - W columns outside the panel."""
    total = 0
    for k in range(0, n, W):
        w = min(W, n - k)
        perm = {}
        for j in range(k, k + w):
            t = j + int(piv[j])
            a, b = perm.get(j, j), perm.get(t, t)
            perm[j], perm[t] = b, a
        moved = sum(1 for r, src in perm.items() if r != src)
        total += moved * (n - w) * 16
    return total


def bench_mixed(fla, np, n, nrhs=8):
    """Mixed-precision solve (fla_hip_dsposv) against the plain FP64 solve
    (FLA_Chol + two triangular solves) on A = Q diag(s) Q^T, Q a product of
    64 Householder reflectors, s log-spaced on [1, kappa]; 8 N(0,1) RHS.
    Time-to-solve of each, normwise backward error max_j ||r_j||/(||A|| ||x_j||)
    (infinity norms) and forward error against the FP64 solution."""
    import ctypes as C
    lib = fla.lib()
    res = {"n": n, "nrhs": nrhs, "cases": []}
    A0 = fla.DeviceMatrix(n, n)
    A = fla.DeviceMatrix(n, n)
    B = fla.DeviceMatrix(n, nrhs)
    X = fla.DeviceMatrix(n, nrhs)
    X64 = fla.DeviceMatrix(n, nrhs)
    R = fla.DeviceMatrix(n, nrhs)
    fla.check(lib.fla_hip_gen_normal(5, n, nrhs, B.ptr, n))
    for kappa in (1e2, 1e4, 1e6):
        log(f"mixed: kappa={kappa:g}, n={n}")
        fla.check(lib.fla_hip_gen_spd_reflectors(11, n, 64, kappa, A0.ptr, n))
        fla.synchronize()
        # plain FP64: FLA_Chol + two triangular solves
        t64 = []
        for rep in range(2):
            copy(fla, A, A0)
            copy(fla, X64, B)
            fla.synchronize()
            t0 = time.perf_counter()
            assert fla.FLA_Chol(fla.FLA_LOWER_TRIANGULAR, A.obj()) == fla.FLA_SUCCESS
            for tr in (fla.FLA_NO_TRANSPOSE, fla.FLA_TRANSPOSE):
                fla.check(lib.fla_hip_dtrsm(fla.FLA_LEFT, fla.FLA_LOWER_TRIANGULAR, tr,
                                            fla.FLA_NONUNIT_DIAG, n, nrhs, 1.0, A.ptr, n,
                                            X64.ptr, n))
            fla.synchronize()
            t64.append(time.perf_counter() - t0)
        ms64 = min(t64) * 1e3
        anorm = _sym_inf_norm(fla, np, A0, n)
        bwd64 = backward_error(fla, np, A0, B, X64, R, n, nrhs, anorm)
        case = {"kappa": kappa, "fp64_ms": round(ms64, 2), "fp64_bwd_err_over_eps": bwd64}
        for name, prec, ref in (("fp32_classic", fla.FLA_MP_FP32, fla.FLA_IR_CLASSIC),
                                ("fp32_gmres", fla.FLA_MP_FP32, fla.FLA_IR_GMRES),
                                ("bf16_gmres", fla.FLA_MP_BF16, fla.FLA_IR_GMRES),
                                ("bf16_classic", fla.FLA_MP_BF16, fla.FLA_IR_CLASSIC)):
            ts = []
            for rep in range(2):
                copy(fla, A, A0)
                fla.synchronize()
                t0 = time.perf_counter()
                it, info, st = fla.dsposv(fla.FLA_LOWER_TRIANGULAR, A, B, X, precision=prec,
                                          refinement=ref, stop_factor=STOP_FACTOR)
                fla.synchronize()
                ts.append(time.perf_counter() - t0)
            bwd = backward_error(fla, np, A0, B, X, R, n, nrhs, anorm)
            xh, x64 = X.download(), X64.download()
            fwd = float(np.max(np.abs(xh - x64)) / np.max(np.abs(x64)))
            case[name] = {"ms": round(min(ts) * 1e3, 2), "iter": it, "info": info,
                          "bwd_err_over_eps": bwd, "fwd_err_vs_fp64": fwd,
                          "factor_ms": round(st["factor_ms"], 2),
                          "refine_ms": round(st["refine_ms"], 2),
                          "prec_solves": int(st["prec_solves"]),
                          "gmres_iters": int(st["gmres_iters"]),
                          "time_ratio_vs_fp64": round(min(ts) * 1e3 / ms64, 3)}
            log(f"  {name}: {case[name]}")
        res["cases"].append(case)
    del A0, A, B, X, X64, R
    return res


def bench_mixed_lu(fla, np, n, nrhs=8):
    """Mixed-precision general solve (fla_hip_dsgesv) against the plain FP64
    solve (FLA_LU_piv, the row interchanges and two triangular solves) on
    A = Q1 diag(s) Q2^T with independent products of 64 reflectors each,
    s log-spaced on [1, kappa]; 8 N(0,1) RHS."""
    lib = fla.lib()
    res = {"n": n, "nrhs": nrhs, "cases": []}
    A0 = fla.DeviceMatrix(n, n)
    A = fla.DeviceMatrix(n, n)
    B = fla.DeviceMatrix(n, nrhs)
    X = fla.DeviceMatrix(n, nrhs)
    X64 = fla.DeviceMatrix(n, nrhs)
    R = fla.DeviceMatrix(n, nrhs)
    p = fla.DeviceMatrix(n, 1, np.int32)
    fla.check(lib.fla_hip_gen_normal(5, n, nrhs, B.ptr, n))
    for kappa in (1e2, 1e4, 1e6):
        log(f"mixed LU: kappa={kappa:g}, n={n}")
        fla.check(lib.fla_hip_gen_ge_reflectors(11, 1011, n, 64, kappa, A0.ptr, n))
        fla.synchronize()
        t64 = []
        for rep in range(2):
            copy(fla, A, A0)
            copy(fla, X64, B)
            fla.synchronize()
            t0 = time.perf_counter()
            info = C.c_int()
            fla.check(lib.fla_hip_dgetrf(n, n, A.ptr, n, p.ptr, C.byref(info)))
            fla.check(lib.fla_hip_dlaswp(nrhs, X64.ptr, n, 0, n - 1, p.ptr, 1))
            fla.check(lib.fla_hip_dtrsm(fla.FLA_LEFT, fla.FLA_LOWER_TRIANGULAR,
                                        fla.FLA_NO_TRANSPOSE, fla.FLA_UNIT_DIAG, n, nrhs, 1.0,
                                        A.ptr, n, X64.ptr, n))
            fla.check(lib.fla_hip_dtrsm(fla.FLA_LEFT, fla.FLA_UPPER_TRIANGULAR,
                                        fla.FLA_NO_TRANSPOSE, fla.FLA_NONUNIT_DIAG, n, nrhs, 1.0,
                                        A.ptr, n, X64.ptr, n))
            fla.synchronize()
            t64.append(time.perf_counter() - t0)
        ms64 = min(t64) * 1e3
        anorm = _sym_inf_norm(fla, np, A0, n)
        bwd64 = backward_error(fla, np, A0, B, X64, R, n, nrhs, anorm, general=True)
        case = {"kappa": kappa, "fp64_ms": round(ms64, 2), "fp64_bwd_err_over_eps": bwd64}
        for name, prec, ref in (("fp32_classic", fla.FLA_MP_FP32, fla.FLA_IR_CLASSIC),
                                ("fp32_gmres", fla.FLA_MP_FP32, fla.FLA_IR_GMRES),
                                ("bf16_gmres", fla.FLA_MP_BF16, fla.FLA_IR_GMRES)):
            ts = []
            for rep in range(2):
                copy(fla, A, A0)
                fla.synchronize()
                t0 = time.perf_counter()
                it, info, st = fla.dsgesv(A, p, B, X, precision=prec, refinement=ref,
                                          stop_factor=STOP_FACTOR)
                fla.synchronize()
                ts.append(time.perf_counter() - t0)
            bwd = backward_error(fla, np, A0, B, X, R, n, nrhs, anorm, general=True)
            xh, x64 = X.download(), X64.download()
            fwd = float(np.max(np.abs(xh - x64)) / np.max(np.abs(x64)))
            case[name] = {"ms": round(min(ts) * 1e3, 2), "iter": it, "info": info,
                          "bwd_err_over_eps": bwd, "fwd_err_vs_fp64": fwd,
                          "factor_ms": round(st["factor_ms"], 2),
                          "refine_ms": round(st["refine_ms"], 2),
                          "prec_solves": int(st["prec_solves"]),
                          "gmres_iters": int(st["gmres_iters"]),
                          "time_ratio_vs_fp64": round(min(ts) * 1e3 / ms64, 3)}
            log(f"  {name}: {case[name]}")
        res["cases"].append(case)
    del A0, A, B, X, X64, R, p
    return res


def backward_error(fla, np, A0, B, X, R, n, nrhs, anorm, general=False):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf) / eps, with the
    product computed in FP64 on the GPU from the full (symmetric or general) A."""
    fla.check(fla.lib().fla_hip_dmatmul_check(0 if general else 1, fla.FLA_LOWER_TRIANGULAR,
                                               fla.FLA_NO_TRANSPOSE,
                                               fla.FLA_NONUNIT_DIAG, n, nrhs, A0.ptr, n, X.ptr,
                                               n, 0.0, R.ptr, n))
    fla.synchronize()
    AX, Bh, Xh = R.download(), B.download(), X.download()
    e = 0.0
    for j in range(nrhs):
        e = max(e, np.max(np.abs(Bh[:, j] - AX[:, j])) / (anorm * np.max(np.abs(Xh[:, j]))))
    return float(e / EPS)


def _sym_inf_norm(fla, np, A0, n):
    """||A||_inf of the symmetric A from its full storage, block rows on the host."""
    best = 0.0
    step = 2048
    for r0 in range(0, n, step):
        rows = min(step, n - r0)
        blk = np.empty((rows, n), dtype=np.float64, order="F")
        fla.check(fla.lib().fla_hip_memcpy2d(blk.ctypes.data, rows * 8,
                                              A0.ptr.value + r0 * 8, n * 8, rows * 8, n, 1))
        best = max(best, float(np.max(np.sum(np.abs(blk), axis=1))))
    return best


def bench_batched(fla, np, count, seed=2026, tune=True):
    """4096 SPD matrices, n_i ~ U{32..512}, A_i = B_i B_i^T/n_i + 0.1 I (FP64 and
    FP32), and 4096 U[-1,1] matrices for LU; each with 1-8 N(0,1) RHS; one
    persistent launch each (autotuned configuration unless tune=False, graph
    replay)."""
    import ctypes as C
    lib = fla.lib()
    rng = np.random.default_rng(seed)
    n = rng.integers(32, 513, size=count)
    r = rng.integers(1, 9, size=count)
    flops_chol = float(np.sum(n.astype(np.float64) ** 3) / 3 + np.sum(2.0 * n * n * r))
    flops_lu = float(np.sum(2.0 * n.astype(np.float64) ** 3 / 3) + np.sum(2.0 * n * n * r))
    tot = int(np.sum(n * n))
    totb = int(np.sum(n * r))
    offs = np.concatenate([[0], np.cumsum(n * n)[:-1]])
    offb = np.concatenate([[0], np.cumsum(n * r)[:-1]])
    out = {"count": count, "n_range": [32, 512], "sum_n3_over_3": float(np.sum(n ** 3.0) / 3)}
    A0 = fla.DeviceBuffer(tot * 8)
    A = fla.DeviceBuffer(tot * 8)
    B0 = fla.DeviceBuffer(totb * 8)
    Bw = fla.DeviceBuffer(totb * 8)
    info = fla.DeviceBuffer(count * 4)
    piv = fla.DeviceBuffer(int(np.sum(n)) * 4)
    pofs = np.concatenate([[0], np.cumsum(n)[:-1]])
    Bg = fla.DeviceBuffer(512 * 512 * 8)
    log(f"batched: generating {count} SPD matrices")
    for i in range(count):
        ni = int(n[i])
        fla.check(lib.fla_hip_gen_normal(seed + i, ni, ni, Bg.ptr, ni))
        fla.check(lib.fla_hip_gram_lower(ni, ni, Bg.ptr, ni, 1.0 / ni, 0.1,
                                         C.c_void_p(A0.addr + int(offs[i]) * 8), ni))
    fla.check(lib.fla_hip_gen_normal(seed + 99991, totb, 1, B0.ptr, totb))
    fla.synchronize()
    ni = (C.c_int * count)(*[int(v) for v in n])
    ri = (C.c_int * count)(*[int(v) for v in r])

    def ptrs(base, offsets, esz):
        return (C.c_void_p * count)(*[base + int(o) * esz for o in offsets])

    def run(fn, Aptrs, Bptrs, esz, steps=3, lu=False):
        ts = []
        for rep in range(steps + 1):
            lib.fla_hip_memcpy_d2d(A.ptr, A0.ptr, tot * esz)
            lib.fla_hip_memcpy_d2d(Bw.ptr, B0.ptr, totb * esz)
            fla.synchronize()
            t0 = time.perf_counter()
            if lu:
                fla.check(fn(count, ni, Aptrs, ni, ptrs(piv.addr, pofs, 4), Bptrs, ni, ri,
                             info.ptr))
            else:
                fla.check(fn(count, ni, Aptrs, ni, Bptrs, ni, ri, info.ptr))
            fla.synchronize()
            if rep > 0:
                ts.append(time.perf_counter() - t0)
        inf = np.empty(count, dtype=np.int32)
        lib.fla_hip_memcpy_d2h(inf.ctypes.data, info.ptr, count * 4)
        return min(ts), int(np.sum(inf != -1))

    for kind, name in ((0, "dposv"), (1, "sposv"), (2, "dgesv")):
        esz = 4 if kind == 1 else 8
        if kind == 1:  # FP32 copies of the same matrices and right-hand sides
            A0d, B0d = A0, B0
            A0, B0 = fla.DeviceBuffer(tot * 4), fla.DeviceBuffer(totb * 4)
            fla.check(lib.fla_hip_convert_d2s(tot, 1, A0d.ptr, tot, A0.ptr, tot))
            fla.check(lib.fla_hip_convert_d2s(totb, 1, B0d.ptr, totb, B0.ptr, totb))
        if kind == 2:
            A0, B0 = A0d, B0d
            log("batched: generating U[-1,1] matrices for LU")
            fla.check(lib.fla_hip_gen_uniform(seed + 7, tot, 1, A0.ptr, tot))
            fla.check(lib.fla_hip_gen_normal(seed + 99991, totb, 1, B0.ptr, totb))
        nb, wg, tms = C.c_int(), C.c_int(), C.c_double()
        if tune:
            fla.check(lib.fla_hip_batched_tune(kind, count, ni, ri, C.byref(nb), C.byref(wg),
                                               C.byref(tms)))
        fn = {0: lib.fla_hip_dposv_batched, 1: lib.fla_hip_sposv_batched,
              2: lib.fla_hip_dgesv_batched}[kind]
        t, bad = run(fn, ptrs(A.addr, offs, esz), ptrs(Bw.addr, offb, esz), esz, lu=(kind == 2))
        fl = flops_lu if kind == 2 else flops_chol
        out[name] = {"ms": round(t * 1e3, 3), "tflops": round(fl / t / 1e12, 3),
                     "failed": bad, "tuned_nb": nb.value if tune else None,
                     "tuned_wg_per_cu": wg.value if tune else None}
        log(f"batched {name}: {out[name]}")
    del A0, A, B0, Bw, info, piv, Bg
    return out


def bench_adjoint(fla, np, n):
    """GP log-marginal-likelihood gradient (RBF, l=0.2, s=1, noise 1e-3) at n
    points in [0,1]^2: kernel matrix, FLA_Chol, solve, Cholesky adjoint,
    solve adjoint, contraction.  Declared flops: n^3/3 (factor) + n^3/3
    (L^T Lbar) + 2 n^3 (two n x n triangular solves)."""
    import ctypes as C
    lib = fla.lib()
    rng = np.random.default_rng(8192)
    x = np.asfortranarray(rng.uniform(0.0, 1.0, size=(n, 2)))
    y = rng.standard_normal(n)
    dx = fla.DeviceMatrix.from_host(x)
    dy = fla.DeviceMatrix.from_host(y.reshape(-1, 1))
    lml, g, info = C.c_double(), (C.c_double * 2)(), C.c_int()
    ts = []
    for rep in range(3):
        fla.synchronize()
        t0 = time.perf_counter()
        fla.check(lib.fla_hip_gp_rbf_lml_grad(n, 2, dx.ptr, dx.ld, dy.ptr, 0.2, 1.0, 1e-3,
                                              C.byref(lml), g, C.byref(info)))
        ts.append(time.perf_counter() - t0)
    t = min(ts[1:])
    flops = n ** 3 / 3 + n ** 3 / 3 + 2.0 * n ** 3
    out = {"n": n, "ms": round(t * 1e3, 2), "tflops_declared": round(flops / t / 1e12, 3),
           "lml": lml.value, "grad_l": g[0], "grad_s": g[1], "info": info.value}
    log(f"adjoint: {out}")
    out["batched_fp32"] = bench_adjoint_batched(fla, np, 1024, 256)
    return out


def bench_spdinv(fla, n):
    """FLA_SPDinv of an n x n SPD matrix (A = B B'/n + I, lower triangle):
    Cholesky, triangular inverse, L'L.  Flops n^3 (n^3/3 per phase).  The
    phases are timed again one call at a time on a fresh copy."""
    import ctypes as C
    lib = fla.lib()
    lo = fla.FLA_LOWER_TRIANGULAR
    info = C.c_int()
    ts = []
    for rep in range(3):
        A = make_spd(fla, n, 77)
        t0 = time.perf_counter()
        fla.check(lib.fla_hip_dspdinv(lo, n, A.ptr, n, C.byref(info)), "dspdinv")
        fla.synchronize()
        ts.append(time.perf_counter() - t0)
        A.buf.free()
    t = min(ts[1:])
    A = make_spd(fla, n, 77)
    phases = {}
    for name, call in (("chol", lambda: lib.fla_hip_dpotrf(lo, n, A.ptr, n, C.byref(info))),
                       ("trinv", lambda: lib.fla_hip_dtrtri(lo, fla.FLA_NONUNIT_DIAG, n, A.ptr, n)),
                       ("ttmm", lambda: lib.fla_hip_dlauum(lo, n, A.ptr, n))):
        t0 = time.perf_counter()
        fla.check(call(), name)
        fla.synchronize()
        dt = time.perf_counter() - t0
        phases[name] = {"ms": round(dt * 1e3, 2), "tflops": round(n ** 3 / 3 / dt / 1e12, 3)}
    A.buf.free()
    out = {"n": n, "ms": round(t * 1e3, 2), "tflops": round(n ** 3 / t / 1e12, 3),
           "info": info.value, "phases": phases}
    log(f"spdinv: {out}")
    return out


def bench_lu_nopiv(fla, np, n):
    """FLA_LU_nopiv of an n x n diagonally dominant matrix (U(-1,1) entries
    plus n on the diagonal, made on the host, uploaded before each timed
    call); flops 2 n^3 / 3."""
    import ctypes as C
    lib = fla.lib()
    info = C.c_int()
    rng = np.random.default_rng(78)
    H = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(n, n)))
    H[np.diag_indices(n)] += n
    A = fla.DeviceMatrix(n, n)
    ts = []
    for rep in range(3):
        A.upload(H)
        fla.synchronize()
        t0 = time.perf_counter()
        fla.check(lib.fla_hip_dgetrf_nopiv(n, n, A.ptr, n, C.byref(info)), "dgetrf_nopiv")
        fla.synchronize()
        ts.append(time.perf_counter() - t0)
    A.buf.free()
    del H
    t = min(ts[1:])
    out = {"n": n, "ms": round(t * 1e3, 2), "tflops": round(2 * n ** 3 / 3 / t / 1e12, 3),
           "info": info.value}
    log(f"lu_nopiv: {out}")
    return out


def bench_flash(fla, np, n, b, n_inc, b_inc, nb_alg):
    """FLASH front-ends: FLASH_Chol on an n x n hierarchical object with
    blocks of b (the flat copy, the factorization and the scatter back are
    all inside the timed call) beside FLA_Chol on the same matrix, and
    FLASH_LU_incpiv at n_inc with tiles of b_inc and inner block nb_alg
    (flops 2 n^3/3, the LU convention, so the tile algorithm's extra work
    is not counted), on U(-1,1) matrices made on the host."""
    import ctypes as C
    lib = fla.lib()
    lo = fla.FLA_LOWER_TRIANGULAR
    A0 = make_spd(fla, n, 91)
    F = fla.DeviceMatrix(n, n)
    Fo = F.obj()
    bf = (C.c_ulong * 1)(b)
    H = fla.FLA_Obj()
    fla.check_fla(lib.FLASH_Obj_create(fla.FLA_DOUBLE, n, n, 1, bf, C.byref(H)), "create")
    tf, th = [], []
    for rep in range(3):
        copy(fla, F, A0)
        fla.synchronize()
        t0 = time.perf_counter()
        assert lib.FLA_Chol(lo, Fo.o) == fla.FLA_SUCCESS
        fla.synchronize()
        tf.append(time.perf_counter() - t0)
        copy(fla, F, A0)
        fla.check_fla(lib.FLASH_Copy_flat_to_hier(Fo.o, 0, 0, H), "hierarchify")
        fla.synchronize()
        t0 = time.perf_counter()
        assert lib.FLASH_Chol(lo, H) == fla.FLA_SUCCESS
        fla.synchronize()
        th.append(time.perf_counter() - t0)
    lib.FLASH_Obj_free(C.byref(H))
    A0.buf.free()
    del F, Fo
    t_flat, t_hier = min(tf[1:]), min(th[1:])
    out = {"chol_n": n, "chol_block": b, "flash_chol_ms": round(t_hier * 1e3, 2),
           "fla_chol_ms": round(t_flat * 1e3, 2),
           "flash_chol_tflops": round(n ** 3 / 3 / t_hier / 1e12, 3),
           "flash_overhead_pct": round(100.0 * (t_hier - t_flat) / t_flat, 2)}
    rng = np.random.default_rng(92)
    Ah = np.asfortranarray(rng.uniform(-1.0, 1.0, size=(n_inc, n_inc)))
    Af = fla.DeviceMatrix.from_host(Ah)
    Afo = Af.obj()
    ti = []
    for rep in range(3):
        HA, Hp, HL = fla.FLA_Obj(), fla.FLA_Obj(), fla.FLA_Obj()
        bi = (C.c_ulong * 1)(b_inc)
        fla.check_fla(lib.FLASH_LU_incpiv_create_hier_matrices(
            Afo.o, 1, bi, nb_alg, C.byref(HA), C.byref(Hp), C.byref(HL)), "create incpiv")
        fla.synchronize()
        t0 = time.perf_counter()
        assert lib.FLASH_LU_incpiv(HA, Hp, HL) == fla.FLA_SUCCESS
        fla.synchronize()
        ti.append(time.perf_counter() - t0)
        for o in (HA, Hp, HL):
            lib.FLASH_Obj_free(C.byref(o))
    t = min(ti[1:])
    out.update({"incpiv_n": n_inc, "incpiv_block": b_inc, "incpiv_nb_alg": nb_alg,
                "incpiv_ms": round(t * 1e3, 2),
                "incpiv_tflops": round(2 * n_inc ** 3 / 3 / t / 1e12, 3)})
    log(f"flash: {out}")
    return out


def bench_adjoint_batched(fla, np, count, n):
    """Batched FP32 GP gradient: count problems of n points in [0,1]^2, l_i
    uniform on [0.05, 0.5], s = 1, noise 1e-3, y ~ N(0,1) (one launch each for
    the kernel matrices, the batched Cholesky + solve and the batched
    adjoint + contraction).  Declared flops per problem as for the single
    problem: n^3/3 + n^3/3 + 2 n^3."""
    lib = fla.lib()
    rng = np.random.default_rng(1024)
    x = rng.uniform(0.0, 1.0, size=count * n * 2).astype(np.float32)
    y = rng.standard_normal(count * n).astype(np.float32)
    ls = rng.uniform(0.05, 0.5, size=count).astype(np.float32)
    dx = fla.DeviceMatrix.from_host(x.reshape(-1, 1), np.float32)
    dy = fla.DeviceMatrix.from_host(y.reshape(-1, 1), np.float32)
    dl = fla.DeviceMatrix.from_host(ls.reshape(-1, 1), np.float32)
    dlml = fla.DeviceMatrix(count, 1, np.float32)
    dg = fla.DeviceMatrix(2 * count, 1, np.float32)
    dinfo = fla.DeviceMatrix(count, 1, np.int32)
    ts = []
    for rep in range(4):
        fla.synchronize()
        t0 = time.perf_counter()
        fla.check(lib.fla_hip_sgp_rbf_lml_grad_batched(count, n, 2, dx.ptr, dy.ptr, dl.ptr, 1.0,
                                                       1e-3, dlml.ptr, dg.ptr, dinfo.ptr))
        fla.synchronize()
        ts.append(time.perf_counter() - t0)
    t = min(ts[1:])
    info = dinfo.download().ravel()
    flops = count * (n ** 3 / 3 + n ** 3 / 3 + 2.0 * n ** 3)
    out = {"count": count, "n": n, "ms": round(t * 1e3, 3),
           "tflops_declared": round(flops / t / 1e12, 3),
           "failed": int(np.sum(info != fla.FLA_SUCCESS)),
           "lml_mean": float(np.nanmean(dlml.download()))}
    log(f"adjoint batched fp32: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--lu-n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--chol-nb", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=8192)
    ap.add_argument("--cpu-lu-n", type=int, default=6144)
    ap.add_argument("--no-lu", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--mp-n", type=int, default=16384)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--adj-n", type=int, default=8192)
    ap.add_argument("--no-inv", action="store_true")
    ap.add_argument("--no-flash", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched leg, untuned (default configuration), and "
                         "print its results; for counter passes, where the tuner's many small "
                         "launches are slow")
    ap.add_argument("--inv-n", type=int, default=16384)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import libflame_amd as fla
    fla.FLA_Init()
    if args.batched_only:
        print(json.dumps({"batched": bench_batched(fla, np, args.batch_count, tune=False)}),
              flush=True)
        fla.FLA_Finalize()
        return
    if args.chol_nb:
        fla.set_blocksizes(chol_nb=args.chol_nb)
    n = args.n
    log(f"generating SPD A (n={n}) on the GPU")
    A0 = make_spd(fla, n, 20261202)
    A = fla.DeviceMatrix(n, n)
    Aobj = A.obj()

    for w in range(args.warmup):
        copy(fla, A, A0)
        r = fla.FLA_Chol(fla.FLA_LOWER_TRIANGULAR, Aobj)
        log(f"warmup {w}: FLA_Chol -> {r}")
        assert r == fla.FLA_SUCCESS
    fla.stats_reset()
    fla.stats_enable(True)
    times = []
    for s in range(args.steps):
        copy(fla, A, A0)
        fla.synchronize()
        t0 = time.perf_counter()
        r = fla.FLA_Chol(fla.FLA_LOWER_TRIANGULAR, Aobj)  # returns after the device is done
        fla.synchronize()
        times.append(time.perf_counter() - t0)
        assert r == fla.FLA_SUCCESS, r
        log(f"step {s}: {times[-1] * 1e3:.1f} ms ({n ** 3 / 3 / times[-1] / 1e12:.2f} TFLOP/s)")
    fla.stats_enable(False)
    upd_flops, upd_ms, upd_n = fla.stats_get(fla.STATS_CHOL_UPDATE)
    t_chol = sum(times) / len(times)
    tflops = n ** 3 / 3.0 / t_chol / 1e12
    achieved = upd_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0
    resid = chol_probe_residual(fla, np, A0, A, n)
    log(f"chol: {tflops:.2f} TFLOP/s, update kernel {achieved:.2f} TFLOP/s over {upd_n} launches, "
        f"probe residual {resid:.3f}")

    # Solve with the factor: dpotrs-style, 16 RHS
    B = fla.DeviceMatrix.from_host(np.random.default_rng(1).standard_normal((n, args.nrhs)))
    X = fla.DeviceMatrix(n, args.nrhs)

    def potrs():
        for tr in (fla.FLA_NO_TRANSPOSE, fla.FLA_TRANSPOSE):
            fla.check(fla.lib().fla_hip_dtrsm(fla.FLA_LEFT, fla.FLA_LOWER_TRIANGULAR, tr,
                                              fla.FLA_NONUNIT_DIAG, n, args.nrhs, 1.0, A.ptr, n,
                                              X.ptr, n))
    copy(fla, X, B)
    potrs()  # warm-up: sizes the workspace
    copy(fla, X, B)
    fla.synchronize()
    t0 = time.perf_counter()
    potrs()
    fla.synchronize()
    t_potrs = time.perf_counter() - t0
    extras = {"chol_probe_residual": round(resid, 4),
              "potrs_ms": round(t_potrs * 1e3, 3),
              "potrs_GBps": round(8.0 * n * n / t_potrs / 1e9, 1)}
    del B, X

    if not args.no_lu:
        m = args.lu_n
        log(f"LU: n={m}")
        L0 = make_lu(fla, m, 7)
        LA = fla.DeviceMatrix(m, m)
        piv = fla.DeviceMatrix(m, 1, np.int32)
        copy(fla, LA, L0)
        r = fla.FLA_LU_piv(LA.obj(), piv.obj())
        assert r == fla.FLA_SUCCESS, r
        fla.stats_reset()
        fla.stats_enable(True)
        lt = []
        for s in range(max(1, args.steps)):
            copy(fla, LA, L0)
            fla.synchronize()
            t0 = time.perf_counter()
            r = fla.FLA_LU_piv(LA.obj(), piv.obj())
            fla.synchronize()
            lt.append(time.perf_counter() - t0)
            log(f"LU step {s}: {lt[-1] * 1e3:.1f} ms ({2 * m ** 3 / 3 / lt[-1] / 1e12:.2f} TFLOP/s)")
        fla.stats_enable(False)
        gf, gms, gn = fla.stats_get(fla.STATS_LU_UPDATE)
        pf, pms, pn = fla.stats_get(fla.STATS_LU_PANEL)
        sf, sms, sn = fla.stats_get(fla.STATS_LU_SWAP)
        t_lu = sum(lt) / len(lt)
        nrep = len(lt)
        swap_bytes = lu_moved_bytes(np, piv.download().ravel(), m)
        # panel kernel: the (m-k) x 128 panel read once and written once
        panel_bytes = sum(2.0 * 8 * (m - k) * min(128, m - k) for k in range(0, m, 128))
        upd_tf = gf / (gms * 1e-3) / 1e12 if gms else 0.0
        extras.update({"lu_n": m, "lu_tflops": round(2 * m ** 3 / 3 / t_lu / 1e12, 3),
                       "lu_ms": round(t_lu * 1e3, 2),
                       "lu_update_kernel_tflops": round(upd_tf, 3),
                       "lu_panel_ms_per_factorization": round(pms / nrep, 2),
                       "lu_panel_GBps": round(panel_bytes / (pms / nrep * 1e-3) / 1e9, 1) if pms else None,
                       "lu_swap_ms_per_factorization": round(sms / nrep, 2),
                       "lu_swap_GBps": round(swap_bytes / (sms / nrep * 1e-3) / 1e9, 1) if sms else None,
                       "lu_roofline": {"bound": "mfma", "achieved": round(upd_tf, 3),
                                       "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                                       "frac": round(upd_tf / PEAK_FP64_MFMA, 4),
                                       "kernel": "dgemm_nt_kp_kernel (LU A22 update, U12 transposed)",
                                       "launches": gn // nrep}})
        log(f"LU: update {upd_tf:.2f} TFLOP/s, panels {pms / nrep:.1f} ms, "
            f"interchanges {sms / nrep:.1f} ms ({swap_bytes / 1e9:.2f} GB)")
        del L0, LA, piv

    if not args.no_mixed:
        extras["mixed"] = bench_mixed(fla, np, args.mp_n)
        extras["mixed_lu"] = bench_mixed_lu(fla, np, args.mp_n)
    if not args.no_batched:
        extras["batched"] = bench_batched(fla, np, args.batch_count)
    if not args.no_adjoint:
        extras["adjoint"] = bench_adjoint(fla, np, args.adj_n)
    if not args.no_inv:
        extras["spdinv"] = bench_spdinv(fla, args.inv_n)
        extras["lu_nopiv"] = bench_lu_nopiv(fla, np, args.inv_n)
    if not args.no_flash:
        extras["flash"] = bench_flash(fla, np, args.inv_n, 1024, 4096, 256, 64)

    cpu = None if args.no_cpu else cpu_baseline(fla, args.cpu_n, args.cpu_lu_n)
    out = {
        "metric": "Chol/LU TFLOP/s @ n=32768 FP64 vs mixed-prec+IR time-to-solve at equal bwd err",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_chol * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FLA_Chol lower FP64, n=32768, A=B*B'/n+I, B~N(0,1)",
                   "n": n, "blocksize": fla.lib().fla_hip_chol_blocksize(n),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_MFMA, 4),
                     "traffic": profiled_traffic(),
                     "kernel": "dgemm_nt_kp_kernel (lower SYRK update)",
                     "launches": upd_n},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    # release the library's streams, events and buffers while the runtime is up
    fla.FLA_Finalize()


if __name__ == "__main__":
    main()
