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
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
EPS = 2.0 ** -53


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


def cpu_baseline(n_cpu):
    """The oracle's FLA_Chol (libflame's control tree, f2c loop order, one
    thread) on a bounded n_cpu sample of the same matrix family."""
    import numpy as np
    from oracle import oracle as orc
    A = orc.gen_spd_c2(11, n_cpu)
    t0 = time.perf_counter()
    _, r = orc.chol(A)
    dt = time.perf_counter() - t0
    assert r == orc.SUCCESS
    return {"value": round(n_cpu ** 3 / 3.0 / dt / 1e12, 6), "unit": "TFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle FLA_Chol lower (var3 128/32 + opt_var2, f2c BLAS order), "
                      f"n={n_cpu}, A=B*B'/n+I, {dt:.1f} s on one host core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--lu-n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--chol-nb", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=3072)
    ap.add_argument("--no-lu", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import libflame_amd as fla
    fla.FLA_Init()
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
        t_lu = sum(lt) / len(lt)
        extras.update({"lu_n": m, "lu_tflops": round(2 * m ** 3 / 3 / t_lu / 1e12, 3),
                       "lu_ms": round(t_lu * 1e3, 2),
                       "lu_update_kernel_tflops": round(gf / (gms * 1e-3) / 1e12, 3) if gms else None,
                       "lu_panel_ms_per_factorization": round(pms / len(lt), 2)})
        del L0, LA, piv

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)
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
                   "n": n, "blocksize": fla.get_blocksizes()[0], "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_MFMA, 4),
                     "traffic": None, "kernel": "dgemm_mfma_kernel<false,true> (lower SYRK update)",
                     "launches": upd_n},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
