"""Helpers for the GPU tests: run libflame_amd entry points on device
buffers made from host arrays, and bring the results back."""
import ctypes as C

import numpy as np

import libflame_amd as fla


def to_dev(a, dtype=np.float64):
    return fla.DeviceMatrix.from_host(np.asfortranarray(a), dtype)


def gpu_chol(A, uplo=fla.FLA_LOWER_TRIANGULAR):
    d = to_dev(A)
    r = fla.FLA_Chol(uplo, d.obj())
    return d.download(), r


def gpu_lu(A):
    m, n = A.shape
    d = to_dev(A)
    p = fla.DeviceMatrix(min(m, n), 1, np.int32)
    r = fla.FLA_LU_piv(d.obj(), p.obj())
    return d.download(), p.download().ravel().copy(), r


def gpu_trsm(side, uplo, trans, diag, T, B, alpha=1.0):
    dT, dB = to_dev(T), to_dev(B)
    m, n = B.shape
    fla.check(fla.lib().fla_hip_dtrsm(side, uplo, trans, diag, m, n, alpha, dT.ptr, dT.ld,
                                      dB.ptr, dB.ld), "dtrsm")
    fla.synchronize()
    return dB.download()


def gpu_gemm(ta, tb, A, B, Cm, alpha=-1.0, beta=1.0):
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(Cm)
    m, n = Cm.shape
    k = A.shape[1] if ta == fla.FLA_NO_TRANSPOSE else A.shape[0]
    fla.check(fla.lib().fla_hip_dgemm(ta, tb, m, n, k, alpha, dA.ptr, dA.ld, dB.ptr, dB.ld, beta,
                                      dC.ptr, dC.ld), "dgemm")
    fla.synchronize()
    return dC.download()


def gpu_syrk(uplo, trans, A, Cm, alpha=-1.0, beta=1.0):
    dA, dC = to_dev(A), to_dev(Cm)
    n = Cm.shape[0]
    k = A.shape[1] if trans == fla.FLA_NO_TRANSPOSE else A.shape[0]
    fla.check(fla.lib().fla_hip_dsyrk(uplo, trans, n, k, alpha, dA.ptr, dA.ld, beta, dC.ptr,
                                      dC.ld), "dsyrk")
    fla.synchronize()
    return dC.download()


def ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))
