/*
 * FLA_hip.h -- thin C ABI between libflame_amd's C host layer and its
 * hand-written gfx950 HIP kernels.  Plain pointers, sizes and an int status
 * (0 = success, otherwise a hipError_t value or one of FLA_HIP_E*).
 *
 * Every matrix argument is a DEVICE pointer to column-major storage with
 * leading dimension ld (element (i,j) at A[i + j*ld]).  All calls are
 * asynchronous on the library's streams unless they return host data
 * (info / pivot indices), in which case they synchronize once at the end.
 *
 * Each entry point replaces one seam of the reference's HIP/SuperMatrix path
 * (paths relative to the libflame source tree):
 *   fla_hip_dgemm        FLA_Gemm_external_hip   src/base/flamec/wrappers/blas/3/hip/FLA_Gemm_external_hip.c:19 (rocblas_dgemm :142)
 *   fla_hip_dsyrk        FLA_Herk_external_hip / FLA_Syrk_external_hip
 *                        src/base/flamec/wrappers/blas/3/hip/FLA_Herk_external_hip.c (rocblas_dsyrk :107)
 *   fla_hip_dtrsm        FLA_Trsm_external_hip   src/base/flamec/wrappers/blas/3/hip/FLA_Trsm_external_hip.c:19 (rocblas_dtrsm :102)
 *   fla_hip_dpotrf       FLA_Chol_blk_external_hip src/base/flamec/wrappers/lapack/hip/FLA_Chol_blk_external_hip.c:30 (rocsolver_dpotrf :75)
 *   fla_hip_dgetrf       FLA_LU_piv_blk_external_hip src/base/flamec/wrappers/lapack/hip/FLA_LU_piv_blk_external_hip.c:20 (rocsolver_dgetrf :75)
 *   fla_hip_dlaswp       FLA_Apply_pivots_unb_external_hip src/base/flamec/wrappers/lapack/hip/FLA_Apply_pivots_unb_external_hip.c:20 (rocsolver_dlaswp :92)
 * and the SuperMatrix HIP scheduler (FLASH_Queue_exec_hip,
 * src/base/flamec/supermatrix/main/FLASH_Queue_exec.c:2302) is replaced by the
 * one-GPU look-ahead schedule inside fla_hip_dpotrf / fla_hip_dgetrf.
 */
#ifndef FLA_HIP_H
#define FLA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLA_HIP_OK 0
#define FLA_HIP_EARG 9001   /* invalid argument */
#define FLA_HIP_ENODEV 9002 /* no gfx950 device available */
#define FLA_HIP_ESPIN 9003  /* a bounded inter-workgroup wait expired */

/* sides / uplo / trans / diag use the FLAME constants (FLAME.h). */

/* ---- runtime ---- */
int fla_hip_init(void);     /* selects device 0, creates the two streams */
int fla_hip_finalize(void); /* releases streams, events and workspaces */
int fla_hip_synchronize(void);
int fla_hip_malloc(void** ptr, size_t bytes);
int fla_hip_free(void* ptr);
int fla_hip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int fla_hip_memcpy_d2h(void* dst, const void* src, size_t bytes);
int fla_hip_memcpy_d2d(void* dst, const void* src, size_t bytes);
/* 2-D strided copies of an m x n column-major double matrix (sync). */
int fla_hip_dmatrix_h2d(int m, int n, const double* h, int ldh, double* d, int ldd);
int fla_hip_dmatrix_d2h(int m, int n, const double* d, int ldd, double* h, int ldh);
int fla_hip_dmatrix_d2d(int m, int n, const double* s, int lds, double* d, int ldd);
/* Pitched copy of `height` rows of `width` bytes; kind 0 = host->device,
   1 = device->host, 2 = device->device (synchronous). */
int fla_hip_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                     size_t height, int kind);
/* 1 if ptr is device memory the library can launch kernels on, else 0. */
int fla_hip_is_device_ptr(const void* ptr);
const char* fla_hip_error_string(int status);

/* ---- level-3 kernels (MFMA FP64) ---- */
/* C := alpha*op(A)*op(B) + beta*C; trans = FLA_NO_TRANSPOSE / FLA_TRANSPOSE. */
int fla_hip_dgemm(int transa, int transb, int m, int n, int k, double alpha,
                  const double* A, int lda, const double* B, int ldb, double beta,
                  double* C, int ldc);
/* C := alpha*op(A)*op(A)' + beta*C on the uplo triangle only. */
int fla_hip_dsyrk(int uplo, int trans, int n, int k, double alpha, const double* A,
                  int lda, double beta, double* C, int ldc);
/* B := alpha*inv(op(A))*B (left) or alpha*B*inv(op(A)) (right). */
int fla_hip_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                  const double* A, int lda, double* B, int ldb);

/* ---- factorizations ---- */
/* Blocked Cholesky, in place on the uplo triangle.  *info = FLA_SUCCESS (-1)
   or the 0-based index of the first column whose pivot is <= 0 or NaN. */
int fla_hip_dpotrf(int uplo, int n, double* A, int lda, int* info);
/* Blocked LU with partial pivoting.  p (device, int[min(m,n)]) receives
   FLAME-native pivots (0-based, relative to the row).  *info = FLA_SUCCESS or
   the 0-based index of the first null pivot. */
int fla_hip_dgetrf(int m, int n, double* A, int lda, int* p, int* info);
/* Row interchanges i <-> i + p[i] for i = k1..k2 (in that order, or reversed
   when incp < 0) on the ncols columns of A.  p is a device pointer. */
int fla_hip_dlaswp(int ncols, double* A, int lda, int k1, int k2, const int* p, int incp);
/* p[i] += (i+1) (to LAPACK) or -= (i+1) (to FLAME) on a device int vector. */
int fla_hip_shift_pivots(int ptype, int m, int* p);

/* ---- tuning ---- */
/* Outer (trailing-update) and inner (panel) block sizes; 0 keeps current. */
int fla_hip_set_blocksizes(int chol_nb, int lu_nb);
int fla_hip_get_blocksizes(int* chol_nb, int* lu_nb);
/* Tile configuration of the MFMA update kernel (0 = default). */
int fla_hip_set_gemm_config(int cfg);

/* ---- measurement ---- */
/* When enabled, the factorization drivers bracket every trailing-update
   launch with HIP events on its stream and accumulate, per kind, the
   algorithmic flops and the elapsed milliseconds.
   kind 0 = Cholesky trailing update (lower SYRK), 1 = LU trailing update
   (GEMM), 2 = LU panel, 3 = Cholesky panel chain. */
int fla_hip_stats_enable(int on);
int fla_hip_stats_reset(void);
int fla_hip_stats_get(int kind, double* flops, double* ms, long* launches);

/* ---- synthetic inputs and checks (device side) ---- */
/* Counter-based generators: identical values to the CPU generators. */
int fla_hip_gen_uniform(unsigned long long seed, int m, int n, double* A, int lda);
int fla_hip_gen_normal(unsigned long long seed, int m, int n, double* A, int lda);
/* A := scale*B*B' + shift*I (lower triangle and diagonal written). */
int fla_hip_gram_lower(int n, int k, const double* B, int ldb, double scale,
                       double shift, double* A, int lda);
/* A(i,:) *= 10^(-i/max(m,n)) */
int fla_hip_scale_rows_logspace(int m, int n, double* A, int lda);
/* Y := op(M)*X + beta*Y with M given by its uplo triangle as a symmetric
   matrix (sym=1), a triangular factor (sym=0, unit diag per `diag`), or a
   full matrix (uplo = FLA_FULL_MATRIX).  FP64, used by residual checks. */
int fla_hip_dmatmul_check(int kind, int uplo, int trans, int diag, int n, int nrhs,
                          const double* M, int ldm, const double* X, int ldx, double beta,
                          double* Y, int ldy);

#ifdef __cplusplus
}
#endif
#endif
