/*
 * lapack2flame.c -- LAPACK-ABI entry points over libflame_amd.
 *
 *   dpotrf_ / dpotf2_  src/map/lapack2flamec/FLA_potrf.c:31-105
 *                      (LAPACK_potrf_body :38-60, info = idx + 1 :57-58)
 *   dgetrf_ / dgetf2_  src/map/lapack2flamec/FLA_getrf.c:32-117
 *                      (LAPACK_getrf_body :40-67: p zeroed, FLA_LU_piv,
 *                       FLA_Shift_pivots_to(FLA_LAPACK_PIVOTS))
 *   dpotrs_            src/map/common/lapacksrc/f2c/dpotrs.c:102-194
 *   dgetrs_            src/map/common/lapacksrc/f2c/dgetrs.c:115-210
 *   dlaswp_            src/map/common/lapacksrc/f2c/dlaswp.c:103-213
 *   dposv_ / dgesv_    src/map/common/lapacksrc/f2c/dposv.c:120-190,
 *                      dgesv.c:113-180 (factor, then solve when info == 0)
 *   dtrtri_ / dlauum_ / dpotri_
 *                      src/map/lapack2flamec/FLA_trtri.c, FLA_lauum.c,
 *                      FLA_potri.c (FLA_Trinv / FLA_Ttmm on the factor)
 *
 * The buffers are host memory (the LAPACK ABI); FLA_Chol / FLA_LU_piv stage
 * them through device memory.  Argument errors follow the netlib checks
 * (dpotrf_check / dgetrf_check in src/map/common/lapackcheck): info = -i for
 * the i-th argument, reported through an xerbla-style message.
 */
#include <stdio.h>
#include <stdlib.h>

#include "FLAME.h"
#include "FLA_hip.h"
#include "fla_internal.h"

static int up(char c) { return (c >= 'a' && c <= 'z') ? c - 32 : c; }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

static void xerbla(const char* name, int info) {
  fprintf(stderr, " ** On entry to %s parameter number %d had an illegal value\n", name, info);
}

static int potrf_body(const char* name, char* uplo, int* n, double* a, int* lda, int* info) {
  FLA_Obj A;
  FLA_Error e, init_result;
  int u = up(*uplo);
  *info = 0;
  if (u != 'U' && u != 'L')
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*lda < imax(1, *n))
    *info = -4;
  if (*info != 0) {
    xerbla(name, -*info);
    return 0;
  }
  if (*n == 0) return 0;
  FLA_Init_safe(&init_result);
  FLA_Obj_create_without_buffer(FLA_DOUBLE, (dim_t)*n, (dim_t)*n, &A);
  FLA_Obj_attach_buffer(a, 1, (dim_t)*lda, &A);
  e = FLA_Chol(u == 'L' ? FLA_LOWER_TRIANGULAR : FLA_UPPER_TRIANGULAR, A);
  FLA_Obj_free_without_buffer(&A);
  FLA_Finalize_safe(init_result);
  *info = (e != FLA_SUCCESS) ? e + 1 : 0;
  return 0;
}

int dpotrf_(char* uplo, int* n, double* buff_A, int* ldim_A, int* info) {
  return potrf_body("DPOTRF", uplo, n, buff_A, ldim_A, info);
}

int dpotf2_(char* uplo, int* n, double* buff_A, int* ldim_A, int* info) {
  return potrf_body("DPOTF2", uplo, n, buff_A, ldim_A, info);
}

static int getrf_body(const char* name, int* m, int* n, double* a, int* lda, int* ipiv,
                      int* info) {
  FLA_Obj A, p;
  FLA_Error e, init_result;
  int mn;
  *info = 0;
  if (*m < 0)
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*lda < imax(1, *m))
    *info = -4;
  if (*info != 0) {
    xerbla(name, -*info);
    return 0;
  }
  if (*m == 0 || *n == 0) return 0;
  mn = imin(*m, *n);
  FLA_Init_safe(&init_result);
  FLA_Obj_create_without_buffer(FLA_DOUBLE, (dim_t)*m, (dim_t)*n, &A);
  FLA_Obj_attach_buffer(a, 1, (dim_t)*lda, &A);
  FLA_Obj_create_without_buffer(FLA_INT, (dim_t)mn, 1, &p);
  FLA_Obj_attach_buffer(ipiv, 1, (dim_t)mn, &p);
  FLA_Set(FLA_ZERO, p);
  e = FLA_LU_piv(A, p);
  FLA_Shift_pivots_to(FLA_LAPACK_PIVOTS, p);
  FLA_Obj_free_without_buffer(&A);
  FLA_Obj_free_without_buffer(&p);
  FLA_Finalize_safe(init_result);
  *info = (e != FLA_SUCCESS) ? e + 1 : 0;
  return 0;
}

int dgetrf_(int* m, int* n, double* buff_A, int* ldim_A, int* buff_p, int* info) {
  return getrf_body("DGETRF", m, n, buff_A, ldim_A, buff_p, info);
}

int dgetf2_(int* m, int* n, double* buff_A, int* ldim_A, int* buff_p, int* info) {
  return getrf_body("DGETF2", m, n, buff_A, ldim_A, buff_p, info);
}

/* Stages A (n x n) and B (n x nrhs) on the device, runs the two (dpotrs) or
   pivot + two (dgetrs) triangular solves, and copies B back. */
int dpotrs_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb,
            int* info) {
  int u = up(*uplo), st;
  double *dA = NULL, *dB = NULL;
  *info = 0;
  if (u != 'U' && u != 'L')
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*nrhs < 0)
    *info = -3;
  else if (*lda < imax(1, *n))
    *info = -5;
  else if (*ldb < imax(1, *n))
    *info = -7;
  if (*info != 0) {
    xerbla("DPOTRS", -*info);
    return 0;
  }
  if (*n == 0 || *nrhs == 0) return 0;
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)*n * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dB, sizeof(double) * (size_t)*n * (size_t)*nrhs);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *n, a, *lda, dA, *n);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *nrhs, b, *ldb, dB, *n);
  if (st == FLA_HIP_OK) {
    if (u == 'U') {
      st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_TRANSPOSE, FLA_NONUNIT_DIAG, *n,
                         *nrhs, 1.0, dA, *n, dB, *n);
      if (st == FLA_HIP_OK)
        st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG,
                           *n, *nrhs, 1.0, dA, *n, dB, *n);
    } else {
      st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG, *n,
                         *nrhs, 1.0, dA, *n, dB, *n);
      if (st == FLA_HIP_OK)
        st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_TRANSPOSE, FLA_NONUNIT_DIAG, *n,
                           *nrhs, 1.0, dA, *n, dB, *n);
    }
  }
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(*n, *nrhs, dB, *n, b, *ldb);
  fla_hip_free(dA);
  fla_hip_free(dB);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: dpotrs_ failed: %s\n", fla_hip_error_string(st));
    *info = -100;
  }
  return 0;
}

int dgetrs_(char* trans, int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b,
            int* ldb, int* info) {
  int t = up(*trans), st, i;
  double *dA = NULL, *dB = NULL;
  int* dp = NULL;
  int* fp = NULL;
  *info = 0;
  if (t != 'N' && t != 'T' && t != 'C')
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*nrhs < 0)
    *info = -3;
  else if (*lda < imax(1, *n))
    *info = -5;
  else if (*ldb < imax(1, *n))
    *info = -8;
  if (*info != 0) {
    xerbla("DGETRS", -*info);
    return 0;
  }
  if (*n == 0 || *nrhs == 0) return 0;
  /* LAPACK (1-based absolute) pivots -> FLAME-native (0-based relative). */
  fp = (int*)malloc(sizeof(int) * (size_t)*n);
  if (!fp) return 0;
  for (i = 0; i < *n; ++i) fp[i] = ipiv[i] - (i + 1);
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)*n * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dB, sizeof(double) * (size_t)*n * (size_t)*nrhs);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dp, sizeof(int) * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *n, a, *lda, dA, *n);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *nrhs, b, *ldb, dB, *n);
  if (st == FLA_HIP_OK) st = fla_hip_memcpy_h2d(dp, fp, sizeof(int) * (size_t)*n);
  if (st == FLA_HIP_OK) {
    if (t == 'N') {
      st = fla_hip_dlaswp(*nrhs, dB, *n, 0, *n - 1, dp, 1);
      if (st == FLA_HIP_OK)
        st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_UNIT_DIAG, *n,
                           *nrhs, 1.0, dA, *n, dB, *n);
      if (st == FLA_HIP_OK)
        st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG,
                           *n, *nrhs, 1.0, dA, *n, dB, *n);
    } else {
      st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_TRANSPOSE, FLA_NONUNIT_DIAG, *n,
                         *nrhs, 1.0, dA, *n, dB, *n);
      if (st == FLA_HIP_OK)
        st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_TRANSPOSE, FLA_UNIT_DIAG, *n,
                           *nrhs, 1.0, dA, *n, dB, *n);
      if (st == FLA_HIP_OK) st = fla_hip_dlaswp(*nrhs, dB, *n, 0, *n - 1, dp, -1);
    }
  }
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(*n, *nrhs, dB, *n, b, *ldb);
  fla_hip_free(dA);
  fla_hip_free(dB);
  fla_hip_free(dp);
  free(fp);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: dgetrs_ failed: %s\n", fla_hip_error_string(st));
    *info = -100;
  }
  return 0;
}

/* dlaswp_ (src/map/common/lapacksrc/f2c/dlaswp.c:103-213): interchange row i
   with row ipiv(ix) for i = k1..k2 (incx > 0, ix = k1 + (i-k1)*incx) or for
   i = k2..k1 (incx < 0, ix = 1 + (i-1)*|incx|); incx == 0 does nothing.  The
   rows that the interchanges reach are staged to the device, permuted there
   by the device interchange kernels, and copied back. */
int dlaswp_(int* n, double* a, int* lda, int* k1, int* k2, int* ipiv, int* incx) {
  int i, st, rows, K, ncols = *n;
  int* rel = NULL;
  int* dp = NULL;
  double* dA = NULL;
  if (*incx == 0 || ncols <= 0 || *k2 < *k1 || *k1 < 1) return 0;
  K = *k2; /* the device vector holds row r's pivot at r (rows 0..k2-1) */
  rel = (int*)calloc((size_t)K, sizeof(int));
  if (!rel) return 0;
  /* FLAME-native (0-based, relative) pivot of each row, and the rows the
     interchanges reach */
  rows = *k2;
  for (i = *k1; i <= *k2; ++i) {
    long ix = *incx > 0 ? (long)*k1 + (long)(i - *k1) * *incx : 1 + (long)(i - 1) * -(long)*incx;
    int ip = ipiv[ix - 1];
    rel[i - 1] = ip - i;
    if (ip > rows) rows = ip;
  }
  if (rows > *lda) rows = *lda; /* the device kernels skip targets beyond the rows */
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)rows * (size_t)ncols);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dp, sizeof(int) * (size_t)K);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(rows, ncols, a, *lda, dA, rows);
  if (st == FLA_HIP_OK) st = fla_hip_memcpy_h2d(dp, rel, sizeof(int) * (size_t)K);
  if (st == FLA_HIP_OK)
    st = fla_hip_dlaswp_rows(rows, ncols, dA, rows, *k1 - 1, *k2 - 1, dp, *incx > 0 ? 1 : -1);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(rows, ncols, dA, rows, a, *lda);
  fla_hip_free(dA);
  fla_hip_free(dp);
  free(rel);
  if (st != FLA_HIP_OK) fprintf(stderr, "libflame_amd: dlaswp_ failed: %s\n", fla_hip_error_string(st));
  return 0;
}

/* dposv_ (src/map/common/lapacksrc/f2c/dposv.c:120-190): the netlib checks,
   then the Cholesky factorization (dpotrf_'s semantics, info = first failing
   column + 1) and, when it succeeded, the two triangular solves -- with A and
   B staged to the device once for both. */
int dposv_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb, int* info) {
  int u = up(*uplo), st, e = FLA_SUCCESS, ul;
  double *dA = NULL, *dB = NULL;
  *info = 0;
  if (u != 'U' && u != 'L')
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*nrhs < 0)
    *info = -3;
  else if (*lda < imax(1, *n))
    *info = -5;
  else if (*ldb < imax(1, *n))
    *info = -7;
  if (*info != 0) {
    xerbla("DPOSV ", -*info);
    return 0;
  }
  if (*n == 0) return 0;
  ul = u == 'U' ? FLA_UPPER_TRIANGULAR : FLA_LOWER_TRIANGULAR;
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)*n * (size_t)*n);
  if (st == FLA_HIP_OK && *nrhs > 0)
    st = fla_hip_malloc((void**)&dB, sizeof(double) * (size_t)*n * (size_t)*nrhs);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *n, a, *lda, dA, *n);
  if (st == FLA_HIP_OK) st = fla_hip_dpotrf(ul, *n, dA, *n, &e);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(*n, *n, dA, *n, a, *lda);
  if (st == FLA_HIP_OK && e == FLA_SUCCESS && *nrhs > 0) {
    st = fla_hip_dmatrix_h2d(*n, *nrhs, b, *ldb, dB, *n);
    if (st == FLA_HIP_OK)
      st = fla_hip_dtrsm(FLA_LEFT, ul, ul == FLA_UPPER_TRIANGULAR ? FLA_TRANSPOSE : FLA_NO_TRANSPOSE,
                         FLA_NONUNIT_DIAG, *n, *nrhs, 1.0, dA, *n, dB, *n);
    if (st == FLA_HIP_OK)
      st = fla_hip_dtrsm(FLA_LEFT, ul, ul == FLA_UPPER_TRIANGULAR ? FLA_NO_TRANSPOSE : FLA_TRANSPOSE,
                         FLA_NONUNIT_DIAG, *n, *nrhs, 1.0, dA, *n, dB, *n);
    if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(*n, *nrhs, dB, *n, b, *ldb);
  }
  fla_hip_free(dA);
  fla_hip_free(dB);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: dposv_ failed: %s\n", fla_hip_error_string(st));
    *info = -100;
    return 0;
  }
  *info = (e != FLA_SUCCESS) ? e + 1 : 0;
  return 0;
}

/* dgesv_ (src/map/common/lapacksrc/f2c/dgesv.c:113-180): the netlib checks,
   LU with partial pivoting (dgetrf_'s semantics: LAPACK pivots in ipiv,
   info = first null pivot + 1) and, when no pivot was null, the pivoted
   unit-lower and upper solves -- A, B and the pivots staying on the device
   between the two steps. */
int dgesv_(int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b, int* ldb, int* info) {
  int st, e = FLA_SUCCESS, i;
  double *dA = NULL, *dB = NULL;
  int* dp = NULL;
  *info = 0;
  if (*n < 0)
    *info = -1;
  else if (*nrhs < 0)
    *info = -2;
  else if (*lda < imax(1, *n))
    *info = -4;
  else if (*ldb < imax(1, *n))
    *info = -7;
  if (*info != 0) {
    xerbla("DGESV ", -*info);
    return 0;
  }
  if (*n == 0) return 0;
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)*n * (size_t)*n);
  if (st == FLA_HIP_OK && *nrhs > 0)
    st = fla_hip_malloc((void**)&dB, sizeof(double) * (size_t)*n * (size_t)*nrhs);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dp, sizeof(int) * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *n, a, *lda, dA, *n);
  if (st == FLA_HIP_OK) st = fla_hip_dgetrf(*n, *n, dA, *n, dp, &e);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(*n, *n, dA, *n, a, *lda);
  if (st == FLA_HIP_OK) st = fla_hip_memcpy_d2h(ipiv, dp, sizeof(int) * (size_t)*n);
  if (st == FLA_HIP_OK && e == FLA_SUCCESS && *nrhs > 0) {
    st = fla_hip_dmatrix_h2d(*n, *nrhs, b, *ldb, dB, *n);
    if (st == FLA_HIP_OK) st = fla_hip_dlaswp(*nrhs, dB, *n, 0, *n - 1, dp, 1);
    if (st == FLA_HIP_OK)
      st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_UNIT_DIAG, *n,
                         *nrhs, 1.0, dA, *n, dB, *n);
    if (st == FLA_HIP_OK)
      st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG, *n,
                         *nrhs, 1.0, dA, *n, dB, *n);
    if (st == FLA_HIP_OK) st = fla_hip_dmatrix_d2h(*n, *nrhs, dB, *n, b, *ldb);
  }
  fla_hip_free(dA);
  fla_hip_free(dB);
  fla_hip_free(dp);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: dgesv_ failed: %s\n", fla_hip_error_string(st));
    *info = -100;
    return 0;
  }
  /* FLAME-native (0-based relative) -> LAPACK (1-based absolute) */
  for (i = 0; i < *n; ++i) ipiv[i] += i + 1;
  *info = (e != FLA_SUCCESS) ? e + 1 : 0;
  return 0;
}

/* dsposv_ (src/map/common/lapacksrc/f2c/dsposv.c:192): reduced-precision
   Cholesky (FP32 trailing updates) + refinement with an FP64 residual, the
   netlib stopping rule, 30-step limit and FP64 fall-back (A then holds the
   FP64 factor).  work / swork are accepted for the ABI and not used: the
   workspace lives on the device. */
int dsposv_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb, double* x,
            int* ldx, double* work, float* swork, int* iter, int* info) {
  int u = up(*uplo), st;
  double *dA = NULL, *dB = NULL, *dX = NULL;
  (void)work;
  (void)swork;
  *info = 0;
  *iter = 0;
  if (u != 'U' && u != 'L')
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*nrhs < 0)
    *info = -3;
  else if (*lda < imax(1, *n))
    *info = -5;
  else if (*ldb < imax(1, *n))
    *info = -7;
  else if (*ldx < imax(1, *n))
    *info = -9;
  if (*info != 0) {
    xerbla("DSPOSV", -*info);
    return 0;
  }
  if (*n == 0) return 0;
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)*n * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dB, sizeof(double) * (size_t)*n * (size_t)imax(1, *nrhs));
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dX, sizeof(double) * (size_t)*n * (size_t)imax(1, *nrhs));
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *n, a, *lda, dA, *n);
  if (st == FLA_HIP_OK && *nrhs > 0) st = fla_hip_dmatrix_h2d(*n, *nrhs, b, *ldb, dB, *n);
  if (st == FLA_HIP_OK)
    st = fla_hip_dsposv(u == 'U' ? FLA_UPPER_TRIANGULAR : FLA_LOWER_TRIANGULAR, *n, *nrhs, dA, *n,
                        dB, *n, dX, *n, NULL, iter, info);
  if (st == FLA_HIP_OK && *nrhs > 0) st = fla_hip_dmatrix_d2h(*n, *nrhs, dX, *n, x, *ldx);
  if (st == FLA_HIP_OK && *iter < 0) st = fla_hip_dmatrix_d2h(*n, *n, dA, *n, a, *lda);
  fla_hip_free(dA);
  fla_hip_free(dB);
  fla_hip_free(dX);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: dsposv_ failed: %s\n", fla_hip_error_string(st));
    *info = -100;
  }
  return 0;
}

/* dsgesv_ (src/map/common/lapacksrc/f2c/dsgesv.c:189): reduced-precision LU
   with partial pivoting (FP32 trailing updates) + refinement with an FP64
   residual, the netlib stopping rule, 30-step limit and FP64 fall-back (A
   then holds the FP64 factors).  ipiv returns LAPACK (1-based absolute)
   pivots of the factorization used.  work / swork are accepted for the ABI
   and not used: the workspace lives on the device. */
int dsgesv_(int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b, int* ldb, double* x,
            int* ldx, double* work, float* swork, int* iter, int* info) {
  int st, i;
  double *dA = NULL, *dB = NULL, *dX = NULL;
  int* dp = NULL;
  (void)work;
  (void)swork;
  *info = 0;
  *iter = 0;
  if (*n < 0)
    *info = -1;
  else if (*nrhs < 0)
    *info = -2;
  else if (*lda < imax(1, *n))
    *info = -4;
  else if (*ldb < imax(1, *n))
    *info = -7;
  else if (*ldx < imax(1, *n))
    *info = -9;
  if (*info != 0) {
    xerbla("DSGESV", -*info);
    return 0;
  }
  if (*n == 0) return 0;
  st = fla_hip_malloc((void**)&dA, sizeof(double) * (size_t)*n * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dB, sizeof(double) * (size_t)*n * (size_t)imax(1, *nrhs));
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dX, sizeof(double) * (size_t)*n * (size_t)imax(1, *nrhs));
  if (st == FLA_HIP_OK) st = fla_hip_malloc((void**)&dp, sizeof(int) * (size_t)*n);
  if (st == FLA_HIP_OK) st = fla_hip_dmatrix_h2d(*n, *n, a, *lda, dA, *n);
  if (st == FLA_HIP_OK && *nrhs > 0) st = fla_hip_dmatrix_h2d(*n, *nrhs, b, *ldb, dB, *n);
  if (st == FLA_HIP_OK)
    st = fla_hip_dsgesv(*n, *nrhs, dA, *n, dp, dB, *n, dX, *n, NULL, iter, info);
  if (st == FLA_HIP_OK && *nrhs > 0) st = fla_hip_dmatrix_d2h(*n, *nrhs, dX, *n, x, *ldx);
  if (st == FLA_HIP_OK && *iter < 0) st = fla_hip_dmatrix_d2h(*n, *n, dA, *n, a, *lda);
  if (st == FLA_HIP_OK && *nrhs > 0) {
    /* FLAME-native (0-based relative) -> LAPACK (1-based absolute) */
    st = fla_hip_memcpy_d2h(ipiv, dp, sizeof(int) * (size_t)*n);
    if (st == FLA_HIP_OK)
      for (i = 0; i < *n; ++i) ipiv[i] += i + 1;
  }
  fla_hip_free(dA);
  fla_hip_free(dB);
  fla_hip_free(dX);
  fla_hip_free(dp);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: dsgesv_ failed: %s\n", fla_hip_error_string(st));
    *info = -100;
  }
  return 0;
}

/* dtrtri_: src/map/lapack2flamec/FLA_trtri.c:33-71 after dtrtri_check
 * (src/map/common/lapackcheck/dtrtri_check.c:3-61): argument checks, then,
 * for a non-unit diagonal, the first exact zero on the diagonal is returned
 * as info = i (1-based) and nothing is computed; else FLA_Trinv, info = 0. */
int dtrtri_(char* uplo, char* diag, int* n, double* a, int* lda, int* info) {
  FLA_Obj A;
  FLA_Error init_result;
  int u = up(*uplo), d = up(*diag), i;
  *info = 0;
  if (u != 'U' && u != 'L')
    *info = -1;
  else if (d != 'N' && d != 'U')
    *info = -2;
  else if (*n < 0)
    *info = -3;
  else if (*lda < imax(1, *n))
    *info = -5;
  if (*info != 0) {
    xerbla("DTRTRI", -*info);
    return 0;
  }
  if (*n == 0) return 0;
  if (d == 'N')
    for (i = 1; i <= *n; ++i)
      if (a[(size_t)(i - 1) + (size_t)(i - 1) * (size_t)*lda] == 0.0) {
        *info = i;
        return 0;
      }
  FLA_Init_safe(&init_result);
  FLA_Obj_create_without_buffer(FLA_DOUBLE, (dim_t)*n, (dim_t)*n, &A);
  FLA_Obj_attach_buffer(a, 1, (dim_t)*lda, &A);
  FLA_Trinv(u == 'L' ? FLA_LOWER_TRIANGULAR : FLA_UPPER_TRIANGULAR,
            d == 'N' ? FLA_NONUNIT_DIAG : FLA_UNIT_DIAG, A);
  FLA_Obj_free_without_buffer(&A);
  FLA_Finalize_safe(init_result);
  *info = 0;
  return 0;
}

/* dlauum_ (FLA_lauum.c, dlauum_check.c: -1 uplo, -2 n, -4 lda) and dpotri_
 * (FLA_potri.c:32-56, dpotri_check.c: the same checks): FLA_Ttmm, and
 * FLA_Trinv (non-unit) then FLA_Ttmm, on the factor in place; info = 0. */
static int inv_body(const char* name, int potri, char* uplo, int* n, double* a, int* lda,
                    int* info) {
  FLA_Obj A;
  FLA_Error init_result;
  int u = up(*uplo);
  FLA_Uplo ufla;
  *info = 0;
  if (u != 'U' && u != 'L')
    *info = -1;
  else if (*n < 0)
    *info = -2;
  else if (*lda < imax(1, *n))
    *info = -4;
  if (*info != 0) {
    xerbla(name, -*info);
    return 0;
  }
  if (*n == 0) return 0;
  ufla = u == 'L' ? FLA_LOWER_TRIANGULAR : FLA_UPPER_TRIANGULAR;
  FLA_Init_safe(&init_result);
  FLA_Obj_create_without_buffer(FLA_DOUBLE, (dim_t)*n, (dim_t)*n, &A);
  FLA_Obj_attach_buffer(a, 1, (dim_t)*lda, &A);
  if (potri) FLA_Trinv(ufla, FLA_NONUNIT_DIAG, A);
  FLA_Ttmm(ufla, A);
  FLA_Obj_free_without_buffer(&A);
  FLA_Finalize_safe(init_result);
  return 0;
}

int dlauum_(char* uplo, int* n, double* a, int* lda, int* info) {
  return inv_body("DLAUUM", 0, uplo, n, a, lda, info);
}

int dpotri_(char* uplo, int* n, double* a, int* lda, int* info) {
  return inv_body("DPOTRI", 1, uplo, n, a, lda, info);
}
