/*
 * flame_oracle.h -- CPU restatement of libflame's FLA_Chol / FLA_LU_piv hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation in libflame_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * All matrices are column-major with leading dimension ld, 0-based C indexing.
 * Return conventions follow libflame: FLA_SUCCESS == -1, a numerical failure
 * returns the 0-based index of the first failing column / first null pivot.
 * Pivots are FLAME-native (0-based, relative to the current row) unless a
 * function says LAPACK (1-based absolute).
 */
#ifndef FLAME_ORACLE_H
#define FLAME_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>

#define OR_SUCCESS (-1)
#define OR_LOWER 300
#define OR_UPPER 301
#define OR_NATIVE_PIVOTS 700
#define OR_LAPACK_PIVOTS 701

/* ---- netlib/f2c-order BLAS restatements (bit-for-bit loop order) ---- */
double or_ddot(int n, const double* x, int incx, const double* y, int incy);
void or_dgemv(char trans, int m, int n, double alpha, const double* a, int lda,
              const double* x, int incx, double beta, double* y, int incy);
void or_dscal(int n, double da, double* x, int incx);
int or_idamax(int n, const double* x, int incx); /* 1-based, 0 if n < 1 */
void or_dsyrk(char uplo, char trans, int n, int k, double alpha, const double* a,
              int lda, double beta, double* c, int ldc);
void or_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
              double alpha, const double* a, int lda, double* b, int ldb);
void or_dgemm(char transa, char transb, int m, int n, int k, double alpha,
              const double* a, int lda, const double* b, int ldb, double beta,
              double* c, int ldc);
void or_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx);

/* Threads used by the level-3 routines: 1 for liboracle.so; the OpenMP build
   liboracle_omp.so (same source, -fopenmp) splits independent columns / rows
   of dgemm, dsyrk and dtrsm over threads without changing any element's
   sequence of operations, so both builds return identical bits. */
int or_threads(void);

/* ---- FLAME layer ---- */
/* FLA_Chol with the fla_chol_cntl2 tree: blocked var3 at nb[0], var3 at nb[1],
   ..., leaf opt_var2.  nlev = number of blocked levels (libflame: 2, {128,32}). */
int or_chol(int uplo, int n, double* a, int lda, const int* nb, int nlev);
/* FLA_LU_piv with the fla_lu_piv_cntl2 tree: var5 at nb[0], var5 at nb[1],
   leaf Crout opt_var4 (libflame: {128,16}).  p gets min(m,n) native pivots. */
int or_lu_piv(int m, int n, double* a, int lda, int* p, const int* nb, int nlev);
/* FLA_Apply_pivots(LEFT, NO_TRANSPOSE): rows k1..k2 of the ncols-wide A. */
void or_apply_pivots_ln(int ncols, double* a, int lda, int k1, int k2,
                        const int* p, int incp);
void or_shift_pivots_to(int ptype, int m, int* p);
int or_chol_solve(int uplo, int n, int nrhs, const double* a, int lda,
                  const double* b, int ldb, double* x, int ldx);
int or_lu_piv_solve(int n, int nrhs, const double* a, int lda, const int* p,
                    const double* b, int ldb, double* x, int ldx);

/* ---- lapack2flame entry points ---- */
int or_dpotrf(char uplo, int n, double* a, int lda, int* info);
int or_dgetrf(int m, int n, double* a, int lda, int* ipiv, int* info);
int or_dpotrs(char uplo, int n, int nrhs, const double* a, int lda, double* b,
              int ldb, int* info);
int or_dgetrs(char trans, int n, int nrhs, const double* a, int lda,
              const int* ipiv, double* b, int ldb, int* info);

/* ---- input generators (shared spec with the device generators) ---- */
uint64_t or_rand_u64(uint64_t seed, uint64_t idx);
double or_uniform_pm1(uint64_t seed, uint64_t idx); /* [-1, 1) */
double or_normal(uint64_t seed, uint64_t idx);      /* N(0,1), Box-Muller */
void or_gen_uniform(uint64_t seed, int m, int n, double* a, int lda);
void or_gen_normal(uint64_t seed, int m, int n, double* a, int lda);
/* C1: A = B*B^T + n*I, B U[-1,1] (full symmetric storage). */
void or_gen_spd_c1(uint64_t seed, int n, double* a, int lda);
/* A = B*B^T/n + I, B N(0,1) (full symmetric storage). */
void or_gen_spd_c2(uint64_t seed, int n, double* a, int lda);
/* LU: U[-1,1] with row i scaled by 10^(-i/n) (well-separated pivots). */
void or_gen_lu_c2(uint64_t seed, int m, int n, double* a, int lda);

/* ---- next-row widening: LU without pivoting, inverses ---- */
/* FLA_LU_nopiv; returns the first exact zero on U's diagonal or OR_SUCCESS. */
int or_lu_nopiv(int m, int n, double* a, int lda, const int* nb, int nlev);
/* FLA_Trinv (diag 500 unit / 501 non-unit), FLA_Ttmm, FLA_SPDinv. */
void or_trinv(int uplo, int diag, int n, double* a, int lda);
void or_ttmm(int uplo, int n, double* a, int lda);
int or_spdinv(int uplo, int n, double* a, int lda, const int* nb, int nlev);

/* ---- LU with incremental pivoting over b x b tiles (FLASH_LU_incpiv) ----
   a: the flat m x n matrix, m and n multiples of b; p: the flattened
   hierarchical pivot object, (m/b)*b x (n/b) ints, tile (i,k) at rows i*b,
   column k; l: the flattened L object, (m/b)*b x (n/b)*nb_alg, tile (i,k) at
   rows i*b, columns k*nb_alg.  Tiles above the diagonal of p and tiles on or
   above the diagonal of l are not written.  Returns OR_SUCCESS, or -2 when m
   or n is not a multiple of b. */
int or_lu_incpiv(int m, int n, double* a, int lda, int b, int nb_alg, int* p, int ldp,
                 double* l, int ldl);
/* FLASH_FS_incpiv on nrhs columns of x (m rows): the incremental pivots and
   unit-lower transformations of or_lu_incpiv, in the factorization's order. */
int or_fs_incpiv(int m, int n, const double* a, int lda, int b, int nb_alg, const int* p,
                 int ldp, const double* l, int ldl, int nrhs, double* x, int ldx);

/* ---- error metrics ---- */
/* ||A - L*L^T||_1 / (n * eps * ||A||_1) using only the uplo triangle of A0. */
double or_chol_residual(int uplo, int n, const double* a0, int lda0,
                        const double* f, int ldf);
/* ||P*A - L*U||_1 / (n * eps * ||A||_1) with FLAME-native pivots. */
double or_lu_residual(int m, int n, const double* a0, int lda0, const double* f,
                      int ldf, const int* p);

#ifdef __cplusplus
}
#endif
#endif
