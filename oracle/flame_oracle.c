/*
 * flame_oracle.c -- CPU restatement of the libflame hot path (TEST INFRASTRUCTURE).
 *
 * This file is the checker for libflame_amd: only tests/, the smoke entry and
 * the CPU-baseline leg of bench.py load it.  It restates, in plain C and in
 * the reference's own arithmetic order, what the reference computes:
 *
 *   FLA_Chol        fla_chol_cntl2: blk var3 (b=128) -> blk var3 (b=32) -> opt var2
 *                   src/base/flamec/control/init/lapack/fla/FLA_Chol_cntl_init.c:25-70
 *   FLA_LU_piv      fla_lu_piv_cntl2: blk var5 (b=128) -> blk var5 (b=16) -> opt var4
 *                   src/base/flamec/control/init/lapack/fla/FLA_LU_piv_cntl_init.c:26-88
 *   leaf BLAS       the builtin f2c reference BLAS, src/flablas/f2c/ (ddot.c, ...)
 *   FLA_Trinv / FLA_Ttmm / FLA_SPDinv   the leaf sweeps (opt var3 / opt var2)
 *                   over the whole matrix; see or_trinv / or_ttmm below
 *
 * Compile with -ffp-contract=off so that every a*b+c rounds twice, as the
 * reference's gcc build on x86-64 (no FMA in the baseline ISA) does.
 */
#include "flame_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(size_t)(i) + (size_t)(j) * (size_t)lda]
#define B_(i, j) b[(size_t)(i) + (size_t)(j) * (size_t)ldb]
#define C_(i, j) c[(size_t)(i) + (size_t)(j) * (size_t)ldc]

static int up(char c) { return (c >= 'a' && c <= 'z') ? c - 32 : c; }
static int imin(int a, int b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* BLAS level 1                                                              */
/* ------------------------------------------------------------------------ */

/* src/flablas/f2c/ddot.c:3-90: strided loop, or the unit-stride loop that
   handles n%5 leading terms and then adds 5 products per step, left to right. */
double or_ddot(int n, const double* x, int incx, const double* y, int incy)
{
    double t = 0.0;
    int i;
    if (n <= 0) return 0.0;
    if (incx == 1 && incy == 1) {
        int m = n % 5;
        for (i = 0; i < m; ++i) t += x[i] * y[i];
        for (i = m; i < n; i += 5)
            t = t + x[i] * y[i] + x[i + 1] * y[i + 1] + x[i + 2] * y[i + 2] +
                x[i + 3] * y[i + 3] + x[i + 4] * y[i + 4];
        return t;
    }
    {
        long ix = incx < 0 ? (long)(-n + 1) * incx : 0;
        long iy = incy < 0 ? (long)(-n + 1) * incy : 0;
        for (i = 0; i < n; ++i) {
            t += x[ix] * y[iy];
            ix += incx;
            iy += incy;
        }
    }
    return t;
}

/* src/flablas/f2c/dscal.c: x := da*x (element order is irrelevant). */
void or_dscal(int n, double da, double* x, int incx)
{
    int i;
    if (n <= 0 || incx <= 0) return;
    for (i = 0; i < n; ++i) x[(size_t)i * incx] = da * x[(size_t)i * incx];
}

/* src/flablas/f2c/idamax.c:3-72: first index of max |x|; an element replaces
   the running max only if NOT (|x_i| <= max), so a NaN is taken when met. */
int or_idamax(int n, const double* x, int incx)
{
    int i, r;
    double dmax;
    if (n < 1 || incx <= 0) return 0;
    r = 1;
    if (n == 1) return 1;
    dmax = fabs(x[0]);
    for (i = 1; i < n; ++i) {
        double v = fabs(x[(size_t)i * incx]);
        if (v <= dmax) continue;
        r = i + 1;
        dmax = v;
    }
    return r;
}

/* ------------------------------------------------------------------------ */
/* BLAS level 2: src/flablas/f2c/dgemv.c                                    */
/* ------------------------------------------------------------------------ */
void or_dgemv(char trans, int m, int n, double alpha, const double* a, int lda,
              const double* x, int incx, double beta, double* y, int incy)
{
    int i, j, lenx, leny;
    long kx, ky;
    int notrans = up(trans) == 'N';
    if (m == 0 || n == 0 || (alpha == 0.0 && beta == 1.0)) return;
    lenx = notrans ? n : m;
    leny = notrans ? m : n;
    kx = incx > 0 ? 0 : -(long)(lenx - 1) * incx;
    ky = incy > 0 ? 0 : -(long)(leny - 1) * incy;
    if (beta != 1.0) {
        long iy = ky;
        for (i = 0; i < leny; ++i) {
            y[iy] = (beta == 0.0) ? 0.0 : beta * y[iy];
            iy += incy;
        }
    }
    if (alpha == 0.0) return;
    if (notrans) {
        long jx = kx;
        for (j = 0; j < n; ++j) {
            if (x[jx] != 0.0) {
                double temp = alpha * x[jx];
                long iy = ky;
                for (i = 0; i < m; ++i) {
                    y[iy] += temp * A_(i, j);
                    iy += incy;
                }
            }
            jx += incx;
        }
    } else {
        long jy = ky;
        for (j = 0; j < n; ++j) {
            double temp = 0.0;
            long ix = kx;
            for (i = 0; i < m; ++i) {
                temp += A_(i, j) * x[ix];
                ix += incx;
            }
            y[jy] += alpha * temp;
            jy += incy;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* BLAS level 3                                                              */
/* ------------------------------------------------------------------------ */

/* src/flablas/f2c/dsyrk.c: C := alpha*A*A' + beta*C ('N') or alpha*A'*A + beta*C. */
void or_dsyrk(char uplo, char trans, int n, int k, double alpha, const double* a,
              int lda, double beta, double* c, int ldc)
{
    int i, j, l;
    int upper = up(uplo) == 'U';
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
    if (alpha == 0.0) {
        for (j = 0; j < n; ++j) {
            int i0 = upper ? 0 : j, i1 = upper ? j : n - 1;
            for (i = i0; i <= i1; ++i) C_(i, j) = (beta == 0.0) ? 0.0 : beta * C_(i, j);
        }
        return;
    }
    if (up(trans) == 'N') {
        /* columns are independent: the OpenMP build splits them over threads
           (each column keeps the serial loop order, so results are identical) */
#pragma omp parallel for private(i, l) schedule(dynamic, 8) if ((double)n * n * k > 4e6)
        for (j = 0; j < n; ++j) {
            int i0 = upper ? 0 : j, i1 = upper ? j : n - 1;
            if (beta == 0.0) {
                for (i = i0; i <= i1; ++i) C_(i, j) = 0.0;
            } else if (beta != 1.0) {
                for (i = i0; i <= i1; ++i) C_(i, j) = beta * C_(i, j);
            }
            for (l = 0; l < k; ++l) {
                if (A_(j, l) != 0.0) {
                    double temp = alpha * A_(j, l);
                    for (i = i0; i <= i1; ++i) C_(i, j) += temp * A_(i, l);
                }
            }
        }
    } else {
        for (j = 0; j < n; ++j) {
            int i0 = upper ? 0 : j, i1 = upper ? j : n - 1;
            for (i = i0; i <= i1; ++i) {
                double temp = 0.0;
                for (l = 0; l < k; ++l) temp += A_(l, i) * A_(l, j);
                if (beta == 0.0)
                    C_(i, j) = alpha * temp;
                else
                    C_(i, j) = alpha * temp + beta * C_(i, j);
            }
        }
    }
}

/* src/flablas/f2c/dgemm.c: C := alpha*op(A)*op(B) + beta*C. */
void or_dgemm(char transa, char transb, int m, int n, int k, double alpha,
              const double* a, int lda, const double* b, int ldb, double beta,
              double* c, int ldc)
{
    int i, j, l;
    int nota = up(transa) == 'N', notb = up(transb) == 'N';
    if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
    if (alpha == 0.0) {
        for (j = 0; j < n; ++j)
            for (i = 0; i < m; ++i) C_(i, j) = (beta == 0.0) ? 0.0 : beta * C_(i, j);
        return;
    }
    if (notb) {
        if (nota) {
#pragma omp parallel for private(i, l) schedule(dynamic, 8) if ((double)m * n * k > 4e6)
            for (j = 0; j < n; ++j) {
                if (beta == 0.0) {
                    for (i = 0; i < m; ++i) C_(i, j) = 0.0;
                } else if (beta != 1.0) {
                    for (i = 0; i < m; ++i) C_(i, j) = beta * C_(i, j);
                }
                for (l = 0; l < k; ++l) {
                    if (B_(l, j) != 0.0) {
                        double temp = alpha * B_(l, j);
                        for (i = 0; i < m; ++i) C_(i, j) += temp * A_(i, l);
                    }
                }
            }
        } else {
#pragma omp parallel for private(i, l) schedule(dynamic, 8) if ((double)m * n * k > 4e6)
            for (j = 0; j < n; ++j)
                for (i = 0; i < m; ++i) {
                    double temp = 0.0;
                    for (l = 0; l < k; ++l) temp += A_(l, i) * B_(l, j);
                    if (beta == 0.0)
                        C_(i, j) = alpha * temp;
                    else
                        C_(i, j) = alpha * temp + beta * C_(i, j);
                }
        }
    } else {
        if (nota) {
#pragma omp parallel for private(i, l) schedule(dynamic, 8) if ((double)m * n * k > 4e6)
            for (j = 0; j < n; ++j) {
                if (beta == 0.0) {
                    for (i = 0; i < m; ++i) C_(i, j) = 0.0;
                } else if (beta != 1.0) {
                    for (i = 0; i < m; ++i) C_(i, j) = beta * C_(i, j);
                }
                for (l = 0; l < k; ++l) {
                    if (B_(j, l) != 0.0) {
                        double temp = alpha * B_(j, l);
                        for (i = 0; i < m; ++i) C_(i, j) += temp * A_(i, l);
                    }
                }
            }
        } else {
            for (j = 0; j < n; ++j)
                for (i = 0; i < m; ++i) {
                    double temp = 0.0;
                    for (l = 0; l < k; ++l) temp += A_(l, i) * B_(j, l);
                    if (beta == 0.0)
                        C_(i, j) = alpha * temp;
                    else
                        C_(i, j) = alpha * temp + beta * C_(i, j);
                }
        }
    }
}

/* src/flablas/f2c/dtrsm.c: op(A)*X = alpha*B (left) or X*op(A) = alpha*B (right).
   Left solves are independent per column of B, right solves per row: the
   OpenMP build runs the unchanged loop nest on disjoint column (row) ranges,
   so every element sees the serial sequence of operations. */
static void dtrsm_part(int lside, int upper, int notr, int nounit, int m, int n,
                       double alpha, const double* a, int lda, double* b, int ldb);

void or_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
              double alpha, const double* a, int lda, double* b, int ldb)
{
    int i, j;
    int lside = up(side) == 'L';
    int upper = up(uplo) == 'U';
    int notr = up(transa) == 'N';
    int nounit = up(diag) == 'N';
    int parts = 1, q;
    if (m == 0 || n == 0) return;
    if (alpha == 0.0) {
        for (j = 0; j < n; ++j)
            for (i = 0; i < m; ++i) B_(i, j) = 0.0;
        return;
    }
#ifdef _OPENMP
    if ((double)m * m * n > 4e6 && (lside ? n : m) >= 16) parts = 64;
#endif
#pragma omp parallel for schedule(dynamic, 1) if (parts > 1)
    for (q = 0; q < parts; ++q) {
        if (lside) {
            int j0 = (int)((long)n * q / parts), j1 = (int)((long)n * (q + 1) / parts);
            if (j1 > j0)
                dtrsm_part(lside, upper, notr, nounit, m, j1 - j0, alpha, a, lda,
                           b + (size_t)j0 * ldb, ldb);
        } else {
            int i0 = (int)((long)m * q / parts), i1 = (int)((long)m * (q + 1) / parts);
            if (i1 > i0)
                dtrsm_part(lside, upper, notr, nounit, i1 - i0, n, alpha, a, lda, b + i0, ldb);
        }
    }
}

static void dtrsm_part(int lside, int upper, int notr, int nounit, int m, int n,
                       double alpha, const double* a, int lda, double* b, int ldb)
{
    int i, j, k;
    if (lside) {
        if (notr) {
            if (upper) {
                for (j = 0; j < n; ++j) {
                    if (alpha != 1.0)
                        for (i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                    for (k = m - 1; k >= 0; --k) {
                        if (B_(k, j) != 0.0) {
                            if (nounit) B_(k, j) /= A_(k, k);
                            for (i = 0; i < k; ++i) B_(i, j) -= B_(k, j) * A_(i, k);
                        }
                    }
                }
            } else {
                for (j = 0; j < n; ++j) {
                    if (alpha != 1.0)
                        for (i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                    for (k = 0; k < m; ++k) {
                        if (B_(k, j) != 0.0) {
                            if (nounit) B_(k, j) /= A_(k, k);
                            for (i = k + 1; i < m; ++i) B_(i, j) -= B_(k, j) * A_(i, k);
                        }
                    }
                }
            }
        } else {
            if (upper) {
                for (j = 0; j < n; ++j)
                    for (i = 0; i < m; ++i) {
                        double temp = alpha * B_(i, j);
                        for (k = 0; k < i; ++k) temp -= A_(k, i) * B_(k, j);
                        if (nounit) temp /= A_(i, i);
                        B_(i, j) = temp;
                    }
            } else {
                for (j = 0; j < n; ++j)
                    for (i = m - 1; i >= 0; --i) {
                        double temp = alpha * B_(i, j);
                        for (k = i + 1; k < m; ++k) temp -= A_(k, i) * B_(k, j);
                        if (nounit) temp /= A_(i, i);
                        B_(i, j) = temp;
                    }
            }
        }
    } else {
        if (notr) {
            if (upper) {
                for (j = 0; j < n; ++j) {
                    if (alpha != 1.0)
                        for (i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                    for (k = 0; k < j; ++k)
                        if (A_(k, j) != 0.0)
                            for (i = 0; i < m; ++i) B_(i, j) -= A_(k, j) * B_(i, k);
                    if (nounit) {
                        double temp = 1.0 / A_(j, j);
                        for (i = 0; i < m; ++i) B_(i, j) = temp * B_(i, j);
                    }
                }
            } else {
                for (j = n - 1; j >= 0; --j) {
                    if (alpha != 1.0)
                        for (i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
                    for (k = j + 1; k < n; ++k)
                        if (A_(k, j) != 0.0)
                            for (i = 0; i < m; ++i) B_(i, j) -= A_(k, j) * B_(i, k);
                    if (nounit) {
                        double temp = 1.0 / A_(j, j);
                        for (i = 0; i < m; ++i) B_(i, j) = temp * B_(i, j);
                    }
                }
            }
        } else {
            if (upper) {
                for (k = n - 1; k >= 0; --k) {
                    if (nounit) {
                        double temp = 1.0 / A_(k, k);
                        for (i = 0; i < m; ++i) B_(i, k) = temp * B_(i, k);
                    }
                    for (j = 0; j < k; ++j)
                        if (A_(j, k) != 0.0) {
                            double temp = A_(j, k);
                            for (i = 0; i < m; ++i) B_(i, j) -= temp * B_(i, k);
                        }
                    if (alpha != 1.0)
                        for (i = 0; i < m; ++i) B_(i, k) = alpha * B_(i, k);
                }
            } else {
                for (k = 0; k < n; ++k) {
                    if (nounit) {
                        double temp = 1.0 / A_(k, k);
                        for (i = 0; i < m; ++i) B_(i, k) = temp * B_(i, k);
                    }
                    for (j = k + 1; j < n; ++j)
                        if (A_(j, k) != 0.0) {
                            double temp = A_(j, k);
                            for (i = 0; i < m; ++i) B_(i, j) -= temp * B_(i, k);
                        }
                    if (alpha != 1.0)
                        for (i = 0; i < m; ++i) B_(i, k) = alpha * B_(i, k);
                }
            }
        }
    }
}

/* src/map/common/lapacksrc/f2c/dlaswp.c:103-213 with 1-based ipiv and k1/k2.
   Row interchanges are exact moves, so the column blocking by 32 does not
   change any value; the interchange order per column (forward for incx > 0,
   backward for incx < 0) is kept. */
void or_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx)
{
    int i, j, i1, i2, inc;
    long ix0;
    if (incx > 0) {
        ix0 = k1;
        i1 = k1;
        i2 = k2;
        inc = 1;
    } else if (incx < 0) {
        ix0 = (long)(1 - k2) * incx + 1;
        i1 = k2;
        i2 = k1;
        inc = -1;
    } else {
        return;
    }
    for (j = 0; j < n; ++j) {
        long ix = ix0;
        for (i = i1; inc > 0 ? i <= i2 : i >= i2; i += inc) {
            int ip = ipiv[ix - 1];
            if (ip != i) {
                double t = A_(i - 1, j);
                A_(i - 1, j) = A_(ip - 1, j);
                A_(ip - 1, j) = t;
            }
            ix += incx;
        }
    }
}

/* Threads the level-3 routines use (1 unless built with OpenMP). */
#ifdef _OPENMP
#include <omp.h>
int or_threads(void) { return omp_get_max_threads(); }
#else
int or_threads(void) { return 1; }
#endif

/* ------------------------------------------------------------------------ */
/* FLAME layer                                                               */
/* ------------------------------------------------------------------------ */

/* bl1_dinvscalv (src/base/flamec/blis/1/bl1_invscalv.c:26-37): skip when
   alpha == 1, else scale by the reciprocal 1.0/alpha through dscal. */
static void invscalv(int n, double alpha, double* x, int incx)
{
    double inv;
    if (alpha == 1.0) return;
    inv = 1.0 / alpha;
    or_dscal(n, inv, x, incx);
}

/* bl1_ddots (src/base/flamec/blis/1/bl1_dots.c:26-37): rho := beta*rho + alpha*dot. */
static void dots(int n, double alpha, const double* x, int incx, const double* y,
                 int incy, double beta, double* rho)
{
    double d = or_ddot(n, x, incx, y, incy);
    *rho = beta * (*rho) + alpha * d;
}

/* bl1_dsqrte (src/base/flamec/include/blis_macro_defs.h:134-136). */
static int sqrte(double* alpha)
{
    if (*alpha <= 0.0 || isnan(*alpha)) return 0;
    *alpha = sqrt(*alpha);
    return 1;
}

/* FLA_Chol_l_opd_var2 (src/lapack/dec/chol/l/flamec/FLA_Chol_l_opt_var2.c:136-193). */
static int chol_l_opd_var2(int n, double* a, int lda)
{
    int i;
    for (i = 0; i < n; ++i) {
        double* a10t = &A_(i, 0);
        double* A20 = &A_(i + 1, 0);
        double* alpha11 = &A_(i, i);
        double* a21 = &A_(i + 1, i);
        int ahead = n - i - 1, behind = i;
        dots(behind, -1.0, a10t, lda, a10t, lda, 1.0, alpha11);
        if (ahead > 0 && behind > 0)
            or_dgemv('N', ahead, behind, -1.0, A20, lda, a10t, lda, 1.0, a21, 1);
        if (!sqrte(alpha11)) return behind;
        invscalv(ahead, *alpha11, a21, 1);
    }
    return OR_SUCCESS;
}

/* FLA_Chol_u_opd_var2 (src/lapack/dec/chol/u/flamec/FLA_Chol_u_opt_var2.c:136-193). */
static int chol_u_opd_var2(int n, double* a, int lda)
{
    int i;
    for (i = 0; i < n; ++i) {
        double* a01 = &A_(0, i);
        double* alpha11 = &A_(i, i);
        double* A02 = &A_(0, i + 1);
        double* a12t = &A_(i, i + 1);
        int ahead = n - i - 1, behind = i;
        dots(behind, -1.0, a01, 1, a01, 1, 1.0, alpha11);
        if (ahead > 0 && behind > 0)
            or_dgemv('T', behind, ahead, -1.0, A02, lda, a01, 1, 1.0, a12t, lda);
        if (!sqrte(alpha11)) return behind;
        invscalv(ahead, *alpha11, a12t, lda);
    }
    return OR_SUCCESS;
}

/* FLA_Chol_l_blk_var3 / FLA_Chol_u_blk_var3
   (src/lapack/dec/chol/l/flamec/FLA_Chol_l_blk_var3.c:13-66,
    src/lapack/dec/chol/u/flamec/FLA_Chol_u_blk_var3.c:38-54), with the block
   size min(remaining, b) of FLA_Determine_blocksize (FLA_Blocksize.c:234-264). */
static int chol_rec(int uplo, int n, double* a, int lda, const int* nb, int nlev)
{
    int k;
    if (nlev <= 0)
        return uplo == OR_LOWER ? chol_l_opd_var2(n, a, lda) : chol_u_opd_var2(n, a, lda);
    for (k = 0; k < n;) {
        int b = imin(n - k, nb[0]);
        int m2 = n - k - b;
        int r = chol_rec(uplo, b, &A_(k, k), lda, nb + 1, nlev - 1);
        if (r != OR_SUCCESS) return k + r;
        if (uplo == OR_LOWER) {
            if (m2 > 0) {
                or_dtrsm('R', 'L', 'C', 'N', m2, b, 1.0, &A_(k, k), lda, &A_(k + b, k), lda);
                or_dsyrk('L', 'N', m2, b, -1.0, &A_(k + b, k), lda, 1.0, &A_(k + b, k + b), lda);
            }
        } else {
            if (m2 > 0) {
                or_dtrsm('L', 'U', 'C', 'N', b, m2, 1.0, &A_(k, k), lda, &A_(k, k + b), lda);
                or_dsyrk('U', 'C', m2, b, -1.0, &A_(k, k + b), lda, 1.0, &A_(k + b, k + b), lda);
            }
        }
        k += b;
    }
    return OR_SUCCESS;
}

int or_chol(int uplo, int n, double* a, int lda, const int* nb, int nlev)
{
    return chol_rec(uplo, n, a, lda, nb, nlev);
}

/* ---- FLA_LU_nopiv --------------------------------------------------------- */

/* Leaf FLA_LU_nopiv_opd_var5 (src/lapack/dec/lu/nopiv/vars/flamec/
   FLA_LU_nopiv_opt_var5.c:132-160): per column, a21 := a21 / alpha11
   (bl1_dinvscalv: by the reciprocal, skipped for 1), then
   A22 := A22 - a21 a12t (dger with alpha -1, columns with a zero a12t entry
   skipped). */
static void lu_nopiv_opd_var5(int m, int n, double* a, int lda)
{
    const int mn = imin(m, n);
    for (int i = 0; i < mn; ++i) {
        const double al = A_(i, i);
        if (al != 1.0) {
            const double s = 1.0 / al;
            for (int r = i + 1; r < m; ++r) A_(r, i) *= s;
        }
        for (int c = i + 1; c < n; ++c) {
            const double t = A_(i, c);
            if (t == 0.0) continue;
            const double tm = -1.0 * t;
            for (int r = i + 1; r < m; ++r) A_(r, c) += A_(r, i) * tm;
        }
    }
}

/* FLA_LU_nopiv_blk_var5 (FLA_LU_nopiv_blk_var5.c:36-52), one level per
   block size: A11 := LU(A11); A12 := trilu(A11) \ A12;
   A21 := A21 / triu(A11); A22 -= A21 A12. */
static void lu_nopiv_rec(int m, int n, double* a, int lda, const int* nb, int nlev)
{
    if (nlev <= 0) {
        lu_nopiv_opd_var5(m, n, a, lda);
        return;
    }
    const int mn = imin(m, n);
    for (int k = 0; k < mn;) {
        const int b = imin(mn - k, nb[0]);
        const int nr = n - k - b, mr = m - k - b;
        lu_nopiv_rec(b, b, &A_(k, k), lda, nb + 1, nlev - 1);
        if (nr > 0) or_dtrsm('L', 'L', 'N', 'U', b, nr, 1.0, &A_(k, k), lda, &A_(k, k + b), lda);
        if (mr > 0) or_dtrsm('R', 'U', 'N', 'N', mr, b, 1.0, &A_(k, k), lda, &A_(k + b, k), lda);
        if (mr > 0 && nr > 0)
            or_dgemm('N', 'N', mr, nr, b, -1.0, &A_(k + b, k), lda, &A_(k, k + b), lda, 1.0,
                     &A_(k + b, k + b), lda);
        k += b;
    }
}

/* FLA_LU_nopiv (src/lapack/dec/lu/nopiv/front/flamec/FLA_LU_nopiv.c:16-36)
   with fla_lu_nopiv_cntl2 (FLA_LU_nopiv_cntl_init.c:28-70: var5 at b, var5
   at b/4, leaf opt var5), returning, as at error level >= MIN, the first
   exact zero on the diagonal (FLA_LU_find_zero_on_diagonal.c:25-46) or
   OR_SUCCESS. */
int or_lu_nopiv(int m, int n, double* a, int lda, const int* nb, int nlev)
{
    lu_nopiv_rec(m, n, a, lda, nb, nlev);
    for (int i = 0; i < imin(m, n); ++i)
        if (A_(i, i) == 0.0) return i;
    return OR_SUCCESS;
}

/* ---- FLA_Trinv / FLA_Ttmm / FLA_SPDinv ---------------------------------- */

/* Exchanges the strictly upper and strictly lower triangles: an upper-stored
   triangle becomes the lower-stored transpose (and back). */
static void tri_swap(int n, double* a, int lda)
{
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) {
            const double t = A_(i, j);
            A_(i, j) = A_(j, i);
            A_(j, i) = t;
        }
}

/* FLA_Trinv on the lower triangle, one unblocked sweep in the order of the
   reference's leaf FLA_Trinv_ln_opd_var3
   (src/lapack/inv/tri/ln/flamec/FLA_Trinv_ln_opt_var3.c:137-176) and, for a
   unit diagonal, FLA_Trinv_lu_opd_var3 (.../lu/flamec/FLA_Trinv_lu_opt_var3.c:
   124-150); per row i:  a21 := a21 / (-alpha11) (unit: -a21),
   A20 += a21 a10t (dger, columns with a zero a10t entry skipped),
   a10t := a10t / alpha11, alpha11 := 1 / alpha11.  The reference runs this
   leaf on 128-blocks inside FLA_Trinv_ln_blk_var3; the sweep here is the same
   inverse with another rounding order.  Upper storage: the transposed
   problem, FLA_Trinv(U) = FLA_Trinv(U')'. */
void or_trinv(int uplo, int diag, int n, double* a, int lda)
{
    if (uplo == OR_UPPER) tri_swap(n, a, lda);
    for (int i = 0; i < n; ++i) {
        const double al = A_(i, i);
        if (diag != 500) {
            const double s = 1.0 / (-al);
            for (int r = i + 1; r < n; ++r) A_(r, i) *= s;
        } else {
            for (int r = i + 1; r < n; ++r) A_(r, i) = -A_(r, i);
        }
        for (int c = 0; c < i; ++c) {
            const double t = A_(i, c);
            if (t == 0.0) continue;
            for (int r = i + 1; r < n; ++r) A_(r, c) += A_(r, i) * t;
        }
        if (diag != 500) {
            const double s = 1.0 / al;
            for (int c = 0; c < i; ++c) A_(i, c) *= s;
            A_(i, i) = 1.0 / al;
        }
    }
    if (uplo == OR_UPPER) tri_swap(n, a, lda);
}

/* FLA_Ttmm on the lower triangle (A := L' L), in the order of the reference's
   leaf FLA_Ttmm_l_opd_var2 (src/lapack/misc/ttmm/l/flamec/
   FLA_Ttmm_l_opt_var2.c:136-176); per row i:  a10t := alpha11 a10t,
   a10t += A20' a21 (dgemv 'T': one dot per column), alpha11 := alpha11^2,
   alpha11 += a21' a21.  Upper storage: A := U U' through the transposed
   problem. */
void or_ttmm(int uplo, int n, double* a, int lda)
{
    if (uplo == OR_UPPER) tri_swap(n, a, lda);
    for (int i = 0; i < n; ++i) {
        const double al = A_(i, i);
        for (int c = 0; c < i; ++c) A_(i, c) *= al;
        for (int c = 0; c < i; ++c) {
            double t = 0.0;
            for (int r = i + 1; r < n; ++r) t += A_(r, c) * A_(r, i);
            A_(i, c) += t;
        }
        double d = 0.0;
        for (int r = i + 1; r < n; ++r) d += A_(r, i) * A_(r, i);
        A_(i, i) = al * al + d;
    }
    if (uplo == OR_UPPER) tri_swap(n, a, lda);
}

/* FLA_SPDinv_internal (src/lapack/inv/spd/front/flamec/FLA_SPDinv_internal.c:
   22-39): Cholesky (the oracle's FLA_Chol), then FLA_Trinv (non-unit) and
   FLA_Ttmm.  Returns the Cholesky's result; on failure nothing else runs. */
int or_spdinv(int uplo, int n, double* a, int lda, const int* nb, int nlev)
{
    const int r = or_chol(uplo, n, a, lda, nb, nlev);
    if (r != OR_SUCCESS) return r;
    or_trinv(uplo, 501, n, a, lda);
    or_ttmm(uplo, n, a, lda);
    return OR_SUCCESS;
}

/* FLA_Apply_pivots_ln_opd_var1 (src/lapack/util/app/piv/ln/flamec/
   FLA_Apply_pivots_ln_opt_var1.c:274-330), column-major branch: for every
   column, swap row i with row i + p[i] for i = k1..k2 (relative pivots). */
void or_apply_pivots_ln(int ncols, double* a, int lda, int k1, int k2,
                        const int* p, int incp)
{
    int i, j;
    int ib, ie, ii, pinc;
    if (incp > 0) {
        ib = k1; ie = k2 + 1; ii = 1; pinc = incp;
    } else {
        ib = k2; ie = k1 - 1; ii = -1; pinc = -incp;
    }
    for (j = 0; j < ncols; ++j) {
        for (i = ib; i != ie; i += ii) {
            int pi = p[(size_t)i * pinc] + i;
            double t = A_(pi, j);
            A_(pi, j) = A_(i, j);
            A_(i, j) = t;
        }
    }
}

/* FLA_LU_piv_opd_var4 (src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_opt_var4.c:218-338):
   Crout LU with partial pivoting; a null pivot records its index (first one
   wins), skips the interchange and the scaling, and the sweep continues. */
static int lu_opd_var4(int m, int n, double* a, int lda, int* p)
{
    int r_val = OR_SUCCESS;
    int mn = imin(m, n);
    int i;
    for (i = 0; i < mn; ++i) {
        double* a10t = &A_(i, 0);
        double* A20 = &A_(i + 1, 0);
        double* a01 = &A_(0, i);
        double* alpha11 = &A_(i, i);
        double* a21 = &A_(i + 1, i);
        double* A02 = &A_(0, i + 1);
        double* a12t = &A_(i, i + 1);
        int* pi1 = &p[i];
        int m_ahead = m - i - 1, n_ahead = n - i - 1, behind = i;
        int null_pivot;

        dots(behind, -1.0, a10t, lda, a01, 1, 1.0, alpha11);
        if (m_ahead > 0 && behind > 0)
            or_dgemv('N', m_ahead, behind, -1.0, A20, lda, a01, 1, 1.0, a21, 1);
        *pi1 = or_idamax(m_ahead + 1, alpha11, 1) - 1;
        null_pivot = (alpha11[*pi1] == 0.0);
        if (null_pivot) {
            if (r_val == OR_SUCCESS) r_val = i;
        } else {
            or_apply_pivots_ln(1, alpha11, lda, 0, 0, pi1, 1);
            or_apply_pivots_ln(behind, a10t, lda, 0, 0, pi1, 1);
            or_apply_pivots_ln(n_ahead, a12t, lda, 0, 0, pi1, 1);
        }
        if (behind > 0 && n_ahead > 0)
            or_dgemv('T', behind, n_ahead, -1.0, A02, lda, a10t, lda, 1.0, a12t, lda);
        if (!null_pivot) invscalv(m_ahead, *alpha11, a21, 1);
    }
    return r_val;
}

/* FLA_LU_piv_blk_var5 (src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_blk_var5.c:13-114). */
static int lu_rec(int m, int n, double* a, int lda, int* p, const int* nb, int nlev)
{
    int r_val = OR_SUCCESS;
    int k;
    if (nlev <= 0) return lu_opd_var4(m, n, a, lda, p);
    for (k = 0; k < m && k < n;) {
        int b = imin(imin(m - k, n - k), nb[0]);
        int r_sub = lu_rec(m - k, b, &A_(k, k), lda, p + k, nb + 1, nlev - 1);
        if (r_val == OR_SUCCESS && r_sub >= 0) r_val = k + r_sub;
        /* AB0 = [A10; A20], AB2 = [A12; A22]: rows k.., pivots relative. */
        or_apply_pivots_ln(k, &A_(k, 0), lda, 0, b - 1, p + k, 1);
        or_apply_pivots_ln(n - k - b, &A_(k, k + b), lda, 0, b - 1, p + k, 1);
        if (n - k - b > 0) {
            or_dtrsm('L', 'L', 'N', 'U', b, n - k - b, 1.0, &A_(k, k), lda, &A_(k, k + b), lda);
            if (m - k - b > 0)
                or_dgemm('N', 'N', m - k - b, n - k - b, b, -1.0, &A_(k + b, k), lda,
                         &A_(k, k + b), lda, 1.0, &A_(k + b, k + b), lda);
        }
        k += b;
    }
    return r_val;
}

int or_lu_piv(int m, int n, double* a, int lda, int* p, const int* nb, int nlev)
{
    return lu_rec(m, n, a, lda, p, nb, nlev);
}

/* ------------------------------------------------------------------------ */
/* LU with incremental pivoting over b x b tiles: FLASH_LU_incpiv            */
/* (src/lapack/dec/lu/incpiv/front/flamec/FLASH_LU_incpiv.c:13-39 ->         */
/* FLASH_LU_incpiv_opt1.c:13-42 -> zaux/flamec/flash/FLASH_LU_incpiv_var2.c, */
/* FLASH_Trsm_piv.c, FLASH_SA_LU.c, FLASH_SA_FS.c; leaves                    */
/* zaux/flamec/fla/FLA_SA_LU_blk.c, FLA_SA_LU_unb.c, FLA_SA_FS_blk.c,        */
/* FLA_SA_Apply_pivots.c; the diagonal tile's LU is FLA_LU_piv_task with     */
/* fla_lu_piv_cntl_leaf = the unblocked Crout opt_var4,                       */
/* FLA_LU_piv_copy_task.c / FLA_LU_piv_task.c).  The operations on a tile    */
/* row (j > k) act column by column, so they run on the whole row at once.   */
/* ------------------------------------------------------------------------ */

/* FLA_SA_Apply_pivots.c:13-105: row i of C (mC rows) <-> row p[i] - (mC - i)
   of E, for every i with p[i] != 0, in order. */
static void sa_apply_pivots(int mC, int nC, double* c, int ldc, double* e, int lde, int mp,
                            const int* p)
{
    int i, j;
    for (i = 0; i < mp; ++i) {
        if (p[i] != 0) {
            int r = p[i] - (mC - i);
            for (j = 0; j < nC; ++j) {
                double t = C_(i, j);
                C_(i, j) = e[(size_t)r + (size_t)j * (size_t)lde];
                e[(size_t)r + (size_t)j * (size_t)lde] = t;
            }
        }
    }
}

/* FLA_SA_LU_unb.c (double branch): LU of [U; D] with U (mU x mU) upper
   triangular; L := triu(U) collects the rows that move, p[j] = ipiv + mU - j
   when row ipiv of D replaces row j (|L(j,j)| < |D(ipiv,j)|), else 0. */
static void sa_lu_unb(int mU, double* u, int ldu, int mD, double* d, int ldd, int* p, double* l,
                      int ldl)
{
    int i, j, c;
    for (c = 0; c < mU; ++c)
        for (i = 0; i < mU; ++i) l[(size_t)i + (size_t)c * ldl] = (i <= c) ? u[(size_t)i + (size_t)c * ldu] : 0.0;
    for (j = 0; j < mU; ++j) {
        int ipiv = or_idamax(mD, d + (size_t)j * ldd, 1) - 1;
        double ltmp = l[(size_t)j + (size_t)j * ldl];
        double dtmp = d[(size_t)ipiv + (size_t)j * ldd];
        if (fabs(ltmp) < fabs(dtmp)) {
            for (c = 0; c < mU; ++c) {
                double t = l[(size_t)j + (size_t)c * ldl];
                l[(size_t)j + (size_t)c * ldl] = d[(size_t)ipiv + (size_t)c * ldd];
                d[(size_t)ipiv + (size_t)c * ldd] = t;
            }
            p[j] = ipiv + mU - j;
        } else {
            p[j] = 0;
        }
        or_dscal(mD, 1.0 / l[(size_t)j + (size_t)j * ldl], d + (size_t)j * ldd, 1);
        /* dger (f2c order): D(:, c) += D(:, j) * (-L(j, c)), c > j */
        for (c = j + 1; c < mU; ++c) {
            double y = l[(size_t)j + (size_t)c * ldl];
            if (y != 0.0) {
                double temp = -1.0 * y;
                for (i = 0; i < mD; ++i)
                    d[(size_t)i + (size_t)c * ldd] += d[(size_t)i + (size_t)j * ldd] * temp;
            }
        }
        for (c = j; c < mU; ++c) u[(size_t)j + (size_t)c * ldu] = l[(size_t)j + (size_t)c * ldl];
    }
}

/* FLA_SA_LU_blk.c:13-109: blocks of nb_alg columns of [U; D] (U nU x nU, D
   mD x nU); L1_sqr is the bs x bs top-left part of L's rows s..s+bs. */
static void sa_lu_blk(int nU, double* u, int ldu, int mD, double* d, int ldd, int* p, double* l,
                      int ldl, int nb_alg)
{
    int s, bs;
    for (s = 0; s < nU; s += bs) {
        int nrest;
        bs = imin(nU - s, nb_alg);
        nrest = nU - s - bs;
        sa_lu_unb(bs, u + (size_t)s + (size_t)s * ldu, ldu, mD, d + (size_t)s * ldd, ldd, p + s,
                  l + s, ldl);
        sa_apply_pivots(bs, nrest, u + (size_t)s + (size_t)(s + bs) * ldu, ldu,
                        d + (size_t)(s + bs) * ldd, ldd, bs, p + s);
        if (nrest > 0) {
            or_dtrsm('L', 'L', 'N', 'U', bs, nrest, 1.0, l + s, ldl,
                     u + (size_t)s + (size_t)(s + bs) * ldu, ldu);
            or_dgemm('N', 'N', mD, nrest, bs, -1.0, d + (size_t)s * ldd, ldd,
                     u + (size_t)s + (size_t)(s + bs) * ldu, ldu, 1.0,
                     d + (size_t)(s + bs) * ldd, ldd);
        }
    }
}

/* FLA_SA_FS_blk.c:13-105: apply the transformations of one SA_LU (L tile
   mL x nb_alg, D mD x mL, pivots p) to C (mL x nC) over E (mD x nC). */
static void sa_fs_blk(int mL, const double* l, int ldl, int mD, const double* d, int ldd,
                      const int* p, int nC, double* c, int ldc, double* e, int lde, int nb_alg)
{
    int s, bs;
    for (s = 0; s < mL; s += bs) {
        bs = imin(mL - s, nb_alg);
        sa_apply_pivots(bs, nC, c + s, ldc, e, lde, bs, p + s);
        or_dtrsm('L', 'L', 'N', 'U', bs, nC, 1.0, l + s, ldl, c + s, ldc);
        or_dgemm('N', 'N', mD, nC, bs, -1.0, d + (size_t)s * ldd, ldd, c + s, ldc, 1.0, e, lde);
    }
}

/* FLASH_LU_incpiv_var2.c:13-114: for each diagonal tile k, its LU (Crout,
   pivots within the tile), the pivoted unit-lower solve of the tile row
   (FLASH_Trsm_piv.c), then for each tile i below, SA_LU of [A_kk; A_ik] and
   SA_FS of the tile rows k and i (FLASH_SA_LU.c, FLASH_SA_FS.c). */
int or_lu_incpiv(int m, int n, double* a, int lda, int b, int nb_alg, int* p, int ldp,
                 double* l, int ldl)
{
    int mt, nt, kt, k, i;
    if (b <= 0 || nb_alg <= 0 || m % b != 0 || n % b != 0) return -2;
    mt = m / b;
    nt = n / b;
    kt = imin(mt, nt);
    for (k = 0; k < kt; ++k) {
        double* akk = &A_(k * b, k * b);
        int* pkk = p + (size_t)k * b + (size_t)k * ldp;
        int nr = n - (k + 1) * b;
        (void)lu_opd_var4(b, b, akk, lda, pkk);
        if (nr > 0) {
            or_apply_pivots_ln(nr, &A_(k * b, (k + 1) * b), lda, 0, b - 1, pkk, 1);
            or_dtrsm('L', 'L', 'N', 'U', b, nr, 1.0, akk, lda, &A_(k * b, (k + 1) * b), lda);
        }
        for (i = k + 1; i < mt; ++i) {
            double* aik = &A_(i * b, k * b);
            int* pik = p + (size_t)i * b + (size_t)k * ldp;
            double* lik = l + (size_t)i * b + (size_t)k * nb_alg * ldl;
            sa_lu_blk(b, akk, lda, b, aik, lda, pik, lik, ldl, nb_alg);
            if (nr > 0)
                sa_fs_blk(b, lik, ldl, b, aik, lda, pik, nr, &A_(k * b, (k + 1) * b), lda,
                          &A_(i * b, (k + 1) * b), lda, nb_alg);
        }
    }
    return OR_SUCCESS;
}

/* FLASH_FS_incpiv.c:13-53 -> zaux/flamec/flash/FLASH_FS_incpiv_aux1.c,
   FLASH_FS_incpiv_aux2.c: per diagonal tile, its pivots and unit-lower solve
   (FLA_Trsv_external; column by column for several right-hand sides), then
   SA_FS of the tiles below on the rows of x. */
int or_fs_incpiv(int m, int n, const double* a, int lda, int b, int nb_alg, const int* p,
                 int ldp, const double* l, int ldl, int nrhs, double* x, int ldx)
{
    int mt, nt, kt, k, i;
    if (b <= 0 || nb_alg <= 0 || m % b != 0 || n % b != 0) return -2;
    mt = m / b;
    nt = n / b;
    kt = imin(mt, nt);
    for (k = 0; k < kt; ++k) {
        const double* akk = &A_(k * b, k * b);
        const int* pkk = p + (size_t)k * b + (size_t)k * ldp;
        double* xk = x + (size_t)k * b;
        or_apply_pivots_ln(nrhs, xk, ldx, 0, b - 1, pkk, 1);
        or_dtrsm('L', 'L', 'N', 'U', b, nrhs, 1.0, akk, lda, xk, ldx);
        for (i = k + 1; i < mt; ++i)
            sa_fs_blk(b, l + (size_t)i * b + (size_t)k * nb_alg * ldl, ldl, b,
                      &A_(i * b, k * b), lda, p + (size_t)i * b + (size_t)k * ldp, nrhs, xk,
                      ldx, x + (size_t)i * b, ldx, nb_alg);
    }
    return OR_SUCCESS;
}

/* FLA_Shift_pivots_to (src/base/flamec/util/lapack/FLA_Shift_pivots_to.c:13-42). */
void or_shift_pivots_to(int ptype, int m, int* p)
{
    int i;
    for (i = 0; i < m; ++i) p[i] += (ptype == OR_LAPACK_PIVOTS) ? (i + 1) : -(i + 1);
}

static void copy_mat(int m, int n, const double* s, int lds, double* d, int ldd)
{
    int j;
    for (j = 0; j < n; ++j) memcpy(d + (size_t)j * ldd, s + (size_t)j * lds, sizeof(double) * (size_t)m);
}

/* FLA_Chol_solve (src/lapack/dec/chol/front/flamec/FLA_Chol_solve.c:13-38). */
int or_chol_solve(int uplo, int n, int nrhs, const double* a, int lda,
                  const double* b, int ldb, double* x, int ldx)
{
    if (b != x) copy_mat(n, nrhs, b, ldb, x, ldx);
    if (uplo == OR_LOWER) {
        or_dtrsm('L', 'L', 'N', 'N', n, nrhs, 1.0, a, lda, x, ldx);
        or_dtrsm('L', 'L', 'C', 'N', n, nrhs, 1.0, a, lda, x, ldx);
    } else {
        or_dtrsm('L', 'U', 'C', 'N', n, nrhs, 1.0, a, lda, x, ldx);
        or_dtrsm('L', 'U', 'N', 'N', n, nrhs, 1.0, a, lda, x, ldx);
    }
    return OR_SUCCESS;
}

/* FLA_LU_piv_solve (src/lapack/dec/lu/piv/front/flamec/FLA_LU_piv_solve.c:13-30). */
int or_lu_piv_solve(int n, int nrhs, const double* a, int lda, const int* p,
                    const double* b, int ldb, double* x, int ldx)
{
    if (b != x) copy_mat(n, nrhs, b, ldb, x, ldx);
    or_apply_pivots_ln(nrhs, x, ldx, 0, n - 1, p, 1);
    or_dtrsm('L', 'L', 'N', 'U', n, nrhs, 1.0, a, lda, x, ldx);
    or_dtrsm('L', 'U', 'N', 'N', n, nrhs, 1.0, a, lda, x, ldx);
    return OR_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* lapack2flame                                                              */
/* ------------------------------------------------------------------------ */
static const int or_nb_chol[2] = {128, 32};
static const int or_nb_lu[2] = {128, 16};

/* dpotrf_ (src/map/lapack2flamec/FLA_potrf.c:38-83): argument check as
   dpotrf_check, then FLA_Chol, info = idx + 1 on failure. */
int or_dpotrf(char uplo, int n, double* a, int lda, int* info)
{
    int u = up(uplo), e;
    *info = 0;
    if (u != 'U' && u != 'L') *info = -1;
    else if (n < 0) *info = -2;
    else if (lda < (n > 1 ? n : 1)) *info = -4;
    if (*info != 0) return 0;
    if (n == 0) return 0;
    e = or_chol(u == 'L' ? OR_LOWER : OR_UPPER, n, a, lda, or_nb_chol, 2);
    *info = (e != OR_SUCCESS) ? e + 1 : 0;
    return 0;
}

/* dgetrf_ (src/map/lapack2flamec/FLA_getrf.c:40-92): p zeroed, FLA_LU_piv,
   pivots shifted to LAPACK form, info = idx + 1 on a null pivot. */
int or_dgetrf(int m, int n, double* a, int lda, int* ipiv, int* info)
{
    int mn = imin(m, n), e;
    *info = 0;
    if (m < 0) *info = -1;
    else if (n < 0) *info = -2;
    else if (lda < (m > 1 ? m : 1)) *info = -4;
    if (*info != 0) return 0;
    if (m == 0 || n == 0) return 0;
    memset(ipiv, 0, sizeof(int) * (size_t)mn);
    e = or_lu_piv(m, n, a, lda, ipiv, or_nb_lu, 2);
    or_shift_pivots_to(OR_LAPACK_PIVOTS, mn, ipiv);
    *info = (e != OR_SUCCESS) ? e + 1 : 0;
    return 0;
}

/* dpotrs_ (src/map/common/lapacksrc/f2c/dpotrs.c:102-194). */
int or_dpotrs(char uplo, int n, int nrhs, const double* a, int lda, double* b,
              int ldb, int* info)
{
    int u = up(uplo);
    *info = 0;
    if (u != 'U' && u != 'L') *info = -1;
    else if (n < 0) *info = -2;
    else if (nrhs < 0) *info = -3;
    else if (lda < (n > 1 ? n : 1)) *info = -5;
    else if (ldb < (n > 1 ? n : 1)) *info = -7;
    if (*info != 0) return 0;
    if (n == 0 || nrhs == 0) return 0;
    if (u == 'U') {
        or_dtrsm('L', 'U', 'T', 'N', n, nrhs, 1.0, a, lda, b, ldb);
        or_dtrsm('L', 'U', 'N', 'N', n, nrhs, 1.0, a, lda, b, ldb);
    } else {
        or_dtrsm('L', 'L', 'N', 'N', n, nrhs, 1.0, a, lda, b, ldb);
        or_dtrsm('L', 'L', 'T', 'N', n, nrhs, 1.0, a, lda, b, ldb);
    }
    return 0;
}

/* dgetrs_ (src/map/common/lapacksrc/f2c/dgetrs.c:115-210). */
int or_dgetrs(char trans, int n, int nrhs, const double* a, int lda,
              const int* ipiv, double* b, int ldb, int* info)
{
    int t = up(trans);
    *info = 0;
    if (t != 'N' && t != 'T' && t != 'C') *info = -1;
    else if (n < 0) *info = -2;
    else if (nrhs < 0) *info = -3;
    else if (lda < (n > 1 ? n : 1)) *info = -5;
    else if (ldb < (n > 1 ? n : 1)) *info = -8;
    if (*info != 0) return 0;
    if (n == 0 || nrhs == 0) return 0;
    if (t == 'N') {
        or_dlaswp(nrhs, b, ldb, 1, n, ipiv, 1);
        or_dtrsm('L', 'L', 'N', 'U', n, nrhs, 1.0, a, lda, b, ldb);
        or_dtrsm('L', 'U', 'N', 'N', n, nrhs, 1.0, a, lda, b, ldb);
    } else {
        or_dtrsm('L', 'U', 'T', 'N', n, nrhs, 1.0, a, lda, b, ldb);
        or_dtrsm('L', 'L', 'T', 'U', n, nrhs, 1.0, a, lda, b, ldb);
        or_dlaswp(nrhs, b, ldb, 1, n, ipiv, -1);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Input generators.  Counter-based: element idx of stream `seed` is a pure   */
/* function, so the host and the device produce identical uniforms.          */
/* ------------------------------------------------------------------------ */
static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t or_rand_u64(uint64_t seed, uint64_t idx)
{
    return mix64(mix64(seed + 0x9E3779B97F4A7C15ULL) ^ (idx * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
}

double or_uniform_pm1(uint64_t seed, uint64_t idx)
{
    /* 53 random bits -> k * 2^-52 - 1 in [-1, 1): exact in binary64. */
    uint64_t k = or_rand_u64(seed, idx) >> 11;
    return (double)k * 0x1.0p-52 - 1.0;
}

double or_normal(uint64_t seed, uint64_t idx)
{
    uint64_t k1 = or_rand_u64(seed, 2 * idx) >> 11;
    uint64_t k2 = or_rand_u64(seed, 2 * idx + 1) >> 11;
    double u1 = ((double)k1 + 1.0) * 0x1.0p-53; /* (0, 1] */
    double u2 = (double)k2 * 0x1.0p-53;         /* [0, 1) */
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void or_gen_uniform(uint64_t seed, int m, int n, double* a, int lda)
{
    int i, j;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) A_(i, j) = or_uniform_pm1(seed, (uint64_t)j * (uint64_t)m + (uint64_t)i);
}

void or_gen_normal(uint64_t seed, int m, int n, double* a, int lda)
{
    int i, j;
    for (j = 0; j < n; ++j)
        for (i = 0; i < m; ++i) A_(i, j) = or_normal(seed, (uint64_t)j * (uint64_t)m + (uint64_t)i);
}

/* A = scale*B*B^T + shift*I with full symmetric storage; the lower triangle is
   computed in a fixed order (dot products over k ascending) and mirrored. */
static void gen_gram(int n, const double* bm, double scale, double shift, double* a, int lda)
{
    int i, j, k;
    for (j = 0; j < n; ++j)
        for (i = j; i < n; ++i) {
            double s = 0.0;
            for (k = 0; k < n; ++k) s += bm[(size_t)i + (size_t)k * n] * bm[(size_t)j + (size_t)k * n];
            s = s * scale;
            if (i == j) s += shift;
            A_(i, j) = s;
            A_(j, i) = s;
        }
}

void or_gen_spd_c1(uint64_t seed, int n, double* a, int lda)
{
    double* bm = (double*)malloc(sizeof(double) * (size_t)n * n);
    or_gen_uniform(seed, n, n, bm, n);
    gen_gram(n, bm, 1.0, (double)n, a, lda);
    free(bm);
}

void or_gen_spd_c2(uint64_t seed, int n, double* a, int lda)
{
    double* bm = (double*)malloc(sizeof(double) * (size_t)n * n);
    or_gen_normal(seed, n, n, bm, n);
    gen_gram(n, bm, 1.0 / (double)n, 1.0, a, lda);
    free(bm);
}

void or_gen_lu_c2(uint64_t seed, int m, int n, double* a, int lda)
{
    int i, j;
    or_gen_uniform(seed, m, n, a, lda);
    for (i = 0; i < m; ++i) {
        double s = pow(10.0, -(double)i / (double)(m > n ? m : n));
        for (j = 0; j < n; ++j) A_(i, j) *= s;
    }
}

/* ------------------------------------------------------------------------ */
/* Error metrics (the dpot01 / dget01 ratios of the LAPACK test suite,       */
/* netlib-test/lapack-3.4.2/TESTING/LIN/dpot01.f:88, dget01.f:200-208).      */
/* ------------------------------------------------------------------------ */
static double sym_norm1(int uplo, int n, const double* a, int lda)
{
    double best = 0.0;
    int i, j;
    for (j = 0; j < n; ++j) {
        double s = 0.0;
        for (i = 0; i < n; ++i) {
            int lo = uplo == OR_LOWER ? (i >= j) : (i <= j);
            s += fabs(lo ? A_(i, j) : A_(j, i));
        }
        if (s > best) best = s;
    }
    return best;
}

double or_chol_residual(int uplo, int n, const double* a0, int lda0, const double* f, int ldf)
{
    /* R = L*L^T - A (lower) or U^T*U - A (upper), in long double. */
    double anorm = sym_norm1(uplo, n, a0, lda0);
    double rmax = 0.0;
    int i, j, k;
    const double eps = 0x1.0p-53;
    long double* col = (long double*)malloc(sizeof(long double) * (size_t)n);
    for (j = 0; j < n; ++j) {
        double s = 0.0;
        for (i = 0; i < n; ++i) {
            long double acc = 0.0L;
            int kmax = i < j ? i : j;
            for (k = 0; k <= kmax; ++k) {
                if (uplo == OR_LOWER)
                    acc += (long double)f[(size_t)i + (size_t)k * ldf] * f[(size_t)j + (size_t)k * ldf];
                else
                    acc += (long double)f[(size_t)k + (size_t)i * ldf] * f[(size_t)k + (size_t)j * ldf];
            }
            {
                int lo = uplo == OR_LOWER ? (i >= j) : (i <= j);
                double av = lo ? a0[(size_t)i + (size_t)j * lda0] : a0[(size_t)j + (size_t)i * lda0];
                col[i] = acc - av;
            }
            s += fabs((double)col[i]);
        }
        if (s > rmax) rmax = s;
    }
    free(col);
    /* dpot01.f:151-156: RESID = 1/EPS when ANORM <= 0 */
    if (anorm <= 0.0) return 1.0 / eps;
    return rmax / ((double)n * eps * anorm);
}

double or_lu_residual(int m, int n, const double* a0, int lda0, const double* f,
                      int ldf, const int* p)
{
    /* Apply the pivots to a copy of A, then compare with L*U. */
    int mn = imin(m, n), i, j, k;
    double anorm = 0.0, rmax = 0.0;
    const double eps = 0x1.0p-53;
    double* pa = (double*)malloc(sizeof(double) * (size_t)m * n);
    copy_mat(m, n, a0, lda0, pa, m);
    or_apply_pivots_ln(n, pa, m, 0, mn - 1, p, 1);
    for (j = 0; j < n; ++j) {
        double s = 0.0, r = 0.0;
        for (i = 0; i < m; ++i) s += fabs(a0[(size_t)i + (size_t)j * lda0]);
        if (s > anorm) anorm = s;
        for (i = 0; i < m; ++i) {
            long double acc = 0.0L;
            int kmax = imin(i, j);
            if (kmax > mn - 1) kmax = mn - 1;
            for (k = 0; k <= kmax; ++k) {
                long double l = (k == i) ? 1.0L : (long double)f[(size_t)i + (size_t)k * ldf];
                long double u = (long double)f[(size_t)k + (size_t)j * ldf];
                acc += l * u;
            }
            r += fabs((double)(acc - pa[(size_t)i + (size_t)j * m]));
        }
        if (r > rmax) rmax = r;
    }
    free(pa);
    /* dget01.f:209-214: RESID / N / ANORM / EPS; 1/EPS when ANORM <= 0 and
       the difference is not zero */
    if (anorm <= 0.0) return rmax == 0.0 ? 0.0 : 1.0 / eps;
    return rmax / ((double)n * eps * anorm);
}
