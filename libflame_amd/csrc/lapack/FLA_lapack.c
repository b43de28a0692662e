/*
 * FLA_lapack.c -- libflame_amd's FLAME-level hot-path entry points.
 *
 * Each function keeps the reference's signature, parameter checks (at error
 * level >= FLA_MIN_ERROR_CHECKING, aborting like FLA_Check_error_code) and
 * return conventions, then hands device buffers to the HIP layer (FLA_hip.h).
 *
 *   FLA_Chol          src/lapack/dec/chol/front/flamec/FLA_Chol.c:16-29
 *                     checks: src/base/flamec/check/lapack/FLA_Chol_check.c:13
 *   FLA_LU_piv        src/lapack/dec/lu/piv/front/flamec/FLA_LU_piv.c:16-33
 *                     checks: src/base/flamec/check/lapack/FLA_LU_piv_check.c:13
 *   FLA_Apply_pivots  src/lapack/util/app/piv/front/flamec/FLA_Apply_pivots.c:15-24
 *                     (the reference calls no check; neither does this one)
 *   FLA_Chol_solve    src/lapack/dec/chol/front/flamec/FLA_Chol_solve.c:13-38
 *   FLA_LU_piv_solve  src/lapack/dec/lu/piv/front/flamec/FLA_LU_piv_solve.c:13-30
 *   FLA_Shift_pivots_to src/base/flamec/util/lapack/FLA_Shift_pivots_to.c:13-42
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "FLAME.h"
#include "FLA_hip.h"
#include "fla_internal.h"

#define CHECK_LEVEL_MIN (FLA_Check_error_level() >= FLA_MIN_ERROR_CHECKING)

static FLA_Error gpu_status(int st, const char* what) {
  if (st == FLA_HIP_OK) return FLA_SUCCESS;
  fprintf(stderr, "libflame_amd: %s failed: %s\n", what, fla_hip_error_string(st));
  return FLA_GPU_ERROR;
}

/* ---- device views of objects ------------------------------------------ */
FLA_Error fla_devmat_acquire(FLA_Obj A, int copy_in, fla_devmat* d) {
  memset(d, 0, sizeof(*d));
  d->obj = A;
  if (FLA_Obj_elemtype(A) == FLA_MATRIX) {
    /* Hierarchical view: a flat column-major device copy of its scalar
       extent, gathered from the leaves in one tile-copy launch and scattered
       back on release. */
    int st;
    FLA_Error e;
    d->m = (int)FLASH_Obj_scalar_length(A);
    d->n = (int)FLASH_Obj_scalar_width(A);
    d->elem = (int)fla_datatype_size(FLA_Obj_datatype(A));
    if (d->m == 0 || d->n == 0) return FLA_SUCCESS;
    st = fla_hip_malloc(&d->dev, (size_t)d->m * (size_t)d->n * (size_t)d->elem);
    if (st != FLA_HIP_OK) return gpu_status(st, "flat copy of a hierarchical object");
    d->ld = d->m;
    d->staged = 2;
    if (copy_in && (e = fla_hier_copy(A, 0, 0, d->m, d->n, d->dev, d->ld, 0)) != FLA_SUCCESS) {
      fla_hip_free(d->dev);
      d->dev = NULL;
      d->staged = 0;
      return e;
    }
    return FLA_SUCCESS;
  }
  d->m = (int)A.m;
  d->n = (int)A.n;
  d->elem = (int)FLA_Obj_elem_size(A);
  if (A.m == 0 || A.n == 0) return FLA_SUCCESS;
  if (FLA_Obj_is_device(A)) {
    if (A.base->rs != 1 && !(A.n == 1)) {
      fprintf(stderr, "libflame_amd: device objects must be column-major (rs == 1)\n");
      return FLA_FAILURE;
    }
    d->dev = FLA_Obj_buffer_at_view(A);
    d->ld = (A.n == 1 && A.base->rs != 1) ? 1 : (int)A.base->cs;
    if (A.n == 1 && A.base->rs != 1) {
      fprintf(stderr, "libflame_amd: device vectors must have unit stride\n");
      return FLA_FAILURE;
    }
    if (d->ld < d->m) d->ld = d->m;
    return FLA_SUCCESS;
  }
  /* Host buffer: stage a packed column-major device copy. */
  {
    size_t bytes = (size_t)d->m * (size_t)d->n * (size_t)d->elem;
    int st = fla_hip_malloc(&d->dev, bytes);
    if (st != FLA_HIP_OK) return gpu_status(st, "staging allocation");
    d->ld = d->m;
    d->staged = 1;
    if (copy_in) {
      const char* src = (const char*)FLA_Obj_buffer_at_view(A);
      if (A.base->rs == 1) {
        st = fla_hip_memcpy2d(d->dev, (size_t)d->m * d->elem, src, (size_t)A.base->cs * d->elem,
                              (size_t)d->m * d->elem, (size_t)d->n, 0);
      } else {
        char* tmp = (char*)malloc(bytes);
        int i, j;
        if (!tmp) return FLA_FAILURE;
        for (j = 0; j < d->n; ++j)
          for (i = 0; i < d->m; ++i)
            memcpy(tmp + ((size_t)j * d->m + i) * d->elem,
                   src + ((size_t)i * A.base->rs + (size_t)j * A.base->cs) * d->elem, d->elem);
        st = fla_hip_memcpy_h2d(d->dev, tmp, bytes);
        free(tmp);
      }
      if (st != FLA_HIP_OK) return gpu_status(st, "staging copy");
    }
  }
  return FLA_SUCCESS;
}

FLA_Error fla_devmat_release(fla_devmat* d, int copy_out) {
  FLA_Error r = FLA_SUCCESS;
  if (!d->staged) return FLA_SUCCESS;
  if (d->staged == 2) {
    /* every call that wrote the flat copy finished on the library's stream
       before the (synchronous) scatter reads it */
    if (copy_out && d->m > 0 && d->n > 0) {
      if (fla_hip_synchronize() != FLA_HIP_OK) r = FLA_GPU_ERROR;
      if (r == FLA_SUCCESS) r = fla_hier_copy(d->obj, 0, 0, d->m, d->n, d->dev, d->ld, 1);
    }
    fla_hip_free(d->dev);
    d->dev = NULL;
    d->staged = 0;
    return r;
  }
  if (copy_out && d->m > 0 && d->n > 0) {
    FLA_Obj A = d->obj;
    char* dst = (char*)FLA_Obj_buffer_at_view(A);
    int st;
    if (A.base->rs == 1) {
      st = fla_hip_memcpy2d(dst, (size_t)A.base->cs * d->elem, d->dev, (size_t)d->m * d->elem,
                            (size_t)d->m * d->elem, (size_t)d->n, 1);
    } else {
      size_t bytes = (size_t)d->m * (size_t)d->n * (size_t)d->elem;
      char* tmp = (char*)malloc(bytes);
      int i, j;
      if (!tmp) return FLA_FAILURE;
      st = fla_hip_memcpy_d2h(tmp, d->dev, bytes);
      for (j = 0; j < d->n; ++j)
        for (i = 0; i < d->m; ++i)
          memcpy(dst + ((size_t)i * A.base->rs + (size_t)j * A.base->cs) * d->elem,
                 tmp + ((size_t)j * d->m + i) * d->elem, d->elem);
      free(tmp);
    }
    if (st != FLA_HIP_OK) r = gpu_status(st, "staging copy-back");
  }
  fla_hip_free(d->dev);
  d->dev = NULL;
  d->staged = 0;
  return r;
}

static FLA_Error require_double(FLA_Obj A) {
  if (FLA_Obj_datatype(A) != FLA_DOUBLE) {
    fprintf(stderr, "libflame_amd: the GPU path computes in FP64; got datatype %d\n",
            FLA_Obj_datatype(A));
    return FLA_UNSUPPORTED_DATATYPE;
  }
  return FLA_SUCCESS;
}

/* ---- FLA_Chol ------------------------------------------------------------ */
FLA_Error FLA_Chol(FLA_Uplo uplo, FLA_Obj A) {
  fla_devmat dA;
  FLA_Error e;
  int info = FLA_SUCCESS, st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_square(A));
  }
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  st = fla_hip_dpotrf(uplo, dA.m, (double*)dA.dev, dA.ld, &info);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Chol");
  if (e != FLA_SUCCESS) return e;
  return info;
}

/* ---- FLA_LU_nopiv / FLA_LU_nopiv_solve ------------------------------------
 * src/lapack/dec/lu/nopiv/front/flamec/FLA_LU_nopiv.c:16-36 (checks
 * FLA_LU_nopiv_check.c: floating, nonconstant; at error level >= MIN the
 * result is FLA_LU_find_zero_on_diagonal: the first exact zero on U's
 * diagonal, else FLA_SUCCESS) and FLA_LU_nopiv_solve.c:13-32 (copy B to X,
 * unit-lower then non-unit-upper triangular solves). */
FLA_Error FLA_LU_nopiv(FLA_Obj A) {
  fla_devmat dA;
  FLA_Error e;
  int info = FLA_SUCCESS, st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
  }
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  st = fla_hip_dgetrf_nopiv(dA.m, dA.n, (double*)dA.dev, dA.ld, &info);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_LU_nopiv");
  if (e != FLA_SUCCESS) return e;
  return CHECK_LEVEL_MIN ? info : FLA_SUCCESS;
}

FLA_Error FLA_LU_nopiv_solve(FLA_Obj A, FLA_Obj B, FLA_Obj X) {
  fla_devmat dA, dX;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, B));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, X));
    FLA_Check_error_code(fla_check_square(A));
    FLA_Check_error_code(fla_check_matrix_matrix_dims(FLA_NO_TRANSPOSE, FLA_NO_TRANSPOSE, A, X, B));
  }
  if (FLA_Obj_is_identical(B, X) == FALSE) {
    e = FLA_Copy(B, X);
    if (e != FLA_SUCCESS) return e;
  }
  if (FLA_Obj_has_zero_dim(X)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(X, 1, &dX)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_UNIT_DIAG, dX.m, dX.n,
                     1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
  if (st == FLA_HIP_OK)
    st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG, dX.m,
                       dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dA, 0);
  e = fla_devmat_release(&dX, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_LU_nopiv_solve");
  return e == FLA_SUCCESS ? FLA_SUCCESS : e;
}

/* ---- FLA_Trinv / FLA_Ttmm / FLA_SPDinv ------------------------------------
 * FLA_Trinv: src/lapack/inv/tri/front/flamec/FLA_Trinv.c:16-46 (checks
 * FLA_Trinv_check.c:13-33; FLA_NONUNIT_DIAG selects the non-unit variants,
 * anything else the unit ones, FLA_Trinv_internal.c; always FLA_SUCCESS).
 * FLA_Ttmm: src/lapack/misc/ttmm/front/flamec/FLA_Ttmm.c (always FLA_SUCCESS).
 * FLA_SPDinv: FLA_SPDinv_internal.c:22-39 -- Chol, then (a failed Cholesky is
 * FLA_CHOL_FAILED_MATRIX_NOT_SPD through FLA_Check_error_code at error level
 * >= MIN) Trinv (non-unit) and Ttmm; returns FLA_SUCCESS. */
static FLA_Error square_checks(FLA_Uplo uplo, FLA_Obj A) {
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_square(A));
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Trinv(FLA_Uplo uplo, FLA_Diag diag, FLA_Obj A) {
  fla_devmat dA;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) FLA_Check_error_code(fla_check_valid_diag(diag));
  square_checks(uplo, A);
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  st = fla_hip_dtrtri(uplo, diag == FLA_NONUNIT_DIAG ? FLA_NONUNIT_DIAG : FLA_UNIT_DIAG, dA.m,
                      (double*)dA.dev, dA.ld);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Trinv");
  if (e != FLA_SUCCESS) return e;
  return FLA_SUCCESS;
}

FLA_Error FLA_Ttmm(FLA_Uplo uplo, FLA_Obj A) {
  fla_devmat dA;
  FLA_Error e;
  int st;
  square_checks(uplo, A);
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  st = fla_hip_dlauum(uplo, dA.m, (double*)dA.dev, dA.ld);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Ttmm");
  if (e != FLA_SUCCESS) return e;
  return FLA_SUCCESS;
}

FLA_Error FLA_SPDinv(FLA_Uplo uplo, FLA_Obj A) {
  fla_devmat dA;
  FLA_Error e;
  int info = FLA_SUCCESS, st;
  square_checks(uplo, A);
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  st = fla_hip_dpotrf(uplo, dA.m, (double*)dA.dev, dA.ld, &info);
  if (st == FLA_HIP_OK && info != FLA_SUCCESS && CHECK_LEVEL_MIN) {
    fla_devmat_release(&dA, 1);
    FLA_Check_error_code(FLA_CHOL_FAILED_MATRIX_NOT_SPD);
  }
  if (st == FLA_HIP_OK) st = fla_hip_dtrtri(uplo, FLA_NONUNIT_DIAG, dA.m, (double*)dA.dev, dA.ld);
  if (st == FLA_HIP_OK) st = fla_hip_dlauum(uplo, dA.m, (double*)dA.dev, dA.ld);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_SPDinv");
  if (e != FLA_SUCCESS) return e;
  return FLA_SUCCESS;
}

/* ---- FLA_LU_piv ---------------------------------------------------------- */
FLA_Error FLA_LU_piv(FLA_Obj A, FLA_Obj p) {
  fla_devmat dA, dp;
  FLA_Error e;
  int info = FLA_SUCCESS, st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_int_object(p));
    FLA_Check_error_code(fla_check_col_vector(p));
    FLA_Check_error_code(fla_check_vector_dim_min(p, FLA_Obj_min_dim(A)));
  }
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(p, 0, &dp)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  st = fla_hip_dgetrf(dA.m, dA.n, (double*)dA.dev, dA.ld, (int*)dp.dev, &info);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  {
    /* only min(m,n) pivots are produced: copy back that many */
    int pm = dp.m;
    const int k = (int)FLASH_Obj_scalar_min_dim(A);
    if (dp.staged) dp.m = k < pm ? k : pm;
    if (dp.staged == 1 && dp.obj.m > (dim_t)dp.m) dp.obj.m = (dim_t)dp.m;
    fla_devmat_release(&dp, st == FLA_HIP_OK);
  }
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_LU_piv");
  if (e != FLA_SUCCESS) return e;
  return info;
}

/* ---- FLA_Apply_pivots ---------------------------------------------------- */
FLA_Error FLA_Apply_pivots(FLA_Side side, FLA_Trans trans, FLA_Obj p, FLA_Obj A) {
  fla_devmat dA, dp;
  FLA_Error e;
  int st;
  /* Only (LEFT, NO_TRANSPOSE) exists in the reference's dispatch
     (src/lapack/util/app/piv/front/flamec/FLA_Apply_pivots_internal.c:13-59). */
  if (side != FLA_LEFT || trans != FLA_NO_TRANSPOSE) {
    fprintf(stderr, "libflame_amd: FLA_Apply_pivots supports FLA_LEFT, FLA_NO_TRANSPOSE\n");
    return FLA_FAILURE;
  }
  if (FLA_Obj_has_zero_dim(A) || FLASH_Obj_scalar_vector_dim(p) == 0) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(p, 1, &dp)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  st = fla_hip_dlaswp_rows(dA.m, dA.n, (double*)dA.dev, dA.ld, 0,
                           (int)FLASH_Obj_scalar_vector_dim(p) - 1, (const int*)dp.dev, 1);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  fla_devmat_release(&dp, 0);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Apply_pivots");
  return e == FLA_SUCCESS ? FLA_SUCCESS : e;
}

/* ---- FLA_Shift_pivots_to ------------------------------------------------- */
FLA_Error FLA_Shift_pivots_to(FLA_Pivot_type ptype, FLA_Obj p) {
  int m_p = (int)FLA_Obj_length(p), n_p = (int)FLA_Obj_width(p);
  if (CHECK_LEVEL_MIN) {
    if (ptype != FLA_NATIVE_PIVOTS && ptype != FLA_LAPACK_PIVOTS)
      FLA_Check_error_code(FLA_INVALID_PIVOT_TYPE);
    FLA_Check_error_code(fla_check_int_object(p));
  }
  if (m_p < 1 || n_p < 1) return FLA_SUCCESS;
  if (FLA_Obj_is_device(p)) {
    int st = fla_hip_shift_pivots(ptype, m_p, (int*)FLA_Obj_buffer_at_view(p));
    return gpu_status(st, "FLA_Shift_pivots_to");
  } else {
    int* b = (int*)FLA_Obj_buffer_at_view(p);
    int i;
    long rs = (long)p.base->rs;
    for (i = 0; i < m_p; ++i) b[i * rs] += (ptype == FLA_LAPACK_PIVOTS) ? (i + 1) : -(i + 1);
  }
  return FLA_SUCCESS;
}

/* ---- level-3 externals (FLA_*_external over the HIP kernels) ------------- */
FLA_Error FLA_Trsm_external(FLA_Side side, FLA_Uplo uplo, FLA_Trans trans, FLA_Diag diag,
                            FLA_Obj alpha, FLA_Obj A, FLA_Obj B) {
  fla_devmat dA, dB;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_side(side));
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(fla_check_valid_trans(trans));
    FLA_Check_error_code(fla_check_valid_diag(diag));
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, B));
    FLA_Check_error_code(fla_check_square(A));
  }
  if (FLA_Obj_has_zero_dim(B)) return FLA_SUCCESS;
  if ((e = require_double(B)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(B, 1, &dB)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  st = fla_hip_dtrsm(side, uplo, trans, diag, dB.m, dB.n, fla_obj_double_value(alpha),
                     (const double*)dA.dev, dA.ld, (double*)dB.dev, dB.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dA, 0);
  e = fla_devmat_release(&dB, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Trsm_external");
  return e;
}

static FLA_Error syrk_common(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A,
                             FLA_Obj beta, FLA_Obj C) {
  fla_devmat dA, dC;
  FLA_Error e;
  int st, k;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(fla_check_valid_trans(trans));
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, C));
    FLA_Check_error_code(fla_check_square(C));
  }
  if (FLA_Obj_has_zero_dim(C)) return FLA_SUCCESS;
  if ((e = require_double(C)) != FLA_SUCCESS) return e;
  k = (trans == FLA_NO_TRANSPOSE) ? (int)A.n : (int)A.m;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(C, 1, &dC)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  st = fla_hip_dsyrk(uplo, trans, dC.m, k, fla_obj_double_value(alpha), (const double*)dA.dev,
                     dA.ld > 0 ? dA.ld : 1, fla_obj_double_value(beta), (double*)dC.dev, dC.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dA, 0);
  e = fla_devmat_release(&dC, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Syrk_external");
  return e;
}

FLA_Error FLA_Herk_external(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A,
                            FLA_Obj beta, FLA_Obj C) {
  return syrk_common(uplo, trans, alpha, A, beta, C);
}

FLA_Error FLA_Syrk_external(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A,
                            FLA_Obj beta, FLA_Obj C) {
  return syrk_common(uplo, trans, alpha, A, beta, C);
}

FLA_Error FLA_Gemm_external(FLA_Trans transa, FLA_Trans transb, FLA_Obj alpha, FLA_Obj A,
                            FLA_Obj B, FLA_Obj beta, FLA_Obj C) {
  fla_devmat dA, dB, dC;
  FLA_Error e;
  int st, k;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_trans(transa));
    FLA_Check_error_code(fla_check_valid_trans(transb));
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, B));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, C));
    FLA_Check_error_code(fla_check_matrix_matrix_dims(transa, transb, A, B, C));
  }
  if (FLA_Obj_has_zero_dim(C)) return FLA_SUCCESS;
  if ((e = require_double(C)) != FLA_SUCCESS) return e;
  k = (transa == FLA_NO_TRANSPOSE) ? (int)A.n : (int)A.m;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(B, 1, &dB)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  if ((e = fla_devmat_acquire(C, 1, &dC)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    fla_devmat_release(&dB, 0);
    return e;
  }
  st = fla_hip_dgemm(transa, transb, dC.m, dC.n, k, fla_obj_double_value(alpha),
                     (const double*)dA.dev, dA.ld > 0 ? dA.ld : 1, (const double*)dB.dev,
                     dB.ld > 0 ? dB.ld : 1, fla_obj_double_value(beta), (double*)dC.dev, dC.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dA, 0);
  fla_devmat_release(&dB, 0);
  e = fla_devmat_release(&dC, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Gemm_external");
  return e;
}

FLA_Error FLA_Copy(FLA_Obj A, FLA_Obj B) {
  fla_devmat dA, dB;
  FLA_Error e;
  int st;
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  /* scalar extents, so that flat and hierarchical objects mix */
  if (FLASH_Obj_scalar_length(A) != FLASH_Obj_scalar_length(B) ||
      FLASH_Obj_scalar_width(A) != FLASH_Obj_scalar_width(B) ||
      fla_datatype_size(FLA_Obj_datatype(A)) != fla_datatype_size(FLA_Obj_datatype(B)))
    return FLA_Check_error_code(FLA_NONCONFORMAL_DIMENSIONS);
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(B, 0, &dB)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  st = fla_hip_memcpy2d(dB.dev, (size_t)dB.ld * dB.elem, dA.dev, (size_t)dA.ld * dA.elem,
                        (size_t)dA.m * dA.elem, (size_t)dA.n, 2);
  fla_devmat_release(&dA, 0);
  e = fla_devmat_release(&dB, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Copy");
  return e;
}

/* FLA_Set for the values used on this path (alpha must be FLA_ZERO-like or a
   constant); integer objects (pivot vectors) and doubles are supported. */
FLA_Error FLA_Set(FLA_Obj alpha, FLA_Obj A) {
  size_t i, n;
  int st;
  if (FLA_Obj_has_zero_dim(A)) return FLA_SUCCESS;
  n = (size_t)A.m * (size_t)A.n;
  if (A.base->rs != 1 || (A.n > 1 && A.base->cs != A.m)) {
    fprintf(stderr, "libflame_amd: FLA_Set supports contiguous objects only\n");
    return FLA_FAILURE;
  }
  if (FLA_Obj_datatype(A) == FLA_INT) {
    int v = (int)fla_obj_double_value(alpha);
    int* tmp = (int*)malloc(n * sizeof(int));
    for (i = 0; i < n; ++i) tmp[i] = v;
    if (FLA_Obj_is_device(A))
      st = fla_hip_memcpy_h2d(FLA_Obj_buffer_at_view(A), tmp, n * sizeof(int));
    else {
      memcpy(FLA_Obj_buffer_at_view(A), tmp, n * sizeof(int));
      st = FLA_HIP_OK;
    }
    free(tmp);
  } else if (FLA_Obj_datatype(A) == FLA_DOUBLE) {
    double v = fla_obj_double_value(alpha);
    double* tmp = (double*)malloc(n * sizeof(double));
    for (i = 0; i < n; ++i) tmp[i] = v;
    if (FLA_Obj_is_device(A))
      st = fla_hip_memcpy_h2d(FLA_Obj_buffer_at_view(A), tmp, n * sizeof(double));
    else {
      memcpy(FLA_Obj_buffer_at_view(A), tmp, n * sizeof(double));
      st = FLA_HIP_OK;
    }
    free(tmp);
  } else {
    return FLA_UNSUPPORTED_DATATYPE;
  }
  return gpu_status(st, "FLA_Set");
}

/* ---- solves -------------------------------------------------------------- */
FLA_Error FLA_Chol_solve(FLA_Uplo uplo, FLA_Obj A, FLA_Obj B, FLA_Obj X) {
  fla_devmat dA, dX;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, B));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, X));
    FLA_Check_error_code(fla_check_square(A));
    FLA_Check_error_code(fla_check_matrix_matrix_dims(FLA_NO_TRANSPOSE, FLA_NO_TRANSPOSE, A, X, B));
  }
  if (FLA_Obj_is_identical(B, X) == FALSE) {
    e = FLA_Copy(B, X);
    if (e != FLA_SUCCESS) return e;
  }
  if (FLA_Obj_has_zero_dim(X)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(X, 1, &dX)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  if (uplo == FLA_LOWER_TRIANGULAR) {
    st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG, dX.m,
                       dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
    if (st == FLA_HIP_OK)
      st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_TRANSPOSE, FLA_NONUNIT_DIAG, dX.m,
                         dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
  } else {
    st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_TRANSPOSE, FLA_NONUNIT_DIAG, dX.m,
                       dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
    if (st == FLA_HIP_OK)
      st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG,
                         dX.m, dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
  }
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dA, 0);
  e = fla_devmat_release(&dX, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Chol_solve");
  return e == FLA_SUCCESS ? FLA_SUCCESS : e;
}

FLA_Error FLA_LU_piv_solve(FLA_Obj A, FLA_Obj p, FLA_Obj B, FLA_Obj X) {
  fla_devmat dA, dX, dp;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, B));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, X));
    FLA_Check_error_code(fla_check_int_object(p));
    FLA_Check_error_code(fla_check_square(A));
    FLA_Check_error_code(fla_check_col_vector(p));
    FLA_Check_error_code(fla_check_vector_dim_min(p, FLA_Obj_min_dim(A)));
    FLA_Check_error_code(fla_check_matrix_matrix_dims(FLA_NO_TRANSPOSE, FLA_NO_TRANSPOSE, A, X, B));
  }
  if (FLA_Obj_is_identical(B, X) == FALSE) {
    e = FLA_Copy(B, X);
    if (e != FLA_SUCCESS) return e;
  }
  if (FLA_Obj_has_zero_dim(X)) return FLA_SUCCESS;
  if ((e = require_double(A)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(X, 1, &dX)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  if ((e = fla_devmat_acquire(p, 1, &dp)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    fla_devmat_release(&dX, 0);
    return e;
  }
  st = fla_hip_dlaswp_rows(dX.m, dX.n, (double*)dX.dev, dX.ld, 0,
                           (int)FLASH_Obj_scalar_vector_dim(p) - 1, (const int*)dp.dev, 1);
  if (st == FLA_HIP_OK)
    st = fla_hip_dtrsm(FLA_LEFT, FLA_LOWER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_UNIT_DIAG, dX.m,
                       dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
  if (st == FLA_HIP_OK)
    st = fla_hip_dtrsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG, dX.m,
                       dX.n, 1.0, (const double*)dA.dev, dA.ld, (double*)dX.dev, dX.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dA, 0);
  fla_devmat_release(&dp, 0);
  e = fla_devmat_release(&dX, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_LU_piv_solve");
  return e == FLA_SUCCESS ? FLA_SUCCESS : e;
}

/* Reverse-mode adjoint of FLA_Chol (no counterpart in the reference):
   Abar := (G + G^T)/2, G = L^-T Phi(L^T Lbar) L^-1, for the lower factor L
   of A = L L^T.  Only FLA_LOWER_TRIANGULAR is supported. */
FLA_Error FLA_Chol_adjoint(FLA_Uplo uplo, FLA_Obj L, FLA_Obj Lbar, FLA_Obj Abar) {
  fla_devmat dL, dLb, dA;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(uplo == FLA_LOWER_TRIANGULAR ? FLA_SUCCESS : FLA_INVALID_UPLO);
    FLA_Check_error_code(fla_check_floating_object(L));
    FLA_Check_error_code(fla_check_identical_object_datatype(L, Lbar));
    FLA_Check_error_code(fla_check_identical_object_datatype(L, Abar));
    FLA_Check_error_code(fla_check_square(L));
    FLA_Check_error_code(fla_check_conformal_dims(FLA_NO_TRANSPOSE, L, Lbar));
    FLA_Check_error_code(fla_check_conformal_dims(FLA_NO_TRANSPOSE, L, Abar));
  }
  if (FLA_Obj_has_zero_dim(L)) return FLA_SUCCESS;
  if ((e = require_double(L)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(L, 1, &dL)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(Lbar, 1, &dLb)) != FLA_SUCCESS) {
    fla_devmat_release(&dL, 0);
    return e;
  }
  if ((e = fla_devmat_acquire(Abar, 0, &dA)) != FLA_SUCCESS) {
    fla_devmat_release(&dL, 0);
    fla_devmat_release(&dLb, 0);
    return e;
  }
  st = fla_hip_dchol_adjoint(dL.m, (const double*)dL.dev, dL.ld, (const double*)dLb.dev, dLb.ld,
                             (double*)dA.dev, dA.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  fla_devmat_release(&dL, 0);
  fla_devmat_release(&dLb, 0);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Chol_adjoint");
  return e;
}

/* Reverse-mode adjoint of FLA_Chol_solve (lower): given the solution X of
   A X = B and Xbar, Bbar := A^-1 Xbar and Abar -= (Bbar X^T + X Bbar^T)/2. */
FLA_Error FLA_Chol_solve_adjoint(FLA_Uplo uplo, FLA_Obj L, FLA_Obj X, FLA_Obj Xbar, FLA_Obj Bbar,
                                 FLA_Obj Abar) {
  fla_devmat dL, dX, dXb, dBb, dA;
  FLA_Error e;
  int st;
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_valid_uplo(uplo));
    FLA_Check_error_code(uplo == FLA_LOWER_TRIANGULAR ? FLA_SUCCESS : FLA_INVALID_UPLO);
    FLA_Check_error_code(fla_check_floating_object(L));
    FLA_Check_error_code(fla_check_square(L));
    FLA_Check_error_code(fla_check_identical_object_datatype(L, X));
    FLA_Check_error_code(fla_check_conformal_dims(FLA_NO_TRANSPOSE, X, Xbar));
    FLA_Check_error_code(fla_check_conformal_dims(FLA_NO_TRANSPOSE, X, Bbar));
    FLA_Check_error_code(fla_check_conformal_dims(FLA_NO_TRANSPOSE, L, Abar));
  }
  if (FLA_Obj_has_zero_dim(X)) return FLA_SUCCESS;
  if ((e = require_double(L)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(L, 1, &dL)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(X, 1, &dX)) != FLA_SUCCESS) goto out1;
  if ((e = fla_devmat_acquire(Xbar, 1, &dXb)) != FLA_SUCCESS) goto out2;
  if ((e = fla_devmat_acquire(Bbar, 0, &dBb)) != FLA_SUCCESS) goto out3;
  if ((e = fla_devmat_acquire(Abar, 1, &dA)) != FLA_SUCCESS) goto out4;
  st = fla_hip_dchol_solve_adjoint(dL.m, dX.n, (const double*)dL.dev, dL.ld,
                                   (const double*)dX.dev, dX.ld, (const double*)dXb.dev, dXb.ld,
                                   (double*)dBb.dev, dBb.ld, (double*)dA.dev, dA.ld);
  if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  fla_devmat_release(&dBb, st == FLA_HIP_OK);
  fla_devmat_release(&dXb, 0);
  fla_devmat_release(&dX, 0);
  fla_devmat_release(&dL, 0);
  if (st != FLA_HIP_OK) return gpu_status(st, "FLA_Chol_solve_adjoint");
  return e;
out4:
  fla_devmat_release(&dBb, 0);
out3:
  fla_devmat_release(&dXb, 0);
out2:
  fla_devmat_release(&dX, 0);
out1:
  fla_devmat_release(&dL, 0);
  return e;
}
