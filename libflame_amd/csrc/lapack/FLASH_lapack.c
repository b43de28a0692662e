/*
 * FLASH_lapack.c -- the FLASH (hierarchical storage) front-ends of the hot path.
 *
 *   FLASH_Chol              src/lapack/dec/chol/front/flamec/FLASH_Chol.c:15-35
 *   FLASH_Chol_solve        src/lapack/dec/chol/front/flamec/FLASH_Chol_solve.c:13-38
 *   FLASH_LU_piv            src/lapack/dec/lu/piv/front/flamec/FLASH_LU_piv.c:15-50
 *   FLASH_LU_piv_solve      src/lapack/dec/lu/piv/front/flamec/FLASH_LU_piv_solve.c:13-30
 *   FLASH_Apply_pivots      src/lapack/util/app/piv/front/flamec/FLASH_Apply_pivots.c:15-45
 *   FLASH_LU_nopiv          src/lapack/dec/lu/nopiv/front/flamec/FLASH_LU_nopiv.c:15-37
 *   FLASH_LU_nopiv_solve    src/lapack/dec/lu/nopiv/front/flamec/FLASH_LU_nopiv_solve.c
 *   FLASH_LU_find_zero_on_diagonal
 *                           src/lapack/dec/lu/nopiv/.../FLASH_LU_find_zero_on_diagonal.c:13-50
 *
 * The reference runs these as SuperMatrix task graphs over the blocks
 * (flash_chol_cntl / flash_lu_piv_cntl: blocked var3 / var5 stepping one
 * block at a time, leaf operations on b x b blocks).  Here the flat-copy
 * helpers of the FLAME layer gather a hierarchical object into one
 * column-major matrix in HBM (one tile-copy launch), the one-GPU look-ahead
 * factorization runs on it, and the result is scattered back (one launch).
 * Flat objects are accepted too and behave as a single leaf.
 *
 * Return values follow the reference's hierarchical conventions:
 *  * FLASH_Chol: under flash_chol_cntl, FLA_Chol_l_blk_var3 returns
 *    FLA_Obj_length(A00) + r_val (FLA_Chol_l_blk_var3.c:42-43), and
 *    FLA_Obj_length of a hierarchical view counts BLOCKS: the value is the
 *    number of block rows above the failing diagonal block at every level of
 *    the hierarchy plus the 0-based column inside the failing leaf.
 *  * FLASH_LU_piv / FLASH_LU_nopiv: at error level >= FLA_MIN_ERROR_CHECKING
 *    the scalar index of the first exact zero on U's diagonal
 *    (FLASH_LU_find_zero_on_diagonal), else FLA_SUCCESS.
 */
#include <stdio.h>
#include <stdlib.h>

#include "FLAME.h"
#include "FLA_hip.h"
#include "fla_internal.h"

#define CHECK_LEVEL_MIN (FLA_Check_error_level() >= FLA_MIN_ERROR_CHECKING)

static FLA_Obj child(FLA_Obj H, dim_t i, dim_t j) {
  const FLA_Obj* buf = (const FLA_Obj*)H.base->buffer;
  return buf[(H.offm + i) * H.base->rs + (H.offn + j) * H.base->cs];
}

/* The reference's return value for a Cholesky failure at scalar column c of
   the hierarchical view H (see the file comment). */
static FLA_Error chol_return_index(FLA_Obj H, dim_t c) {
  dim_t i, acc = 0;
  if (FLA_Obj_elemtype(H) == FLA_SCALAR) return (FLA_Error)c;
  for (i = 0; i < H.m; ++i) {
    const FLA_Obj d = child(H, i, i);
    const dim_t h = FLASH_Obj_scalar_length(d);
    if (c < acc + h) return (FLA_Error)i + chol_return_index(d, c - acc);
    acc += h;
  }
  return (FLA_Error)c;
}

/* The reference's depth restriction of the LU and pivot front-ends
   (FLASH_LU_piv.c:26-33, FLASH_Apply_pivots.c:24-31). */
static void require_depth_one(FLA_Obj A, const char* fn) {
  if (FLASH_Obj_depth(A) != 1) {
    fprintf(stderr, "libflame_amd: %s() currently only supports matrices of depth 1\n", fn);
    abort();
  }
}

FLA_Error FLASH_Chol(FLA_Uplo uplo, FLA_Obj A) {
  FLA_Error r = FLA_Chol(uplo, A);
  if (r >= 0 && FLA_Obj_elemtype(A) == FLA_MATRIX) r = chol_return_index(A, (dim_t)r);
  return r;
}

FLA_Error FLASH_Chol_solve(FLA_Uplo uplo, FLA_Obj A, FLA_Obj B, FLA_Obj X) {
  return FLA_Chol_solve(uplo, A, B, X);
}

FLA_Error FLASH_LU_piv(FLA_Obj A, FLA_Obj p) {
  FLA_Error r;
  require_depth_one(A, "FLASH_LU_piv");
  r = FLA_LU_piv(A, p);
  if (r < FLA_SUCCESS) return r; /* a library error code */
  /* the first null pivot is the first exact zero on U's diagonal */
  return CHECK_LEVEL_MIN ? r : FLA_SUCCESS;
}

FLA_Error FLASH_LU_piv_solve(FLA_Obj A, FLA_Obj p, FLA_Obj B, FLA_Obj X) {
  return FLA_LU_piv_solve(A, p, B, X);
}

FLA_Error FLASH_Apply_pivots(FLA_Side side, FLA_Trans trans, FLA_Obj p, FLA_Obj A) {
  require_depth_one(A, "FLASH_Apply_pivots");
  return FLA_Apply_pivots(side, trans, p, A);
}

FLA_Error FLASH_LU_nopiv(FLA_Obj A) {
  FLA_Error r = FLA_LU_nopiv(A);
  if (r < FLA_SUCCESS) return r;
  return CHECK_LEVEL_MIN ? r : FLA_SUCCESS;
}

FLA_Error FLASH_LU_nopiv_solve(FLA_Obj A, FLA_Obj B, FLA_Obj X) {
  return FLA_LU_nopiv_solve(A, B, X);
}

/* The scalar index of the first exact zero on the diagonal, else FLA_SUCCESS. */
FLA_Error FLASH_LU_find_zero_on_diagonal(FLA_Obj A) {
  fla_devmat dA;
  FLA_Error e;
  int k, i, st;
  double* diag;
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  k = dA.m < dA.n ? dA.m : dA.n;
  if (k == 0) {
    fla_devmat_release(&dA, 0);
    return FLA_SUCCESS;
  }
  diag = (double*)malloc(sizeof(double) * (size_t)k);
  if (diag == NULL) {
    fla_devmat_release(&dA, 0);
    return FLA_FAILURE;
  }
  /* the diagonal is a 1 x k matrix with stride ld + 1 */
  st = fla_hip_memcpy2d(diag, sizeof(double), dA.dev, (size_t)(dA.ld + 1) * sizeof(double),
                        sizeof(double), (size_t)k, 1);
  fla_devmat_release(&dA, 0);
  if (st != FLA_HIP_OK) {
    free(diag);
    fprintf(stderr, "libflame_amd: FLASH_LU_find_zero_on_diagonal: %s\n",
            fla_hip_error_string(st));
    return FLA_GPU_ERROR;
  }
  for (i = 0; i < k; ++i)
    if (diag[i] == 0.0) break;
  free(diag);
  return i < k ? (FLA_Error)i : FLA_SUCCESS;
}

/* ---- hierarchical level-3 front-ends (src/blas/3/{trsm,gemm,herk}/front/flamec/FLASH_*.c):
   the flat-copy helpers make the FLAME-level kernels accept hierarchical views. */
FLA_Error FLASH_Trsm(FLA_Side side, FLA_Uplo uplo, FLA_Trans trans, FLA_Diag diag, FLA_Obj alpha,
                     FLA_Obj A, FLA_Obj B) {
  return FLA_Trsm_external(side, uplo, trans, diag, alpha, A, B);
}
FLA_Error FLASH_Gemm(FLA_Trans transa, FLA_Trans transb, FLA_Obj alpha, FLA_Obj A, FLA_Obj B,
                     FLA_Obj beta, FLA_Obj C) {
  return FLA_Gemm_external(transa, transb, alpha, A, B, beta, C);
}
FLA_Error FLASH_Herk(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A, FLA_Obj beta,
                     FLA_Obj C) {
  return FLA_Herk_external(uplo, trans, alpha, A, beta, C);
}

/* ---- LU with incremental pivoting ---------------------------------------------
 *   FLASH_LU_incpiv        src/lapack/dec/lu/incpiv/front/flamec/FLASH_LU_incpiv.c:13-39
 *                          (-> FLASH_LU_incpiv_opt1.c, zaux/flamec/flash/FLASH_LU_incpiv_var2.c)
 *   create_hier_matrices   .../front/flamec/FLASH_LU_incpiv_create_hier_matrices.c:13-75
 *   FLASH_FS_incpiv        .../front/flamec/FLASH_FS_incpiv.c:13-53
 *   FLASH_LU_incpiv_solve  .../front/flamec/FLASH_LU_incpiv_solve.c:13-27
 * The tile size b is A's storage block (its top-left leaf's order) and the
 * inner block nb_alg is the width of L's top-left leaf, as in the reference.
 * The GPU path needs m and n to be multiples of b: the reference's SA_FS_blk
 * loop runs over the full b rows of an L tile and has no ragged tile. */

dim_t FLASH_LU_incpiv_determine_alg_blocksize(FLA_Obj A) {
  const double b = (double)FLA_Obj_length(*FLASH_OBJ_PTR_AT(A)) * 0.25; /* FLA_LU_INNER_TO_OUTER_B_RATIO */
  return (dim_t)(b > 1.0 ? b : 1.0);
}

/* Zeroes the one device allocation behind a freshly created depth-1
   hierarchy (its leaves are consecutive, starting at the top-left leaf). */
static FLA_Error zero_leaves(FLA_Obj H) {
  const size_t count = (size_t)FLASH_Obj_scalar_length(H) * (size_t)FLASH_Obj_scalar_width(H);
  void* buf;
  int st;
  if (count == 0) return FLA_SUCCESS;
  buf = FLA_Obj_buffer_at_view(*FLASH_OBJ_PTR_AT(H));
  if (FLA_Obj_datatype(H) == FLA_DOUBLE && count <= 0x7fffffffu) {
    st = fla_hip_dlaset(0, (int)count, 1, 0.0, 0.0, (double*)buf, (int)count);
    if (st == FLA_HIP_OK) st = fla_hip_synchronize();
  } else {
    const size_t bytes = count * (size_t)FLA_Obj_elem_size(*FLASH_OBJ_PTR_AT(H));
    void* z = calloc(bytes, 1);
    if (!z) return FLA_FAILURE;
    st = fla_hip_memcpy_h2d(buf, z, bytes);
    free(z);
  }
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: FLASH_LU_incpiv_create_hier_matrices: %s\n",
            fla_hip_error_string(st));
    return FLA_GPU_ERROR;
  }
  return FLA_SUCCESS;
}

/* p and L are zeroed on creation: the reference leaves them uninitialised
   (FLASH_Obj_create_ext), and the parts of L that the tile algorithm never
   writes (beyond a ragged last inner block) would otherwise hold whatever
   the allocation held before. */
FLA_Error FLASH_LU_incpiv_create_hier_matrices(FLA_Obj A_flat, dim_t depth, dim_t* b_flash,
                                               dim_t b_alg, FLA_Obj* A, FLA_Obj* p, FLA_Obj* L) {
  dim_t one = 1, m, n;
  FLA_Error e;
  if (depth != 1) {
    fprintf(stderr, "libflame_amd: FLASH_LU_incpiv() currently only supports matrices of depth 1\n");
    abort();
  }
  if ((e = FLASH_Obj_create_hier_copy_of_flat(A_flat, depth, b_flash, A)) != FLA_SUCCESS) return e;
  if (b_alg == 0) b_alg = FLASH_LU_incpiv_determine_alg_blocksize(*A);
  m = FLA_Obj_length(*A); /* in blocks: p and L get full blocks on the edges */
  n = FLA_Obj_width(*A);
  if ((e = FLASH_Obj_create_ext(FLA_INT, m * b_flash[0], n, depth, b_flash, &one, p)) !=
      FLA_SUCCESS)
    return e;
  if ((e = zero_leaves(*p)) != FLA_SUCCESS) return e;
  if ((e = FLASH_Obj_create_ext(FLA_Obj_datatype(A_flat), m * b_flash[0], n * b_alg, depth,
                                b_flash, &b_alg, L)) != FLA_SUCCESS)
    return e;
  return zero_leaves(*L);
}

static FLA_Error incpiv_checks(FLA_Obj A, FLA_Obj p, FLA_Obj L, const char* fn) {
  require_depth_one(A, fn);
  if (CHECK_LEVEL_MIN) {
    FLA_Check_error_code(fla_check_floating_object(A));
    FLA_Check_error_code(fla_check_nonconstant_object(A));
    FLA_Check_error_code(fla_check_int_object(p));
    FLA_Check_error_code(fla_check_identical_object_datatype(A, L));
  }
  if (FLA_Obj_datatype(A) != FLA_DOUBLE) {
    fprintf(stderr, "libflame_amd: the GPU path computes in FP64; got datatype %d\n",
            FLA_Obj_datatype(A));
    return FLA_UNSUPPORTED_DATATYPE;
  }
  return FLA_SUCCESS;
}

FLA_Error FLASH_LU_incpiv(FLA_Obj A, FLA_Obj p, FLA_Obj L) {
  fla_devmat dA, dp, dL;
  FLA_Error e;
  int st, b, nb_alg;
  if ((e = incpiv_checks(A, p, L, "FLASH_LU_incpiv")) != FLA_SUCCESS) return e;
  b = (int)FLASH_Obj_scalar_length_tl(A);
  nb_alg = (int)FLASH_Obj_scalar_width_tl(L);
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(p, 1, &dp)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  if ((e = fla_devmat_acquire(L, 1, &dL)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    fla_devmat_release(&dp, 0);
    return e;
  }
  st = (dA.m == 0 || dA.n == 0)
           ? FLA_HIP_OK
           : fla_hip_dgetrf_incpiv(dA.m, dA.n, (double*)dA.dev, dA.ld, b, nb_alg, (int*)dp.dev,
                                   dp.ld, (double*)dL.dev, dL.ld);
  e = fla_devmat_release(&dA, st == FLA_HIP_OK);
  fla_devmat_release(&dp, st == FLA_HIP_OK);
  fla_devmat_release(&dL, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: FLASH_LU_incpiv failed: %s\n", fla_hip_error_string(st));
    return FLA_GPU_ERROR;
  }
  return e;
}

FLA_Error FLASH_FS_incpiv(FLA_Obj A, FLA_Obj p, FLA_Obj L, FLA_Obj b) {
  fla_devmat dA, dp, dL, db;
  FLA_Error e;
  int st, bs, nb_alg;
  if ((e = incpiv_checks(A, p, L, "FLASH_FS_incpiv")) != FLA_SUCCESS) return e;
  bs = (int)FLASH_Obj_scalar_length_tl(A);
  nb_alg = (int)FLASH_Obj_scalar_width_tl(L);
  if ((e = fla_devmat_acquire(A, 1, &dA)) != FLA_SUCCESS) return e;
  if ((e = fla_devmat_acquire(p, 1, &dp)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    return e;
  }
  if ((e = fla_devmat_acquire(L, 1, &dL)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    fla_devmat_release(&dp, 0);
    return e;
  }
  if ((e = fla_devmat_acquire(b, 1, &db)) != FLA_SUCCESS) {
    fla_devmat_release(&dA, 0);
    fla_devmat_release(&dp, 0);
    fla_devmat_release(&dL, 0);
    return e;
  }
  st = (dA.m == 0 || db.n == 0)
           ? FLA_HIP_OK
           : fla_hip_dfs_incpiv(dA.m, dA.n, (const double*)dA.dev, dA.ld, bs, nb_alg,
                                (const int*)dp.dev, dp.ld, (const double*)dL.dev, dL.ld, db.n,
                                (double*)db.dev, db.ld);
  fla_devmat_release(&dA, 0);
  fla_devmat_release(&dp, 0);
  fla_devmat_release(&dL, 0);
  e = fla_devmat_release(&db, st == FLA_HIP_OK);
  if (st != FLA_HIP_OK) {
    fprintf(stderr, "libflame_amd: FLASH_FS_incpiv failed: %s\n", fla_hip_error_string(st));
    return FLA_GPU_ERROR;
  }
  return e;
}

FLA_Error FLASH_LU_incpiv_solve(FLA_Obj A, FLA_Obj p, FLA_Obj L, FLA_Obj B, FLA_Obj X) {
  FLA_Error e;
  if ((e = FLASH_Copy(B, X)) != FLA_SUCCESS) return e;
  if ((e = FLASH_FS_incpiv(A, p, L, X)) != FLA_SUCCESS) return e;
  return FLASH_Trsm(FLA_LEFT, FLA_UPPER_TRIANGULAR, FLA_NO_TRANSPOSE, FLA_NONUNIT_DIAG, FLA_ONE, A,
                    X);
}
