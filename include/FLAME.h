/*
 * FLAME.h -- public C API of libflame_amd, the MI355X implementation of
 * libflame's dense Cholesky / LU hot path.
 *
 * Drop-in subset of libflame's API: the same type names, struct layout,
 * constants (numeric values from src/base/flamec/include/FLA_macro_defs.h),
 * entry points and return conventions.  Difference: FLA_Obj_create puts the
 * buffer in GPU memory (hipMalloc), and every factorization runs on the GPU.
 * FLA_Obj_attach_buffer accepts device or host pointers; host buffers are
 * staged through device memory by the call that uses them.
 *
 * Interfaces replaced (paths relative to the libflame source tree):
 *   FLA_Obj / FLA_Base_obj     src/base/flamec/include/FLA_type_defs.h:132-180
 *   FLA_Init / FLA_Finalize    src/base/flamec/main/FLA_Init.c:47,76
 *   FLA_Obj_create ... free    src/base/flamec/main/FLA_Obj.c:55,362,522,588,615
 *   views (Part/Repart/...)    src/base/flamec/main/FLA_View.c:17-562
 *   FLA_Chol                   src/lapack/dec/chol/front/flamec/FLA_Chol.c:16-29
 *   FLA_LU_piv                 src/lapack/dec/lu/piv/front/flamec/FLA_LU_piv.c:16-33
 *   FLA_Apply_pivots           src/lapack/util/app/piv/front/flamec/FLA_Apply_pivots.c:15-24
 *   FLA_Chol_solve             src/lapack/dec/chol/front/flamec/FLA_Chol_solve.c:13-38
 *   FLA_LU_piv_solve           src/lapack/dec/lu/piv/front/flamec/FLA_LU_piv_solve.c:13-30
 *   FLA_Shift_pivots_to        src/base/flamec/util/lapack/FLA_Shift_pivots_to.c:13-42
 *   FLA_Trsm/Herk/Syrk/Gemm_external  src/base/flamec/wrappers/blas/3/external/
 *   dpotrf_ / dgetrf_          src/map/lapack2flamec/FLA_potrf.c:31-105, FLA_getrf.c:32-117
 *   dpotrs_ / dgetrs_          src/map/common/lapacksrc/f2c/dpotrs.c:102, dgetrs.c:115
 *   dlaswp_                    src/map/common/lapacksrc/f2c/dlaswp.c:103
 *   dposv_ / dgesv_            src/map/common/lapacksrc/f2c/dposv.c:120, dgesv.c:113
 *   dsposv_                    src/map/common/lapacksrc/f2c/dsposv.c:192-415
 *   dsgesv_                    src/map/common/lapacksrc/f2c/dsgesv.c:189-407
 */
#ifndef LIBFLAME_AMD_FLAME_H
#define LIBFLAME_AMD_FLAME_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- types (FLA_type_defs.h) ---- */
typedef int FLA_Bool;
typedef int FLA_Error;
typedef int FLA_Quadrant;
typedef int FLA_Datatype;
typedef int FLA_Elemtype;
typedef int FLA_Side;
typedef int FLA_Uplo;
typedef int FLA_Trans;
typedef int FLA_Conj;
typedef int FLA_Diag;
typedef int FLA_Dimension;
typedef int FLA_Pivot_type;
typedef unsigned long dim_t;

typedef struct FLA_Obj_struct {
  FLA_Datatype datatype;
  FLA_Elemtype elemtype;
  dim_t m;
  dim_t n;
  dim_t rs;
  dim_t cs;
  dim_t m_inner;
  dim_t n_inner;
  unsigned long id;
  dim_t m_index;
  dim_t n_index;
  dim_t n_elem_alloc;
  void* buffer;
  int buffer_info; /* FLA_BUFFER_DEVICE / FLA_BUFFER_HOST | FLA_BUFFER_ATTACHED */
  FLA_Uplo uplo;
} FLA_Base_obj;

typedef struct FLA_Obj_view {
  dim_t offm;
  dim_t offn;
  dim_t m;
  dim_t n;
  dim_t m_inner;
  dim_t n_inner;
  FLA_Base_obj* base;
} FLA_Obj;

/* ---- constants (FLA_macro_defs.h) ---- */
#undef TRUE
#undef FALSE
#define TRUE 1
#define FALSE 0
#define FLA_SUCCESS (-1)
#define FLA_FAILURE (-2)
#define FLA_TL 11
#define FLA_TR 12
#define FLA_BL 21
#define FLA_BR 22
#define FLA_FLOAT 100
#define FLA_DOUBLE 101
#define FLA_COMPLEX 102
#define FLA_DOUBLE_COMPLEX 103
#define FLA_INT 104
#define FLA_CONSTANT 105
#define FLA_MATRIX 150
#define FLA_SCALAR 151
#define FLA_TOP 200
#define FLA_BOTTOM 201
#define FLA_LEFT 210
#define FLA_RIGHT 211
#define FLA_LOWER_TRIANGULAR 300
#define FLA_UPPER_TRIANGULAR 301
#define FLA_ZERO_MATRIX 310
#define FLA_FULL_MATRIX 311
#define FLA_NO_TRANSPOSE 400
#define FLA_TRANSPOSE 401
#define FLA_CONJ_TRANSPOSE 402
#define FLA_CONJ_NO_TRANSPOSE 403
#define FLA_NO_CONJUGATE 450
#define FLA_CONJUGATE 451
#define FLA_UNIT_DIAG 500
#define FLA_NONUNIT_DIAG 501
#define FLA_ZERO_DIAG 502
#define FLA_DIMENSION_M 600
#define FLA_DIMENSION_K 601
#define FLA_DIMENSION_N 602
#define FLA_DIMENSION_MIN 603
#define FLA_NATIVE_PIVOTS 700
#define FLA_LAPACK_PIVOTS 701
#define FLA_FULL_ERROR_CHECKING 2
#define FLA_MIN_ERROR_CHECKING 1
#define FLA_NO_ERROR_CHECKING 0

/* error codes used by the hot-path checks (FLA_macro_defs.h:208-309) */
#define FLA_INVALID_SIDE (-10)
#define FLA_INVALID_UPLO (-11)
#define FLA_INVALID_TRANS (-12)
#define FLA_OBJECT_NOT_INTEGER (-21)
#define FLA_INVALID_DATATYPE (-17)
#define FLA_OBJECT_NOT_SQUARE (-24)
#define FLA_OBJECT_NOT_VECTOR (-26)
#define FLA_INCONSISTENT_DATATYPES (-27)
#define FLA_NONCONFORMAL_DIMENSIONS (-28)
#define FLA_NULL_POINTER (-32)
#define FLA_INVALID_DIAG (-58)
#define FLA_INVALID_QUADRANT (-53)
#define FLA_UNDEFINED_ERROR_CODE (-57)
#define FLA_INVALID_PIVOT_TYPE (-35)
#define FLA_OBJECT_NOT_FLOATING_POINT (-45)
#define FLA_OBJECT_NOT_NONCONSTANT (-80)
#define FLA_INVALID_VECTOR_DIM (-97)
#define FLA_CHOL_FAILED_MATRIX_NOT_SPD (-67)
#define FLA_ERROR_CODE_MIN (-10)
#define FLA_ERROR_CODE_MAX (-111)
/* libflame_amd codes, outside libflame's [-111, -10] range */
#define FLA_UNSUPPORTED_DATATYPE (-201) /* the GPU path computes in FP64 */
#define FLA_GPU_ERROR (-202)            /* HIP runtime failure */

/* buffer_info bits (libflame leaves this field unused in flat objects) */
#define FLA_BUFFER_HOST 0
#define FLA_BUFFER_DEVICE 1
#define FLA_BUFFER_ATTACHED 2
#define FLA_BUFFER_TREE 4 /* top-left leaf of a FLASH tree that owns its allocation */

extern FLA_Obj FLA_ONE;
extern FLA_Obj FLA_ZERO;
extern FLA_Obj FLA_MINUS_ONE;

/* ---- init / error checking ---- */
void FLA_Init(void);
void FLA_Finalize(void);
FLA_Bool FLA_Initialized(void);
void FLA_Init_safe(FLA_Error* init_result);
void FLA_Finalize_safe(FLA_Error init_result);
unsigned int FLA_Check_error_level(void);
unsigned int FLA_Check_error_level_set(unsigned int new_level);

/* ---- objects ---- */
FLA_Error FLA_Obj_create(FLA_Datatype datatype, dim_t m, dim_t n, dim_t rs, dim_t cs,
                         FLA_Obj* obj);
FLA_Error FLA_Obj_create_without_buffer(FLA_Datatype datatype, dim_t m, dim_t n, FLA_Obj* obj);
FLA_Error FLA_Obj_attach_buffer(void* buffer, dim_t rs, dim_t cs, FLA_Obj* obj);
FLA_Error FLA_Obj_free(FLA_Obj* obj);
FLA_Error FLA_Obj_free_without_buffer(FLA_Obj* obj);
FLA_Datatype FLA_Obj_datatype(FLA_Obj obj);
dim_t FLA_Obj_length(FLA_Obj obj);
dim_t FLA_Obj_width(FLA_Obj obj);
dim_t FLA_Obj_min_dim(FLA_Obj obj);
dim_t FLA_Obj_max_dim(FLA_Obj obj);
dim_t FLA_Obj_vector_dim(FLA_Obj obj);
dim_t FLA_Obj_vector_inc(FLA_Obj obj);
dim_t FLA_Obj_row_stride(FLA_Obj obj);
dim_t FLA_Obj_col_stride(FLA_Obj obj);
dim_t FLA_Obj_elem_size(FLA_Obj obj);
void* FLA_Obj_buffer_at_view(FLA_Obj obj);
FLA_Bool FLA_Obj_is_identical(FLA_Obj A, FLA_Obj B);
FLA_Bool FLA_Obj_has_zero_dim(FLA_Obj A);
FLA_Bool FLA_Obj_is_device(FLA_Obj obj); /* libflame_amd: buffer in GPU memory */
FLA_Elemtype FLA_Obj_elemtype(FLA_Obj obj);
void* FLA_Obj_base_buffer(FLA_Obj obj);
dim_t FLA_Obj_base_length(FLA_Obj obj);
dim_t FLA_Obj_base_width(FLA_Obj obj);
dim_t FLA_Obj_row_offset(FLA_Obj obj);
dim_t FLA_Obj_col_offset(FLA_Obj obj);
dim_t FLA_Obj_num_elem_alloc(FLA_Obj obj);

/* ---- views (FLA_View.c) ---- */
FLA_Error FLA_Part_2x2(FLA_Obj A, FLA_Obj* A11, FLA_Obj* A12, FLA_Obj* A21, FLA_Obj* A22,
                       dim_t mb, dim_t nb, FLA_Quadrant quadrant);
FLA_Error FLA_Part_2x1(FLA_Obj A, FLA_Obj* A1, FLA_Obj* A2, dim_t mb, FLA_Side side);
FLA_Error FLA_Part_1x2(FLA_Obj A, FLA_Obj* A1, FLA_Obj* A2, dim_t nb, FLA_Side side);
FLA_Error FLA_Merge_2x2(FLA_Obj A11, FLA_Obj A12, FLA_Obj A21, FLA_Obj A22, FLA_Obj* A);
FLA_Error FLA_Merge_2x1(FLA_Obj AT, FLA_Obj AB, FLA_Obj* A);
FLA_Error FLA_Merge_1x2(FLA_Obj AL, FLA_Obj AR, FLA_Obj* A);
FLA_Error FLA_Repart_2x2_to_3x3(FLA_Obj ATL, FLA_Obj ATR, FLA_Obj* A00, FLA_Obj* A01,
                                FLA_Obj* A02, FLA_Obj* A10, FLA_Obj* A11, FLA_Obj* A12,
                                FLA_Obj ABL, FLA_Obj ABR, FLA_Obj* A20, FLA_Obj* A21,
                                FLA_Obj* A22, dim_t mb, dim_t nb, FLA_Quadrant quadrant);
FLA_Error FLA_Cont_with_3x3_to_2x2(FLA_Obj* ATL, FLA_Obj* ATR, FLA_Obj A00, FLA_Obj A01,
                                   FLA_Obj A02, FLA_Obj A10, FLA_Obj A11, FLA_Obj A12,
                                   FLA_Obj* ABL, FLA_Obj* ABR, FLA_Obj A20, FLA_Obj A21,
                                   FLA_Obj A22, FLA_Quadrant quadrant);
FLA_Error FLA_Repart_2x1_to_3x1(FLA_Obj AT, FLA_Obj* A0, FLA_Obj* A1, FLA_Obj AB, FLA_Obj* A2,
                                dim_t mb, FLA_Side side);
FLA_Error FLA_Cont_with_3x1_to_2x1(FLA_Obj* AT, FLA_Obj A0, FLA_Obj A1, FLA_Obj* AB, FLA_Obj A2,
                                   FLA_Side side);
FLA_Error FLA_Repart_1x2_to_1x3(FLA_Obj AL, FLA_Obj AR, FLA_Obj* A0, FLA_Obj* A1, FLA_Obj* A2,
                                dim_t nb, FLA_Side side);
FLA_Error FLA_Cont_with_1x3_to_1x2(FLA_Obj* AL, FLA_Obj* AR, FLA_Obj A0, FLA_Obj A1,
                                   FLA_Obj A2, FLA_Side side);

/* ---- hot path ---- */
FLA_Error FLA_Chol(FLA_Uplo uplo, FLA_Obj A);
FLA_Error FLA_LU_piv(FLA_Obj A, FLA_Obj p);
FLA_Error FLA_Apply_pivots(FLA_Side side, FLA_Trans trans, FLA_Obj p, FLA_Obj A);
FLA_Error FLA_Chol_solve(FLA_Uplo uplo, FLA_Obj A, FLA_Obj B, FLA_Obj X);
FLA_Error FLA_LU_piv_solve(FLA_Obj A, FLA_Obj p, FLA_Obj B, FLA_Obj X);
/* Reverse-mode adjoints (no counterpart in the reference; lower factor only):
   Abar := (G + G')/2, G = L^-T Phi(L' Lbar) L^-1 (Phi: lower part, diagonal
   halved); and for X = A^-1 B: Bbar := A^-1 Xbar, Abar -= (Bbar X' + X Bbar')/2. */
FLA_Error FLA_Chol_adjoint(FLA_Uplo uplo, FLA_Obj L, FLA_Obj Lbar, FLA_Obj Abar);
FLA_Error FLA_Chol_solve_adjoint(FLA_Uplo uplo, FLA_Obj L, FLA_Obj X, FLA_Obj Xbar, FLA_Obj Bbar,
                                 FLA_Obj Abar);
FLA_Error FLA_Shift_pivots_to(FLA_Pivot_type ptype, FLA_Obj p);
FLA_Error FLA_LU_nopiv(FLA_Obj A);
FLA_Error FLA_LU_nopiv_solve(FLA_Obj A, FLA_Obj B, FLA_Obj X);
FLA_Error FLA_Trinv(FLA_Uplo uplo, FLA_Diag diag, FLA_Obj A);
FLA_Error FLA_Ttmm(FLA_Uplo uplo, FLA_Obj A);
FLA_Error FLA_SPDinv(FLA_Uplo uplo, FLA_Obj A);
FLA_Error FLA_Trsm_external(FLA_Side side, FLA_Uplo uplo, FLA_Trans trans, FLA_Diag diag,
                            FLA_Obj alpha, FLA_Obj A, FLA_Obj B);
FLA_Error FLA_Herk_external(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A,
                            FLA_Obj beta, FLA_Obj C);
FLA_Error FLA_Syrk_external(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A,
                            FLA_Obj beta, FLA_Obj C);
FLA_Error FLA_Gemm_external(FLA_Trans transa, FLA_Trans transb, FLA_Obj alpha, FLA_Obj A,
                            FLA_Obj B, FLA_Obj beta, FLA_Obj C);
FLA_Error FLA_Copy(FLA_Obj A, FLA_Obj B);
FLA_Error FLA_Set(FLA_Obj alpha, FLA_Obj A);

/* ---- hierarchical (FLASH) objects ----
   src/base/flamec/hierarchy/main/FLASH_Obj.c, FLASH_View.c, FLASH_Copy_other.c.
   Inner nodes (elemtype FLA_MATRIX) hold host arrays of child objects in
   units of blocks; the leaves are column-major blocks of one device
   allocation, stored in the reference's order (block-column-major at every
   level, leading dimension = the leaf's row count). */
#define FLASH_OBJ_PTR_AT(A) ((FLA_Obj*)FLA_Obj_buffer_at_view(A))
FLA_Error FLASH_Obj_create(FLA_Datatype datatype, dim_t m, dim_t n, dim_t depth, dim_t* b_mn,
                           FLA_Obj* H);
FLA_Error FLASH_Obj_create_ext(FLA_Datatype datatype, dim_t m, dim_t n, dim_t depth, dim_t* b_m,
                               dim_t* b_n, FLA_Obj* H);
FLA_Error FLASH_Obj_create_without_buffer(FLA_Datatype datatype, dim_t m, dim_t n, dim_t depth,
                                          dim_t* b_mn, FLA_Obj* H);
FLA_Error FLASH_Obj_create_without_buffer_ext(FLA_Datatype datatype, dim_t m, dim_t n,
                                              dim_t depth, dim_t* b_m, dim_t* b_n, FLA_Obj* H);
FLA_Error FLASH_Obj_create_conf_to(FLA_Trans trans, FLA_Obj H, FLA_Obj* H_new);
FLA_Error FLASH_Obj_create_hier_conf_to_flat(FLA_Trans trans, FLA_Obj F, dim_t depth,
                                             dim_t* b_mn, FLA_Obj* H);
FLA_Error FLASH_Obj_create_hier_conf_to_flat_ext(FLA_Trans trans, FLA_Obj F, dim_t depth,
                                                 dim_t* b_m, dim_t* b_n, FLA_Obj* H);
FLA_Error FLASH_Obj_create_flat_conf_to_hier(FLA_Trans trans, FLA_Obj H, FLA_Obj* F);
FLA_Error FLASH_Obj_create_copy_of(FLA_Trans trans, FLA_Obj H_cur, FLA_Obj* H_new);
FLA_Error FLASH_Obj_create_hier_copy_of_flat(FLA_Obj F, dim_t depth, dim_t* b_mn, FLA_Obj* H);
FLA_Error FLASH_Obj_create_hier_copy_of_flat_ext(FLA_Obj F, dim_t depth, dim_t* b_m,
                                                 dim_t* b_n, FLA_Obj* H);
FLA_Error FLASH_Obj_create_flat_copy_of_hier(FLA_Obj H, FLA_Obj* F);
void FLASH_Obj_free(FLA_Obj* H);
void FLASH_Obj_free_without_buffer(FLA_Obj* H);
void FLASH_Obj_free_hierarchy(FLA_Obj* H);
void* FLASH_Obj_extract_buffer(FLA_Obj H);
FLA_Error FLASH_Obj_attach_buffer(void* buffer, dim_t rs, dim_t cs, FLA_Obj* H);
FLA_Error FLASH_Obj_attach_buffer_hierarchy(FLA_Obj F, FLA_Obj* H);
FLA_Error FLASH_Obj_flatten(FLA_Obj H, FLA_Obj F);
FLA_Error FLASH_Obj_hierarchify(FLA_Obj F, FLA_Obj H);
FLA_Error FLASH_Copy_flat_to_hier(FLA_Obj F, dim_t i, dim_t j, FLA_Obj H);
FLA_Error FLASH_Copy_hier_to_flat(dim_t i, dim_t j, FLA_Obj H, FLA_Obj F);
FLA_Error FLASH_Copy(FLA_Obj A, FLA_Obj B);
FLA_Datatype FLASH_Obj_datatype(FLA_Obj H);
dim_t FLASH_Obj_depth(FLA_Obj H);
dim_t FLASH_Obj_blocksizes(FLA_Obj H, dim_t* b_m, dim_t* b_n);
dim_t FLASH_Obj_base_scalar_length(FLA_Obj H);
dim_t FLASH_Obj_base_scalar_width(FLA_Obj H);
dim_t FLASH_Obj_scalar_length(FLA_Obj H);
dim_t FLASH_Obj_scalar_width(FLA_Obj H);
dim_t FLASH_Obj_scalar_min_dim(FLA_Obj H);
dim_t FLASH_Obj_scalar_max_dim(FLA_Obj H);
dim_t FLASH_Obj_scalar_vector_dim(FLA_Obj H);
dim_t FLASH_Obj_scalar_row_offset(FLA_Obj H);
dim_t FLASH_Obj_scalar_col_offset(FLA_Obj H);
dim_t FLASH_Obj_scalar_length_tl(FLA_Obj H);
dim_t FLASH_Obj_scalar_width_tl(FLA_Obj H);

/* ---- FLASH front-ends (hierarchical objects; flat objects act as one leaf) ---- */
FLA_Error FLASH_Chol(FLA_Uplo uplo, FLA_Obj A);
FLA_Error FLASH_Chol_solve(FLA_Uplo uplo, FLA_Obj A, FLA_Obj B, FLA_Obj X);
FLA_Error FLASH_LU_piv(FLA_Obj A, FLA_Obj p);
FLA_Error FLASH_LU_piv_solve(FLA_Obj A, FLA_Obj p, FLA_Obj B, FLA_Obj X);
FLA_Error FLASH_Apply_pivots(FLA_Side side, FLA_Trans trans, FLA_Obj p, FLA_Obj A);
FLA_Error FLASH_LU_nopiv(FLA_Obj A);
FLA_Error FLASH_LU_nopiv_solve(FLA_Obj A, FLA_Obj B, FLA_Obj X);
FLA_Error FLASH_LU_find_zero_on_diagonal(FLA_Obj A);
FLA_Error FLASH_LU_incpiv(FLA_Obj A, FLA_Obj p, FLA_Obj L);
FLA_Error FLASH_LU_incpiv_create_hier_matrices(FLA_Obj A_flat, dim_t depth, dim_t* b_flash,
                                               dim_t b_alg, FLA_Obj* A, FLA_Obj* p, FLA_Obj* L);
dim_t FLASH_LU_incpiv_determine_alg_blocksize(FLA_Obj A);
FLA_Error FLASH_FS_incpiv(FLA_Obj A, FLA_Obj p, FLA_Obj L, FLA_Obj b);
FLA_Error FLASH_LU_incpiv_solve(FLA_Obj A, FLA_Obj p, FLA_Obj L, FLA_Obj B, FLA_Obj X);
/* hierarchical level-3 front-ends used by the FLASH solves */
FLA_Error FLASH_Trsm(FLA_Side side, FLA_Uplo uplo, FLA_Trans trans, FLA_Diag diag, FLA_Obj alpha,
                     FLA_Obj A, FLA_Obj B);
FLA_Error FLASH_Gemm(FLA_Trans transa, FLA_Trans transb, FLA_Obj alpha, FLA_Obj A, FLA_Obj B,
                     FLA_Obj beta, FLA_Obj C);
FLA_Error FLASH_Herk(FLA_Uplo uplo, FLA_Trans trans, FLA_Obj alpha, FLA_Obj A, FLA_Obj beta,
                     FLA_Obj C);

/* ---- lapack2flame (host pointers) ---- */
int dpotrf_(char* uplo, int* n, double* buff_A, int* ldim_A, int* info);
int dpotf2_(char* uplo, int* n, double* buff_A, int* ldim_A, int* info);
int dgetrf_(int* m, int* n, double* buff_A, int* ldim_A, int* buff_p, int* info);
int dgetf2_(int* m, int* n, double* buff_A, int* ldim_A, int* buff_p, int* info);
int dpotrs_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb,
            int* info);
int dgetrs_(char* trans, int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b,
            int* ldb, int* info);
int dlaswp_(int* n, double* a, int* lda, int* k1, int* k2, int* ipiv, int* incx);
int dposv_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb, int* info);
int dgesv_(int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b, int* ldb, int* info);
int dsposv_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb, double* x,
            int* ldx, double* work, float* swork, int* iter, int* info);
int dsgesv_(int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b, int* ldb, double* x,
            int* ldx, double* work, float* swork, int* iter, int* info);
int dtrtri_(char* uplo, char* diag, int* n, double* a, int* lda, int* info);
int dlauum_(char* uplo, int* n, double* a, int* lda, int* info);
int dpotri_(char* uplo, int* n, double* a, int* lda, int* info);

#ifdef __cplusplus
}
#endif
#endif
