/*
 * FLA_base.c -- libflame_amd's base layer: initialization, FLA_Obj lifecycle,
 * queries, views and parameter checks, with libflame's semantics.
 *
 * Follows (paths relative to the libflame source tree):
 *   FLA_Init / FLA_Finalize / *_safe   src/base/flamec/main/FLA_Init.c:47-122
 *   FLA_Obj_create_ext / attach / free src/base/flamec/main/FLA_Obj.c:64-140,522-628
 *   FLA_Obj_buffer_at_view             src/base/flamec/main/FLA_Query.c:215-235
 *   views                              src/base/flamec/main/FLA_View.c:17-562
 *   FLA_Check_error_code_helper        src/base/flamec/main/FLA_Check.c:38-63
 *
 * The buffer of an object made by FLA_Obj_create lives in GPU memory; the
 * base's buffer_info field (unused by flat libflame objects) records whether
 * a buffer is device-resident and whether it was attached.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "FLAME.h"
#include "FLA_hip.h"
#include "fla_internal.h"

static int fla_initialized = 0;
static unsigned int fla_error_level = FLA_FULL_ERROR_CHECKING;
static unsigned long fla_next_id = 1;

FLA_Obj FLA_ONE;
FLA_Obj FLA_ZERO;
FLA_Obj FLA_MINUS_ONE;

/* Constants carry one value per datatype, as libflame's FLA_CONSTANT objects
   (src/base/flamec/main/FLA_Init.c:16-34). */
typedef struct {
  int i;
  float s;
  double d;
  float c[2];
  double z[2];
} fla_const_vals;

static fla_const_vals fla_cv_one = {1, 1.0f, 1.0, {1.0f, 0.0f}, {1.0, 0.0}};
static fla_const_vals fla_cv_zero = {0, 0.0f, 0.0, {0.0f, 0.0f}, {0.0, 0.0}};
static fla_const_vals fla_cv_minus_one = {-1, -1.0f, -1.0, {-1.0f, 0.0f}, {-1.0, 0.0}};
static FLA_Base_obj fla_base_one, fla_base_zero, fla_base_minus_one;

static void make_const(FLA_Obj* obj, FLA_Base_obj* base, fla_const_vals* v) {
  memset(base, 0, sizeof(*base));
  base->datatype = FLA_CONSTANT;
  base->elemtype = FLA_SCALAR;
  base->m = base->n = 1;
  base->m_inner = base->n_inner = 1;
  base->rs = 1;
  base->cs = 1;
  base->buffer = v;
  base->buffer_info = FLA_BUFFER_HOST | FLA_BUFFER_ATTACHED;
  base->uplo = FLA_FULL_MATRIX;
  memset(obj, 0, sizeof(*obj));
  obj->m = obj->n = 1;
  obj->m_inner = obj->n_inner = 1;
  obj->base = base;
}

double fla_obj_double_value(FLA_Obj obj) {
  if (obj.base == NULL || obj.base->buffer == NULL) return 0.0;
  if (obj.base->datatype == FLA_CONSTANT) return ((fla_const_vals*)obj.base->buffer)->d;
  if (obj.base->datatype == FLA_DOUBLE) {
    double v = 0.0;
    const double* p = (const double*)FLA_Obj_buffer_at_view(obj);
    if (FLA_Obj_is_device(obj))
      fla_hip_memcpy_d2h(&v, p, sizeof(double));
    else
      v = *p;
    return v;
  }
  if (obj.base->datatype == FLA_FLOAT) return (double)*(const float*)FLA_Obj_buffer_at_view(obj);
  return 0.0;
}

void FLA_Init(void) {
  if (fla_initialized) return;
  make_const(&FLA_ONE, &fla_base_one, &fla_cv_one);
  make_const(&FLA_ZERO, &fla_base_zero, &fla_cv_zero);
  make_const(&FLA_MINUS_ONE, &fla_base_minus_one, &fla_cv_minus_one);
  fla_initialized = 1;
}

void FLA_Finalize(void) {
  if (!fla_initialized) return;
  fla_hip_finalize();
  fla_initialized = 0;
}

FLA_Bool FLA_Initialized(void) { return fla_initialized ? TRUE : FALSE; }

void FLA_Init_safe(FLA_Error* init_result) {
  if (FLA_Initialized()) {
    *init_result = FLA_FAILURE;
  } else {
    FLA_Init();
    *init_result = FLA_SUCCESS;
  }
}

/* The *_safe pair brackets one library call made without FLA_Init (the
   LAPACK entry points): it sets up and clears only the FLAME-level state.
   The device context (streams, events, workspaces) is kept for the next
   call and released by FLA_Finalize or at process exit, so a LAPACK caller
   does not recreate streams on every call. */
void FLA_Finalize_safe(FLA_Error init_result) {
  if (init_result == FLA_SUCCESS) fla_initialized = 0;
}

unsigned int FLA_Check_error_level(void) { return fla_error_level; }

unsigned int FLA_Check_error_level_set(unsigned int new_level) {
  unsigned int old = fla_error_level;
  if (new_level > FLA_FULL_ERROR_CHECKING) {
    fprintf(stderr, "libflame: invalid error checking level\n");
    abort();
  }
  fla_error_level = new_level;
  return old;
}

/* ---- error reporting (FLA_Check.c:38-63) ---- */
static const char* fla_error_string(int code) {
  switch (code) {
    case FLA_INVALID_SIDE: return "Invalid side parameter value.";
    case FLA_INVALID_UPLO: return "Invalid uplo parameter value.";
    case FLA_INVALID_TRANS: return "Invalid trans parameter value.";
    case FLA_INVALID_DATATYPE: return "Invalid datatype value.";
    case FLA_OBJECT_NOT_INTEGER: return "Expected integer object.";
    case FLA_OBJECT_NOT_SQUARE: return "Expected square matrix object.";
    case FLA_OBJECT_NOT_VECTOR: return "Expected vector object.";
    case FLA_INCONSISTENT_DATATYPES: return "Objects have inconsistent datatypes.";
    case FLA_NONCONFORMAL_DIMENSIONS: return "Objects have non-conformal dimensions.";
    case FLA_NULL_POINTER: return "Encountered NULL pointer.";
    case FLA_INVALID_PIVOT_TYPE: return "Invalid pivot index type specified.";
    case FLA_OBJECT_NOT_FLOATING_POINT: return "Expected floating-point object.";
    case FLA_OBJECT_NOT_NONCONSTANT: return "Expected non-constant object.";
    case FLA_INVALID_VECTOR_DIM: return "Invalid vector dimension.";
    case FLA_INVALID_DIAG: return "Invalid diag parameter value.";
    case FLA_INVALID_QUADRANT: return "Invalid quadrant parameter value.";
    case FLA_UNSUPPORTED_DATATYPE: return "The GPU path supports FLA_DOUBLE matrices only.";
    case FLA_GPU_ERROR: return "HIP runtime error.";
    default: return "Undefined error code.";
  }
}

FLA_Error fla_check_error_code_helper(int code, const char* file, int line) {
  if (code == FLA_SUCCESS) return code;
  fprintf(stderr, "libflame: %s (line %d):\nlibflame: %s\n", file, line, fla_error_string(code));
  fflush(stderr);
  abort();
  return code;
}

/* ---- checks (src/base/flamec/check/base/...) ---- */
static int is_floating(FLA_Datatype d) {
  return d == FLA_FLOAT || d == FLA_DOUBLE || d == FLA_COMPLEX || d == FLA_DOUBLE_COMPLEX ||
         d == FLA_CONSTANT;
}

FLA_Error fla_check_valid_uplo(FLA_Uplo u) {
  return (u == FLA_LOWER_TRIANGULAR || u == FLA_UPPER_TRIANGULAR) ? FLA_SUCCESS
                                                                   : FLA_INVALID_UPLO;
}
FLA_Error fla_check_valid_side(FLA_Side s) {
  return (s == FLA_LEFT || s == FLA_RIGHT) ? FLA_SUCCESS : FLA_INVALID_SIDE;
}
FLA_Error fla_check_valid_trans(FLA_Trans t) {
  return (t == FLA_NO_TRANSPOSE || t == FLA_TRANSPOSE || t == FLA_CONJ_TRANSPOSE ||
          t == FLA_CONJ_NO_TRANSPOSE)
             ? FLA_SUCCESS
             : FLA_INVALID_TRANS;
}
FLA_Error fla_check_valid_diag(FLA_Diag d) {
  return (d == FLA_UNIT_DIAG || d == FLA_NONUNIT_DIAG || d == FLA_ZERO_DIAG) ? FLA_SUCCESS
                                                                             : FLA_INVALID_DIAG;
}
FLA_Error fla_check_floating_object(FLA_Obj A) {
  return is_floating(FLA_Obj_datatype(A)) ? FLA_SUCCESS : FLA_OBJECT_NOT_FLOATING_POINT;
}
FLA_Error fla_check_nonconstant_object(FLA_Obj A) {
  return FLA_Obj_datatype(A) == FLA_CONSTANT ? FLA_OBJECT_NOT_NONCONSTANT : FLA_SUCCESS;
}
FLA_Error fla_check_int_object(FLA_Obj A) {
  return FLA_Obj_datatype(A) == FLA_INT ? FLA_SUCCESS : FLA_OBJECT_NOT_INTEGER;
}
FLA_Error fla_check_square(FLA_Obj A) {
  return FLA_Obj_length(A) == FLA_Obj_width(A) ? FLA_SUCCESS : FLA_OBJECT_NOT_SQUARE;
}
FLA_Error fla_check_conformal_dims(FLA_Trans trans, FLA_Obj A, FLA_Obj B) {
  const dim_t m = (trans == FLA_NO_TRANSPOSE) ? FLA_Obj_length(A) : FLA_Obj_width(A);
  const dim_t n = (trans == FLA_NO_TRANSPOSE) ? FLA_Obj_width(A) : FLA_Obj_length(A);
  return (m == FLA_Obj_length(B) && n == FLA_Obj_width(B)) ? FLA_SUCCESS
                                                           : FLA_NONCONFORMAL_DIMENSIONS;
}
FLA_Error fla_check_col_vector(FLA_Obj A) {
  return FLA_Obj_width(A) == 1 ? FLA_SUCCESS : FLA_OBJECT_NOT_VECTOR;
}
FLA_Error fla_check_if_vector(FLA_Obj A) {
  return (FLA_Obj_width(A) == 1 || FLA_Obj_length(A) == 1) ? FLA_SUCCESS : FLA_OBJECT_NOT_VECTOR;
}
FLA_Error fla_check_vector_dim_min(FLA_Obj x, dim_t dim) {
  return FLA_Obj_vector_dim(x) >= dim ? FLA_SUCCESS : FLA_INVALID_VECTOR_DIM;
}
FLA_Error fla_check_identical_object_datatype(FLA_Obj A, FLA_Obj B) {
  return FLA_Obj_datatype(A) == FLA_Obj_datatype(B) ? FLA_SUCCESS : FLA_INCONSISTENT_DATATYPES;
}
/* op(A) * X == B dimensions (FLA_Check_matrix_matrix_dims). */
FLA_Error fla_check_matrix_matrix_dims(FLA_Trans ta, FLA_Trans tb, FLA_Obj A, FLA_Obj B,
                                       FLA_Obj C) {
  dim_t ma = (ta == FLA_NO_TRANSPOSE || ta == FLA_CONJ_NO_TRANSPOSE) ? FLA_Obj_length(A)
                                                                     : FLA_Obj_width(A);
  dim_t ka = (ta == FLA_NO_TRANSPOSE || ta == FLA_CONJ_NO_TRANSPOSE) ? FLA_Obj_width(A)
                                                                     : FLA_Obj_length(A);
  dim_t kb = (tb == FLA_NO_TRANSPOSE || tb == FLA_CONJ_NO_TRANSPOSE) ? FLA_Obj_length(B)
                                                                     : FLA_Obj_width(B);
  dim_t nb = (tb == FLA_NO_TRANSPOSE || tb == FLA_CONJ_NO_TRANSPOSE) ? FLA_Obj_width(B)
                                                                     : FLA_Obj_length(B);
  if (ma != FLA_Obj_length(C) || nb != FLA_Obj_width(C) || ka != kb)
    return FLA_NONCONFORMAL_DIMENSIONS;
  return FLA_SUCCESS;
}

/* ---- objects ---- */
static dim_t elem_size_of(FLA_Datatype dt) {
  switch (dt) {
    case FLA_INT: return sizeof(int);
    case FLA_FLOAT: return sizeof(float);
    case FLA_DOUBLE: return sizeof(double);
    case FLA_COMPLEX: return 2 * sizeof(float);
    case FLA_DOUBLE_COMPLEX: return 2 * sizeof(double);
    case FLA_CONSTANT: return sizeof(fla_const_vals);
    default: return 0;
  }
}

static void fill_base(FLA_Base_obj* b, FLA_Datatype dt, dim_t m, dim_t n) {
  memset(b, 0, sizeof(*b));
  b->datatype = dt;
  b->elemtype = FLA_SCALAR;
  b->m = m;
  b->n = n;
  b->m_inner = m;
  b->n_inner = n;
  b->id = fla_next_id++;
  b->uplo = FLA_FULL_MATRIX;
}

static void default_strides(dim_t m, dim_t n, dim_t* rs, dim_t* cs) {
  if (*rs == 0 && *cs == 0) {
    *rs = 1;
    *cs = m > 0 ? m : 1;
  }
  (void)n;
}

FLA_Error FLA_Obj_create_without_buffer(FLA_Datatype datatype, dim_t m, dim_t n, FLA_Obj* obj) {
  FLA_Base_obj* b = (FLA_Base_obj*)malloc(sizeof(FLA_Base_obj));
  if (b == NULL) return FLA_FAILURE;
  fill_base(b, datatype, m, n);
  b->buffer_info = FLA_BUFFER_HOST | FLA_BUFFER_ATTACHED;
  memset(obj, 0, sizeof(*obj));
  obj->m = m;
  obj->n = n;
  obj->m_inner = m;
  obj->n_inner = n;
  obj->base = b;
  return FLA_SUCCESS;
}

FLA_Error FLA_Obj_create(FLA_Datatype datatype, dim_t m, dim_t n, dim_t rs, dim_t cs,
                         FLA_Obj* obj) {
  FLA_Error e = FLA_Obj_create_without_buffer(datatype, m, n, obj);
  if (e != FLA_SUCCESS) return e;
  default_strides(m, n, &rs, &cs);
  obj->base->rs = rs;
  obj->base->cs = cs;
  obj->base->buffer_info = FLA_BUFFER_DEVICE;
  if (m > 0 && n > 0) {
    dim_t nelem = (m - 1) * rs + (n - 1) * cs + 1;
    void* p = NULL;
    int st = fla_hip_malloc(&p, (size_t)(nelem * elem_size_of(datatype)));
    if (st != FLA_HIP_OK || p == NULL) {
      fprintf(stderr, "libflame_amd: device allocation of %lu bytes failed (%s)\n",
              (unsigned long)(nelem * elem_size_of(datatype)), fla_hip_error_string(st));
      free(obj->base);
      obj->base = NULL;
      return FLA_GPU_ERROR;
    }
    obj->base->buffer = p;
    obj->base->n_elem_alloc = nelem;
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Obj_attach_buffer(void* buffer, dim_t rs, dim_t cs, FLA_Obj* obj) {
  dim_t m = obj->base->m, n = obj->base->n;
  default_strides(m, n, &rs, &cs);
  obj->base->buffer = buffer;
  obj->base->rs = rs;
  obj->base->cs = cs;
  obj->base->buffer_info =
      FLA_BUFFER_ATTACHED | (fla_hip_is_device_ptr(buffer) ? FLA_BUFFER_DEVICE : FLA_BUFFER_HOST);
  return FLA_SUCCESS;
}

FLA_Error FLA_Obj_free(FLA_Obj* obj) {
  if (obj->base != NULL) {
    if (!(obj->base->buffer_info & FLA_BUFFER_ATTACHED) && obj->base->buffer != NULL)
      fla_hip_free(obj->base->buffer);
    free(obj->base);
  }
  obj->offm = obj->offn = obj->m = obj->n = 0;
  obj->base = NULL;
  return FLA_SUCCESS;
}

FLA_Error FLA_Obj_free_without_buffer(FLA_Obj* obj) {
  if (obj->base != NULL) free(obj->base);
  obj->offm = obj->offn = obj->m = obj->n = 0;
  obj->base = NULL;
  return FLA_SUCCESS;
}

FLA_Datatype FLA_Obj_datatype(FLA_Obj obj) { return obj.base ? obj.base->datatype : 0; }
dim_t FLA_Obj_length(FLA_Obj obj) { return obj.m; }
dim_t FLA_Obj_width(FLA_Obj obj) { return obj.n; }
dim_t FLA_Obj_min_dim(FLA_Obj obj) { return obj.m < obj.n ? obj.m : obj.n; }
dim_t FLA_Obj_max_dim(FLA_Obj obj) { return obj.m > obj.n ? obj.m : obj.n; }
dim_t FLA_Obj_vector_dim(FLA_Obj obj) { return obj.m == 1 ? obj.n : obj.m; }
dim_t FLA_Obj_vector_inc(FLA_Obj obj) { return obj.m == 1 ? obj.base->cs : obj.base->rs; }
dim_t FLA_Obj_row_stride(FLA_Obj obj) { return obj.base->rs; }
dim_t FLA_Obj_col_stride(FLA_Obj obj) { return obj.base->cs; }
dim_t fla_datatype_size(FLA_Datatype dt) { return elem_size_of(dt); }

/* Elements of a hierarchical node (elemtype FLA_MATRIX) are FLA_Obj. */
static dim_t obj_elem_size(const FLA_Base_obj* b) {
  return b->elemtype == FLA_MATRIX ? (dim_t)sizeof(FLA_Obj) : elem_size_of(b->datatype);
}
dim_t FLA_Obj_elem_size(FLA_Obj obj) { return obj.base ? obj_elem_size(obj.base) : 0; }

void* FLA_Obj_buffer_at_view(FLA_Obj obj) {
  char* p = (char*)obj.base->buffer;
  if (p == NULL) return NULL;
  return p + (obj.offn * obj.base->cs + obj.offm * obj.base->rs) * obj_elem_size(obj.base);
}

FLA_Elemtype FLA_Obj_elemtype(FLA_Obj obj) { return obj.base ? obj.base->elemtype : FLA_SCALAR; }
void* FLA_Obj_base_buffer(FLA_Obj obj) { return obj.base ? obj.base->buffer : NULL; }
dim_t FLA_Obj_base_length(FLA_Obj obj) { return obj.base ? obj.base->m : 0; }
dim_t FLA_Obj_base_width(FLA_Obj obj) { return obj.base ? obj.base->n : 0; }
dim_t FLA_Obj_row_offset(FLA_Obj obj) { return obj.offm; }
dim_t FLA_Obj_col_offset(FLA_Obj obj) { return obj.offn; }
dim_t FLA_Obj_num_elem_alloc(FLA_Obj obj) { return obj.base ? obj.base->n_elem_alloc : 0; }

FLA_Bool FLA_Obj_is_identical(FLA_Obj A, FLA_Obj B) {
  if (A.base != B.base) return FALSE;
  return (A.offm == B.offm && A.offn == B.offn && A.m == B.m && A.n == B.n) ? TRUE : FALSE;
}

FLA_Bool FLA_Obj_has_zero_dim(FLA_Obj A) { return (A.m == 0 || A.n == 0) ? TRUE : FALSE; }

FLA_Bool FLA_Obj_is_device(FLA_Obj obj) {
  return (obj.base && (obj.base->buffer_info & FLA_BUFFER_DEVICE)) ? TRUE : FALSE;
}

/* ---- views (FLA_View.c) ---- */
static FLA_Obj sub(FLA_Obj A, dim_t offm, dim_t offn, dim_t m, dim_t n) {
  FLA_Obj r = A;
  r.offm = A.offm + offm;
  r.offn = A.offn + offn;
  r.m = m;
  r.n = n;
  r.m_inner = m;
  r.n_inner = n;
  return r;
}

FLA_Error FLA_Part_2x2(FLA_Obj A, FLA_Obj* A11, FLA_Obj* A12, FLA_Obj* A21, FLA_Obj* A22,
                       dim_t mb, dim_t nb, FLA_Quadrant quadrant) {
  dim_t m = A.m, n = A.n;
  if (mb > m) mb = m;
  if (nb > n) nb = n;
  if (quadrant == FLA_BL || quadrant == FLA_BR) mb = m - mb;
  if (quadrant == FLA_TR || quadrant == FLA_BR) nb = n - nb;
  *A11 = sub(A, 0, 0, mb, nb);
  *A12 = sub(A, 0, nb, mb, n - nb);
  *A21 = sub(A, mb, 0, m - mb, nb);
  *A22 = sub(A, mb, nb, m - mb, n - nb);
  return FLA_SUCCESS;
}

FLA_Error FLA_Part_2x1(FLA_Obj A, FLA_Obj* A1, FLA_Obj* A2, dim_t mb, FLA_Side side) {
  if (mb > A.m) mb = A.m;
  if (side == FLA_BOTTOM) mb = A.m - mb;
  *A1 = sub(A, 0, 0, mb, A.n);
  *A2 = sub(A, mb, 0, A.m - mb, A.n);
  return FLA_SUCCESS;
}

FLA_Error FLA_Part_1x2(FLA_Obj A, FLA_Obj* A1, FLA_Obj* A2, dim_t nb, FLA_Side side) {
  if (nb > A.n) nb = A.n;
  if (side == FLA_RIGHT) nb = A.n - nb;
  *A1 = sub(A, 0, 0, A.m, nb);
  *A2 = sub(A, 0, nb, A.m, A.n - nb);
  return FLA_SUCCESS;
}

FLA_Error FLA_Merge_2x2(FLA_Obj A11, FLA_Obj A12, FLA_Obj A21, FLA_Obj A22, FLA_Obj* A) {
  (void)A12;
  (void)A21;
  *A = sub(A11, 0, 0, A11.m + A22.m, A11.n + A22.n);
  A->offm = A11.offm;
  A->offn = A11.offn;
  return FLA_SUCCESS;
}

FLA_Error FLA_Merge_2x1(FLA_Obj AT, FLA_Obj AB, FLA_Obj* A) {
  *A = AT;
  A->m = AT.m + AB.m;
  A->m_inner = A->m;
  return FLA_SUCCESS;
}

FLA_Error FLA_Merge_1x2(FLA_Obj AL, FLA_Obj AR, FLA_Obj* A) {
  *A = AL;
  A->n = AL.n + AR.n;
  A->n_inner = A->n;
  return FLA_SUCCESS;
}

FLA_Error FLA_Repart_2x2_to_3x3(FLA_Obj ATL, FLA_Obj ATR, FLA_Obj* A00, FLA_Obj* A01,
                                FLA_Obj* A02, FLA_Obj* A10, FLA_Obj* A11, FLA_Obj* A12,
                                FLA_Obj ABL, FLA_Obj ABR, FLA_Obj* A20, FLA_Obj* A21,
                                FLA_Obj* A22, dim_t mb, dim_t nb, FLA_Quadrant quadrant) {
  if (quadrant == FLA_BR) {
    if (mb > ABR.m) mb = ABR.m;
    if (nb > ABR.n) nb = ABR.n;
    *A00 = ATL;
    *A01 = sub(ATR, 0, 0, ATR.m, nb);
    *A02 = sub(ATR, 0, nb, ATR.m, ATR.n - nb);
    *A10 = sub(ABL, 0, 0, mb, ABL.n);
    *A11 = sub(ABR, 0, 0, mb, nb);
    *A12 = sub(ABR, 0, nb, mb, ABR.n - nb);
    *A20 = sub(ABL, mb, 0, ABL.m - mb, ABL.n);
    *A21 = sub(ABR, mb, 0, ABR.m - mb, nb);
    *A22 = sub(ABR, mb, nb, ABR.m - mb, ABR.n - nb);
  } else if (quadrant == FLA_TL) {
    if (mb > ATL.m) mb = ATL.m;
    if (nb > ATL.n) nb = ATL.n;
    *A00 = sub(ATL, 0, 0, ATL.m - mb, ATL.n - nb);
    *A01 = sub(ATL, 0, ATL.n - nb, ATL.m - mb, nb);
    *A02 = sub(ATR, 0, 0, ATR.m - mb, ATR.n);
    *A10 = sub(ATL, ATL.m - mb, 0, mb, ATL.n - nb);
    *A11 = sub(ATL, ATL.m - mb, ATL.n - nb, mb, nb);
    *A12 = sub(ATR, ATR.m - mb, 0, mb, ATR.n);
    *A20 = sub(ABL, 0, 0, ABL.m, ABL.n - nb);
    *A21 = sub(ABL, 0, ABL.n - nb, ABL.m, nb);
    *A22 = ABR;
  } else {
    return fla_check_error_code_helper(FLA_INVALID_QUADRANT, __FILE__, __LINE__);
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Cont_with_3x3_to_2x2(FLA_Obj* ATL, FLA_Obj* ATR, FLA_Obj A00, FLA_Obj A01,
                                   FLA_Obj A02, FLA_Obj A10, FLA_Obj A11, FLA_Obj A12,
                                   FLA_Obj* ABL, FLA_Obj* ABR, FLA_Obj A20, FLA_Obj A21,
                                   FLA_Obj A22, FLA_Quadrant quadrant) {
  (void)A01;
  (void)A10;
  (void)A12;
  (void)A21;
  if (quadrant == FLA_TL) {
    /* ATL = [A00 A01; A10 A11], ATR = [A02; A12], ABL = [A20 A21], ABR = A22 */
    *ATL = sub(A00, 0, 0, A00.m + A11.m, A00.n + A11.n);
    *ATR = sub(A02, 0, 0, A02.m + A12.m, A02.n);
    *ABL = sub(A20, 0, 0, A20.m, A20.n + A21.n);
    *ABR = A22;
  } else if (quadrant == FLA_BR) {
    *ATL = A00;
    *ATR = sub(A01, 0, 0, A01.m, A01.n + A02.n);
    *ABL = sub(A10, 0, 0, A10.m + A20.m, A10.n);
    *ABR = sub(A11, 0, 0, A11.m + A22.m, A11.n + A22.n);
  } else {
    return fla_check_error_code_helper(FLA_INVALID_QUADRANT, __FILE__, __LINE__);
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Repart_2x1_to_3x1(FLA_Obj AT, FLA_Obj* A0, FLA_Obj* A1, FLA_Obj AB, FLA_Obj* A2,
                                dim_t mb, FLA_Side side) {
  if (side == FLA_BOTTOM) {
    if (mb > AB.m) mb = AB.m;
    *A0 = AT;
    *A1 = sub(AB, 0, 0, mb, AB.n);
    *A2 = sub(AB, mb, 0, AB.m - mb, AB.n);
  } else {
    if (mb > AT.m) mb = AT.m;
    *A0 = sub(AT, 0, 0, AT.m - mb, AT.n);
    *A1 = sub(AT, AT.m - mb, 0, mb, AT.n);
    *A2 = AB;
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Cont_with_3x1_to_2x1(FLA_Obj* AT, FLA_Obj A0, FLA_Obj A1, FLA_Obj* AB, FLA_Obj A2,
                                   FLA_Side side) {
  if (side == FLA_TOP) {
    *AT = sub(A0, 0, 0, A0.m + A1.m, A0.n);
    *AB = A2;
  } else {
    *AT = A0;
    *AB = sub(A1, 0, 0, A1.m + A2.m, A1.n);
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Repart_1x2_to_1x3(FLA_Obj AL, FLA_Obj AR, FLA_Obj* A0, FLA_Obj* A1, FLA_Obj* A2,
                                dim_t nb, FLA_Side side) {
  if (side == FLA_RIGHT) {
    if (nb > AR.n) nb = AR.n;
    *A0 = AL;
    *A1 = sub(AR, 0, 0, AR.m, nb);
    *A2 = sub(AR, 0, nb, AR.m, AR.n - nb);
  } else {
    if (nb > AL.n) nb = AL.n;
    *A0 = sub(AL, 0, 0, AL.m, AL.n - nb);
    *A1 = sub(AL, 0, AL.n - nb, AL.m, nb);
    *A2 = AR;
  }
  return FLA_SUCCESS;
}

FLA_Error FLA_Cont_with_1x3_to_1x2(FLA_Obj* AL, FLA_Obj* AR, FLA_Obj A0, FLA_Obj A1,
                                   FLA_Obj A2, FLA_Side side) {
  if (side == FLA_LEFT) {
    *AL = sub(A0, 0, 0, A0.m, A0.n + A1.n);
    *AR = A2;
  } else {
    *AL = A0;
    *AR = sub(A1, 0, 0, A1.m, A1.n + A2.n);
  }
  return FLA_SUCCESS;
}
