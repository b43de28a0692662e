/* Internal helpers of libflame_amd's C host layer (not installed). */
#ifndef LIBFLAME_AMD_FLA_INTERNAL_H
#define LIBFLAME_AMD_FLA_INTERNAL_H

#include "FLAME.h"

FLA_Error fla_check_error_code_helper(int code, const char* file, int line);
#define FLA_Check_error_code(code) fla_check_error_code_helper((code), __FILE__, __LINE__)

FLA_Error fla_check_valid_uplo(FLA_Uplo u);
FLA_Error fla_check_valid_side(FLA_Side s);
FLA_Error fla_check_valid_trans(FLA_Trans t);
FLA_Error fla_check_valid_diag(FLA_Diag d);
FLA_Error fla_check_floating_object(FLA_Obj A);
FLA_Error fla_check_nonconstant_object(FLA_Obj A);
FLA_Error fla_check_int_object(FLA_Obj A);
FLA_Error fla_check_square(FLA_Obj A);
FLA_Error fla_check_conformal_dims(FLA_Trans trans, FLA_Obj A, FLA_Obj B);
FLA_Error fla_check_col_vector(FLA_Obj A);
FLA_Error fla_check_if_vector(FLA_Obj A);
FLA_Error fla_check_vector_dim_min(FLA_Obj x, dim_t dim);
FLA_Error fla_check_identical_object_datatype(FLA_Obj A, FLA_Obj B);
FLA_Error fla_check_matrix_matrix_dims(FLA_Trans ta, FLA_Trans tb, FLA_Obj A, FLA_Obj B,
                                       FLA_Obj C);
double fla_obj_double_value(FLA_Obj obj);

/* A device-resident, column-major view of an object's data for one call:
   the object's own buffer when it is device memory with unit row stride,
   otherwise a staged device copy that is written back on release. */
typedef struct {
  void* dev;      /* device pointer to element (0,0) */
  int ld;         /* leading dimension (elements) */
  int m, n;
  int elem;       /* element size in bytes */
  int staged;     /* 1 when dev is a temporary copy */
  FLA_Obj obj;
} fla_devmat;

FLA_Error fla_devmat_acquire(FLA_Obj A, int copy_in, fla_devmat* d);
FLA_Error fla_devmat_release(fla_devmat* d, int copy_out);

/* Bytes per element of a scalar datatype. */
dim_t fla_datatype_size(FLA_Datatype dt);

/* Copies between the leaves covering the scalar region (i, j, m, n) of a
   hierarchical view and a device column-major matrix (to_leaves: write the
   leaves).  Asynchronous on the library's update stream. */
FLA_Error fla_hier_copy(FLA_Obj H, dim_t i, dim_t j, dim_t m, dim_t n, void* dev, int ld,
                        int to_leaves);

#endif
