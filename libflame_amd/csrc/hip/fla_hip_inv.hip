// Triangular inverse (FLA_Trinv), triangle times its transpose (FLA_Ttmm)
// and the SPD inverse built from them (FLA_SPDinv) on one MI355X.
//
// Reference algorithms (lower triangle; upper storage is handled by
// exchanging the triangles in place, running the lower algorithm on U', and
// exchanging back, so the strictly lower part of an upper-stored matrix is
// preserved):
//   FLA_Trinv_ln_blk_var3 (src/lapack/inv/tri/ln/flamec/FLA_Trinv_ln_blk_var3.c:36-52):
//     A21 := -A21 inv(A11);  A20 := A20 + A21 A10;  A10 := inv(A11) A10;
//     A11 := inv(A11)
//   FLA_Ttmm_l_blk_var1 (src/lapack/misc/ttmm/l/flamec/FLA_Ttmm_l_blk_var1.c:36-49):
//     A00 := A00 + A10' A10;  A10 := tril(A11)' A10;  A11 := tril(A11)' tril(A11)
//   FLA_SPDinv_internal (src/lapack/inv/spd/front/flamec/FLA_SPDinv_internal.c:22-39):
//     Chol, then Trinv (non-unit), then Ttmm.
// Blocks are 128 wide.  inv(A11) is formed once per step by the diagonal-block
// inverse kernel (LDS, one workgroup) into a 128 x 128 buffer; every product
// is the MFMA GEMM.  The two triangular products whose output overwrites
// their right operand (A21 inv(A11) and inv(A11) A10, tril(A11)' A10) run in
// place: their contracted dimension and one output dimension are one block,
// so each workgroup reads exactly the rows or columns it writes, all before
// its first store.
#include "fla_hip_internal.h"

namespace {

constexpr int TB = 128;

// dst (ld TB) := tril(src), nb x nb, zeros above the diagonal.
__global__ __launch_bounds__(256) void tril_copy_kernel(int nb, const double* __restrict__ src,
                                                        long lds, double* __restrict__ dst) {
  for (int e = blockIdx.x * 256 + threadIdx.x; e < TB * TB; e += gridDim.x * 256) {
    const int i = e % TB, j = e / TB;
    dst[e] = (i < nb && j < nb && i >= j) ? src[(long)i + (long)j * lds] : 0.0;
  }
}

// A11 (ld lda) := lower triangle of W (ld TB); the diagonal is left alone for
// a unit-diagonal matrix (it is not part of the stored data).
__global__ __launch_bounds__(256) void lower_copy_kernel(int nb, const double* __restrict__ W,
                                                         double* __restrict__ A, long lda,
                                                         int unit) {
  for (int e = blockIdx.x * 256 + threadIdx.x; e < nb * nb; e += gridDim.x * 256) {
    const int i = e % nb, j = e / nb;
    if (i > j || (i == j && !unit)) A[(long)i + (long)j * lda] = W[i + j * TB];
  }
}

}  // namespace

// A := inv(A) on the lower triangle (order n, ld lda); W: TB x TB scratch.
static int trinv_lower(hipStream_t s, int n, double* A, long lda, bool unit, double* W) {
  for (int k = 0; k < n; k += TB) {
    const int b = (n - k) < TB ? (n - k) : TB;
    const int mr = n - k - b;
    double* A10 = A + (long)k;
    double* A11 = A + (long)k + (long)k * lda;
    double* A20 = A + (long)(k + b);
    double* A21 = A + (long)(k + b) + (long)k * lda;
    FLA_TRY(launch_trtri_blocks(s, b, A11, lda, false, unit, W));
    if (mr > 0) {
      FLA_TRY(launch_dgemm_mfma(s, false, false, 0, mr, b, b, -1.0, A21, lda, W, TB, 0.0, A21, lda,
                                nullptr));
      if (k > 0)
        FLA_TRY(launch_dgemm_mfma(s, false, false, 0, mr, k, b, 1.0, A21, lda, A10, lda, 1.0, A20,
                                  lda, nullptr));
    }
    if (k > 0)
      FLA_TRY(launch_dgemm_mfma(s, false, false, 0, b, k, b, 1.0, W, TB, A10, lda, 0.0, A10, lda,
                                nullptr));
    hipLaunchKernelGGL(lower_copy_kernel, dim3((b * b + 255) / 256), dim3(256), 0, s, b, W, A11,
                       lda, unit ? 1 : 0);
    FLA_HIP_CHECK(hipGetLastError());
  }
  return FLA_HIP_OK;
}

// A := L' L on the lower triangle (L = tril(A), order n); Z: TB x TB scratch.
static int ttmm_lower(hipStream_t s, int n, double* A, long lda, double* Z) {
  for (int k = 0; k < n; k += TB) {
    const int b = (n - k) < TB ? (n - k) : TB;
    double* A10 = A + (long)k;
    double* A11 = A + (long)k + (long)k * lda;
    hipLaunchKernelGGL(tril_copy_kernel, dim3(64), dim3(256), 0, s, b, A11, lda, Z);
    FLA_HIP_CHECK(hipGetLastError());
    if (k > 0) {
      FLA_TRY(launch_dgemm_mfma(s, true, false, 1, k, k, b, 1.0, A10, lda, A10, lda, 1.0, A, lda,
                                nullptr));
      FLA_TRY(launch_dgemm_mfma(s, true, false, 0, b, k, b, 1.0, Z, TB, A10, lda, 0.0, A10, lda,
                                nullptr));
    }
    FLA_TRY(launch_dgemm_mfma(s, true, false, 1, b, b, b, 1.0, Z, TB, Z, TB, 0.0, A11, lda,
                              nullptr));
  }
  return FLA_HIP_OK;
}

static int with_lower_view(int uplo, int n, double* A, long lda,
                           int (*op)(hipStream_t, int, double*, long, bool, double*), bool unit) {
  FlaHipCtx& c = fla_ctx();
  hipStream_t s = c.s_update;
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(3, (size_t)TB * TB * sizeof(double), &w));
  if (uplo == kUpper) FLA_TRY(launch_transpose_swap(s, n, A, lda));
  FLA_TRY(op(s, n, A, lda, unit, static_cast<double*>(w)));
  if (uplo == kUpper) FLA_TRY(launch_transpose_swap(s, n, A, lda));
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  return FLA_HIP_OK;
}

static int ttmm_op(hipStream_t s, int n, double* A, long lda, bool, double* Z) {
  return ttmm_lower(s, n, A, lda, Z);
}

extern "C" int fla_hip_dtrtri(int uplo, int diag, int n, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if ((uplo != kLower && uplo != kUpper) || (diag != kUnit && diag != kNonUnit)) return FLA_HIP_EARG;
  if (n < 0 || lda < (n > 1 ? n : 1)) return FLA_HIP_EARG;
  if (n == 0) return FLA_HIP_OK;
  return with_lower_view(uplo, n, A, lda, trinv_lower, diag == kUnit);
}

extern "C" int fla_hip_dlauum(int uplo, int n, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (uplo != kLower && uplo != kUpper) return FLA_HIP_EARG;
  if (n < 0 || lda < (n > 1 ? n : 1)) return FLA_HIP_EARG;
  if (n == 0) return FLA_HIP_OK;
  return with_lower_view(uplo, n, A, lda, ttmm_op, false);
}

extern "C" int fla_hip_dspdinv(int uplo, int n, double* A, int lda, int* info) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_hip_dpotrf(uplo, n, A, lda, info));
  if (*info != kSuccess || n == 0) return FLA_HIP_OK;
  FLA_TRY(fla_hip_dtrtri(uplo, kNonUnit, n, A, lda));
  return fla_hip_dlauum(uplo, n, A, lda);
}
