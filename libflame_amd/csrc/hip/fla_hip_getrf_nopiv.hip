// LU without pivoting (FLA_LU_nopiv) on one MI355X.
//
// Reference: FLA_LU_nopiv_blk_var5
//   (src/lapack/dec/lu/nopiv/vars/flamec/FLA_LU_nopiv_blk_var5.c:36-52):
//     A11 := LU_nopiv(A11);  A12 := trilu(A11) \ A12;  A21 := A21 / triu(A11);
//     A22 := A22 - A21 A12
// with the leaf FLA_LU_nopiv_opd_var5 (FLA_LU_nopiv_opt_var5.c:132-160):
// per column, a21 := a21 / alpha11 (bl1_dinvscalv: by the reciprocal) and
// A22 -= a21 a12t.  The front end (src/lapack/dec/lu/nopiv/front/flamec/
// FLA_LU_nopiv.c:27-35) then reports the first exact zero on the diagonal
// (FLA_LU_find_zero_on_diagonal.c:25-46).
//
// Here: 256-column blocks, each panel factored in 128-column steps; a
// 128 x 128 diagonal block is factored by one workgroup in LDS; the panel
// solves run by substitution in the reference's dtrsm_ order (one lane per
// column / row: without pivoting the triangles need not be well conditioned,
// and an explicit block inverse would amplify that); the trailing update is
// the MFMA GEMM with U12 read transposed and K = 256.
#include "fla_hip_internal.h"

namespace {

constexpr int TB = 128;
constexpr int SLD = TB + 1;

// Unpivoted LU of the nb x nb block at A (nb <= 128), right-looking in LDS.
__global__ __launch_bounds__(256) void lu_nopiv_diag_kernel(int nb, double* __restrict__ A,
                                                            long lda) {
  __shared__ double S[TB * SLD];  // S[c * SLD + r] = A(r, c)
  const int t = threadIdx.x;
  for (int e = t; e < nb * nb; e += 256) {
    const int r = e % nb, c = e / nb;
    S[c * SLD + r] = A[(long)r + (long)c * lda];
  }
  __syncthreads();
  for (int j = 0; j < nb - 1; ++j) {
    const int w = nb - 1 - j;  // trailing order
    const double piv = S[j * SLD + j];
    if (t < w && piv != 1.0) S[j * SLD + j + 1 + t] *= 1.0 / piv;
    __syncthreads();
    for (int e = t; e < w * w; e += 256) {
      const int r = j + 1 + e % w, c = j + 1 + e / w;
      S[c * SLD + r] -= S[j * SLD + r] * S[c * SLD + j];
    }
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += 256) {
    const int r = e % nb, c = e / nb;
    A[(long)r + (long)c * lda] = S[c * SLD + r];
  }
}

// First exact zero on the diagonal of the min(m, n) x min(m, n) leading
// block (atomic minimum of its index; *idx starts at INT_MAX).
__global__ void zero_diag_kernel(int mn, const double* __restrict__ A, long lda, int* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < mn && A[(long)i + (long)i * lda] == 0.0) atomicMin(idx, i);
}

__global__ void set_int_max_kernel(int* p) { *p = 0x7fffffff; }

// dst (ld ldd) := src', an r x c block read as c x r (32 x 32 LDS tiles).
__global__ __launch_bounds__(256) void transpose_nopiv_kernel(int r, int c,
                                                              const double* __restrict__ src,
                                                              long lds, double* __restrict__ dst,
                                                              long ldd) {
  __shared__ double tile[32][33];
  const int l0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    const int l = l0 + tx, j = j0 + q;
    tile[q][tx] = (l < r && j < c) ? src[(long)l + (long)j * lds] : 0.0;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int j = j0 + tx, l = l0 + q;
    if (l < r && j < c) dst[(long)j + (long)l * ldd] = tile[tx][q];
  }
}

}  // namespace

extern "C" int fla_hip_dgetrf_nopiv(int m, int n, double* A, int lda, int* info) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  *info = kSuccess;
  const int mn = m < n ? m : n;
  if (mn == 0) return FLA_HIP_OK;
  FlaHipCtx& c = fla_ctx();
  hipStream_t s = c.s_update;
  // workspace: U12 transposed (n x 2 TB)
  void* w = nullptr;
  FLA_TRY(fla_workspace((size_t)n * 2 * TB * sizeof(double), &w));
  double* Ut = static_cast<double*>(w);
  const long ldu = n;
  auto at = [&](int i, int j) { return A + (long)i + (long)j * lda; };
  // 256-column blocks: the (m - k) x kb panel by 128-column steps (their
  // updates confined to the panel), then U12 := inv(L11) A12 for the whole
  // block (two substitutions around B_b -= L_ba X_a) and one K = 256 update.
  for (int k = 0; k < mn; k += 2 * TB) {
    const int kb = (mn - k) < 2 * TB ? (mn - k) : 2 * TB;
    for (int j = k; j < k + kb; j += TB) {
      const int jb = (k + kb - j) < TB ? (k + kb - j) : TB;
      const int nc = k + kb - j - jb, mr = m - j - jb;
      hipLaunchKernelGGL(lu_nopiv_diag_kernel, dim3(1), dim3(256), 0, s, jb, at(j, j), (long)lda);
      FLA_HIP_CHECK(hipGetLastError());
      if (nc > 0) FLA_TRY(launch_trsm_llnu_subst(s, nc, jb, at(j, j), lda, at(j, j + jb), lda));
      if (mr > 0) FLA_TRY(launch_trsm_runn_rows(s, mr, jb, at(j, j), lda, at(j + jb, j), lda));
      if (mr > 0 && nc > 0)
        FLA_TRY(launch_dgemm_mfma(s, false, false, 0, mr, nc, jb, -1.0, at(j + jb, j), lda,
                                  at(j, j + jb), lda, 1.0, at(j + jb, j + jb), lda, nullptr));
    }
    const int nc = n - k - kb, mr = m - k - kb;
    if (nc <= 0) continue;
    const int ka = kb < TB ? kb : TB, kbb = kb - ka;
    FLA_TRY(launch_trsm_llnu_subst(s, nc, ka, at(k, k), lda, at(k, k + kb), lda));
    if (kbb > 0) {
      FLA_TRY(launch_dgemm_mfma(s, false, false, 0, kbb, nc, ka, -1.0, at(k + ka, k), lda,
                                at(k, k + kb), lda, 1.0, at(k + ka, k + kb), lda, nullptr));
      FLA_TRY(launch_trsm_llnu_subst(s, nc, kbb, at(k + ka, k + ka), lda, at(k + ka, k + kb), lda));
    }
    if (mr > 0) {
      hipLaunchKernelGGL(transpose_nopiv_kernel, dim3((kb + 31) / 32, (nc + 31) / 32), dim3(256),
                         0, s, kb, nc, at(k, k + kb), (long)lda, Ut, ldu);
      FLA_HIP_CHECK(hipGetLastError());
      FLA_TRY(launch_dgemm_mfma(s, false, true, 0, mr, nc, kb, -1.0, at(k + kb, k), lda, Ut, ldu,
                                1.0, at(k + kb, k + kb), lda, nullptr));
    }
  }
  int* d = c.dscratch + 8;
  hipLaunchKernelGGL(set_int_max_kernel, dim3(1), dim3(1), 0, s, d);
  hipLaunchKernelGGL(zero_diag_kernel, dim3((mn + 255) / 256), dim3(256), 0, s, mn, A, (long)lda,
                     d);
  FLA_HIP_CHECK(hipGetLastError());
  int h = 0x7fffffff;
  FLA_HIP_CHECK(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  *info = (h == 0x7fffffff) ? kSuccess : h;
  return FLA_HIP_OK;
}
