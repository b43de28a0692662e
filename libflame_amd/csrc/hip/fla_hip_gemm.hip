// FP64 MFMA trailing-update kernel for gfx950 (CDNA4).
//
// One kernel serves every level-3 update on the hot path:
//   * FLA_Chol_l_blk_var3's A22 := A22 - A21*A21'   (lower triangle only)
//     src/lapack/dec/chol/l/flamec/FLA_Chol_l_blk_var3.c:52-54 -> dsyrk_('L','N')
//   * FLA_LU_piv_blk_var5's A22 := A22 - A21*A12
//     src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_blk_var5.c:94-96 -> dgemm_('N','N')
//   * the FLA_Gemm / FLA_Syrk entry points of the thin C ABI.
//
// Tiling: a 256-thread workgroup (4 wave64s in a 2x2 grid) owns a 128x128 tile
// of C; each wave owns 64x64 = 4x4 blocks of v_mfma_f64_16x16x4_f64.  K is
// streamed in chunks of 16 through two LDS buffers (load chunk t+1 into
// registers while chunk t feeds the MFMAs, then one barrier per chunk).
// The MFMA is issued with the B fragment as its "A" operand so that each
// lane's four accumulators are four columns of one row index: for fixed
// accumulator slot, lanes 0..15 hold 16 consecutive rows of one column, and
// the epilogue stores are 128-byte contiguous runs of column-major C.
#include "fla_hip_internal.h"

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int NTHR = 256;
constexpr int LS = BM + 1;  // LDS row stride (doubles); the pad spreads transposed stores

// Loads the 128 x 16 chunk of op(X) starting at (r0, k0) into registers.
//   T == false: op(X)(r,l) = X[r + l*ld]   (rows contiguous)
//   T == true : op(X)(r,l) = X[l + r*ld]   (k contiguous)
template <bool T>
__device__ __forceinline__ void load_chunk(const double* X, long ld, int rows,
                                           int kdim, int r0, int k0, double (&reg)[8], int tid) {
  const bool full = (r0 + BM <= rows) && (k0 + BK <= kdim);
  if (!T) {
    const int r = tid & (BM - 1);
    const int lb = tid >> 7;  // 0..1
    const double* p = X + (long)(r0 + r) + (long)(k0 + lb) * ld;
    if (full) {
#pragma unroll
      for (int q = 0; q < 8; ++q) reg[q] = p[(long)(2 * q) * ld];
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int l = lb + 2 * q;
        reg[q] = (r0 + r < rows && k0 + l < kdim) ? p[(long)(2 * q) * ld] : 0.0;
      }
    }
  } else {
    const int l = tid & (BK - 1);
    const int rb = tid >> 4;  // 0..15
    const double* p = X + (long)(k0 + l) + (long)(r0 + rb) * ld;
    if (full) {
#pragma unroll
      for (int q = 0; q < 8; ++q) reg[q] = p[(long)(16 * q) * ld];
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int r = rb + 16 * q;
        reg[q] = (r0 + r < rows && k0 + l < kdim) ? p[(long)(16 * q) * ld] : 0.0;
      }
    }
  }
}

// Stores registers from load_chunk into LDS laid out as S[l][r] (stride LS).
template <bool T>
__device__ __forceinline__ void store_chunk(double* S, const double (&reg)[8], int tid) {
  if (!T) {
    const int r = tid & (BM - 1);
    const int lb = tid >> 7;
#pragma unroll
    for (int q = 0; q < 8; ++q) S[(lb + 2 * q) * LS + r] = reg[q];
  } else {
    const int l = tid & (BK - 1);
    const int rb = tid >> 4;
#pragma unroll
    for (int q = 0; q < 8; ++q) S[l * LS + rb + 16 * q] = reg[q];
  }
}

// Maps a linear workgroup id to a C tile.  Workgroups are dealt to the eight
// XCDs round-robin; the id is first permuted so that each XCD receives a
// contiguous run of tiles (which share operand panels in its own L2).
// mode 0: column-major over a Tm x Tn grid; mode 1: lower tiles (ti >= tj) of
// a Tm x Tn grid, column by column.
__device__ __forceinline__ void tile_of(int bid, int nblk, int mode, int Tm, int Tn, int& ti,
                                        int& tj) {
  const int x = bid & 7;
  const int q = nblk >> 3, rr = nblk & 7;
  const int t = x * q + (x < rr ? x : rr) + (bid >> 3);
  if (mode == 0) {
    tj = t / Tm;
    ti = t - tj * Tm;
  } else {
    // offset(c) = c*Tm - c*(c-1)/2 tiles precede column c.
    const double a = 2.0 * Tm + 1.0;
    int c = (int)((a - sqrt(a * a - 8.0 * (double)t)) * 0.5);
    if (c < 0) c = 0;
    if (c > Tn - 1) c = Tn - 1;
    while (c > 0 && (long)c * Tm - (long)c * (c - 1) / 2 > t) --c;
    while (c + 1 < Tn && (long)(c + 1) * Tm - (long)(c + 1) * c / 2 <= t) ++c;
    tj = c;
    ti = c + (t - (c * Tm - c * (c - 1) / 2));
  }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(NTHR) void dgemm_mfma_kernel(
    int mode, int m, int n, int k, double alpha, const double* A, long lda,
    const double* B, long ldb, double beta, double* C, long ldc,
    int Tm, int Tn, const int* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag >= 0) return;
  __shared__ double As[2][BK * LS];
  __shared__ double Bs[2][BK * LS];

  int ti, tj;
  tile_of(blockIdx.x, gridDim.x, mode, Tm, Tn, ti, tj);
  const int i0 = ti * BM, j0 = tj * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;

  v4d acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  const int nk = (k + BK - 1) / BK;
  double ra[8], rb[8];
  if (nk > 0) {
    load_chunk<TA>(A, lda, m, k, i0, 0, ra, tid);
    load_chunk<!TB>(B, ldb, n, k, j0, 0, rb, tid);
    store_chunk<TA>(As[0], ra, tid);
    store_chunk<!TB>(Bs[0], rb, tid);
  }
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      load_chunk<TA>(A, lda, m, k, i0, (kt + 1) * BK, ra, tid);
      load_chunk<!TB>(B, ldb, n, k, j0, (kt + 1) * BK, rb, tid);
    }
    const double* as = As[cur];
    const double* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int row = (kk * 4 + lk) * LS;
      double fa[4], fb[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        fa[t] = as[row + wr * 64 + t * 16 + li];
        fb[t] = bs[row + wc * 64 + t * 16 + li];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b], fa[a], acc[a][b], 0, 0, 0);
    }
    if (more) {
      store_chunk<TA>(As[cur ^ 1], ra, tid);
      store_chunk<!TB>(Bs[cur ^ 1], rb, tid);
    }
    __syncthreads();
  }

  // Epilogue.  v_mfma_f64_16x16x4f64 leaves D(x, y) in lane 16*(x%4) + y,
  // register x/4 (measured on gfx950); with x = column, y = row here, lane
  // holds C(i0+wr*64+a*16+li, j0+wc*64+b*16+lk+4*r) in acc[a][b][r].
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + wr * 64 + a * 16 + li;
    if (i >= m) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = j0 + wc * 64 + b * 16 + lk + 4 * r;
        if (j >= n) continue;
        if (mode == 1 && i < j) continue;
        double* c = C + (long)i + (long)j * ldc;
        const double v = alpha * acc[a][b][r];
        *c = (beta == 0.0) ? v : v + beta * (*c);
      }
    }
  }
}

}  // namespace

int launch_dgemm_mfma(hipStream_t s, bool ta, bool tb, int mode, int m, int n, int k,
                      double alpha, const double* A, long lda, const double* B, long ldb,
                      double beta, double* C, long ldc, const int* abort_flag) {
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  if (mode == 1 && m < n) return FLA_HIP_EARG;  // lower mode expects a tall C
  const int Tm = (m + BM - 1) / BM, Tn = (n + BN - 1) / BN;
  long nblk;
  if (mode == 0)
    nblk = (long)Tm * Tn;
  else
    nblk = (long)Tn * Tm - (long)Tn * (Tn - 1) / 2;
  if (nblk <= 0) return FLA_HIP_OK;
  if (nblk > 0x7fffffffL) return FLA_HIP_EARG;
  dim3 grid((unsigned)nblk), block(NTHR);
  if (!ta && !tb)
    hipLaunchKernelGGL((dgemm_mfma_kernel<false, false>), grid, block, 0, s, mode, m, n, k,
                       alpha, A, lda, B, ldb, beta, C, ldc, Tm, Tn, abort_flag);
  else if (!ta && tb)
    hipLaunchKernelGGL((dgemm_mfma_kernel<false, true>), grid, block, 0, s, mode, m, n, k,
                       alpha, A, lda, B, ldb, beta, C, ldc, Tm, Tn, abort_flag);
  else if (ta && !tb)
    hipLaunchKernelGGL((dgemm_mfma_kernel<true, false>), grid, block, 0, s, mode, m, n, k,
                       alpha, A, lda, B, ldb, beta, C, ldc, Tm, Tn, abort_flag);
  else
    hipLaunchKernelGGL((dgemm_mfma_kernel<true, true>), grid, block, 0, s, mode, m, n, k,
                       alpha, A, lda, B, ldb, beta, C, ldc, Tm, Tn, abort_flag);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

static bool is_trans(int t) { return t == kTrans || t == kConjTrans; }

extern "C" int fla_hip_dgemm(int transa, int transb, int m, int n, int k, double alpha,
                             const double* A, int lda, const double* B, int ldb, double beta,
                             double* C, int ldc) {
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || k < 0) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const bool ta = is_trans(transa), tb = is_trans(transb);
  if (ldc < m || lda < (ta ? k : m) || ldb < (tb ? n : k)) return FLA_HIP_EARG;
  if (k == 0 || alpha == 0.0) {
    if (beta == 1.0) return FLA_HIP_OK;
    // k == 0 still scales C: run with k = 0 (loop skipped, epilogue applies beta).
  }
  return launch_dgemm_mfma(fla_ctx().s_update, ta, tb, 0, m, n, k, alpha, A, lda, B, ldb,
                           beta, C, ldc, nullptr);
}

extern "C" int fla_hip_dsyrk(int uplo, int trans, int n, int k, double alpha,
                             const double* A, int lda, double beta, double* C, int ldc) {
  FLA_TRY(fla_ensure_ready());
  if (n < 0 || k < 0) return FLA_HIP_EARG;
  if (n == 0) return FLA_HIP_OK;
  if (uplo != kLower && uplo != kUpper) return FLA_HIP_EARG;
  const bool t = is_trans(trans);
  if (ldc < n || lda < (t ? k : n)) return FLA_HIP_EARG;
  if ((k == 0 || alpha == 0.0) && beta == 1.0) return FLA_HIP_OK;
  hipStream_t s = fla_ctx().s_update;
  if (uplo == kLower) {
    // C(lower) += alpha * op(A) op(A)'   with op(A) n x k.
    return launch_dgemm_mfma(s, t, !t, 1, n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                             nullptr);
  }
  // Upper triangle of C is the lower triangle of C' (same storage viewed with
  // swapped strides); C(i,j) for i <= j equals C'(j,i): compute lower of
  // C' = alpha * op(A) op(A)' + beta*C'  ==  identical values, transposed
  // storage.  Exchange triangles, update the lower one, exchange back.
  FLA_TRY(launch_transpose_swap(s, n, C, ldc));
  FLA_TRY(launch_dgemm_mfma(s, t, !t, 1, n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                            nullptr));
  return launch_transpose_swap(s, n, C, ldc);
}
