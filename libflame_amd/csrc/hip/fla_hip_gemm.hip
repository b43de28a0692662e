// FP64 MFMA trailing-update kernel for gfx950 (CDNA4).
//
// One kernel serves every level-3 update on the hot path:
//   * FLA_Chol_l_blk_var3's A22 := A22 - A21*A21'   (lower triangle only)
//     src/lapack/dec/chol/l/flamec/FLA_Chol_l_blk_var3.c:52-54 -> dsyrk_('L','N')
//   * FLA_LU_piv_blk_var5's A22 := A22 - A21*A12
//     src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_blk_var5.c:94-96 -> dgemm_('N','N')
//   * the FLA_Gemm / FLA_Syrk entry points of the thin C ABI.
//
// Tiling (default configuration): a 512-thread workgroup (8 wave64s in a 4x2
// grid) owns a 128x128 tile of C; each wave owns 32x64 = 2x4 blocks of
// v_mfma_f64_16x16x4_f64.  K is streamed in chunks of 16 (8 when B is
// k-contiguous) through two LDS buffers (load chunk t+1 into registers while
// chunk t feeds the MFMAs, then one barrier per chunk).
// The MFMA is issued with the B fragment as its "A" operand so that each
// lane's four accumulators are four columns of one row index: for fixed
// accumulator slot, lanes 0..15 hold 16 consecutive rows of one column, and
// the epilogue stores are 128-byte contiguous runs of column-major C.
#include "fla_hip_internal.h"

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

// Loads the BM x BK chunk of op(X) starting at (r0, k0) into registers
// (NT threads, PER = BM*BK/NT values each).
//   T == false: op(X)(r,l) = X[r + l*ld]   (rows contiguous: lanes walk r)
//   T == true : op(X)(r,l) = X[l + r*ld]   (k contiguous: lanes walk l)
template <bool T, int BM, int BK, int NT>
__device__ __forceinline__ void load_chunk(const double* X, long ld, int rows, int kdim, int r0,
                                           int k0, double (&reg)[BM * BK / NT], int tid) {
  constexpr int PER = BM * BK / NT;
  const bool full = (r0 + BM <= rows) && (k0 + BK <= kdim);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + q * NT;
    const int r = T ? (e / BK) : (e % BM);
    const int l = T ? (e % BK) : (e / BM);
    const double* p = T ? X + (long)(k0 + l) + (long)(r0 + r) * ld
                        : X + (long)(r0 + r) + (long)(k0 + l) * ld;
    if (full)
      reg[q] = *p;
    else
      reg[q] = (r0 + r < rows && k0 + l < kdim) ? *p : 0.0;
  }
}

// Stores registers from load_chunk into LDS laid out as S[l][r] (stride BM+1).
template <bool T, int BM, int BK, int NT>
__device__ __forceinline__ void store_chunk(double* S, const double (&reg)[BM * BK / NT], int tid) {
  constexpr int PER = BM * BK / NT;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + q * NT;
    const int r = T ? (e / BK) : (e % BM);
    const int l = T ? (e % BK) : (e / BM);
    S[l * (BM + 1) + r] = reg[q];
  }
}

// C := alpha*op(A)*op(B) + beta*C on a BM x BN tile per workgroup of
// WM x WN waves; each wave owns (BM/WM) x (BN/WN) = TM x TN MFMA blocks.
template <bool TA, bool TB, int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN)
__attribute__((amdgpu_waves_per_eu(WM * WN >= 8 ? 4 : 2))) void dgemm_mfma_kernel(
    int mode, int m, int n, int k, double alpha, const double* A, long lda, const double* B,
    long ldb, double beta, double* C, long ldc, int Tm, int Tn, int rpc,
    const int* __restrict__ abort_flag) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int LA = BM + 1, LB = BN + 1;
  constexpr int PA = BM * BK / NT, PB = BN * BK / NT;
  if (abort_flag && *abort_flag >= 0) return;
  __shared__ double As[2][BK * LA];
  __shared__ double Bs[2][BK * LB];

  int ti, tj;
  fla_tile_of(blockIdx.x, gridDim.x, mode, Tm, Tn, rpc, ti, tj);
  const int i0 = ti * BM, j0 = tj * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;

  v4d acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  const int nk = (k + BK - 1) / BK;
  double ra[PA], rb[PB];
  if (nk > 0) {
    load_chunk<TA, BM, BK, NT>(A, lda, m, k, i0, 0, ra, tid);
    load_chunk<!TB, BN, BK, NT>(B, ldb, n, k, j0, 0, rb, tid);
    store_chunk<TA, BM, BK, NT>(As[0], ra, tid);
    store_chunk<!TB, BN, BK, NT>(Bs[0], rb, tid);
  }
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      load_chunk<TA, BM, BK, NT>(A, lda, m, k, i0, (kt + 1) * BK, ra, tid);
      load_chunk<!TB, BN, BK, NT>(B, ldb, n, k, j0, (kt + 1) * BK, rb, tid);
    }
    const double* as = As[cur];
    const double* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      double fa[TM], fb[TN];
#pragma unroll
      for (int t = 0; t < TM; ++t) fa[t] = as[(kk * 4 + lk) * LA + wr * (BM / WM) + t * 16 + li];
#pragma unroll
      for (int t = 0; t < TN; ++t) fb[t] = bs[(kk * 4 + lk) * LB + wc * (BN / WN) + t * 16 + li];
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b], fa[a], acc[a][b], 0, 0, 0);
    }
    if (more) {
      store_chunk<TA, BM, BK, NT>(As[cur ^ 1], ra, tid);
      store_chunk<!TB, BN, BK, NT>(Bs[cur ^ 1], rb, tid);
    }
    __syncthreads();
  }

  // Epilogue.  v_mfma_f64_16x16x4f64 leaves D(x, y) in lane 16*(x%4) + y,
  // register x/4 (measured on gfx950); with x = column, y = row here, lane
  // holds C(i0+wr*BM/WM+a*16+li, j0+wc*BN/WN+b*16+lk+4*r) in acc[a][b][r].
  // Interior tiles (inside C and, in lower mode, strictly below the
  // diagonal) take a branch-free path: per 16-row block a and group of EB
  // column blocks every C value the lane needs is loaded first and only then
  // combined and stored, so the loads are in flight together (one memory
  // round trip per group instead of one per element; the group size keeps
  // the kernel within 128 VGPRs).  Edge and diagonal tiles store through
  // per-element guards.  Addresses are the tile's base (uniform) plus a
  // 32-bit lane offset.
  constexpr int EB = TN >= 2 ? 2 : 1;
  double* Ct = C + (long)i0 + (long)j0 * ldc;
  const int ldci = (int)ldc;
  // (32-bit lane offsets: the tile spans fewer than 2^31 elements)
  const bool interior = (i0 + BM <= m) && (j0 + BN <= n) && (mode != 1 || i0 >= j0 + BN - 1) &&
                        ldc < (1L << 23);
  if (interior) {
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const int di = wr * (BM / WM) + a * 16 + li;
#pragma unroll
      for (int b0 = 0; b0 < TN; b0 += EB) {
        double cv[EB][4];
        if (beta != 0.0) {
#pragma unroll
          for (int b = 0; b < EB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              cv[b][r] = Ct[(unsigned)(di + (wc * (BN / WN) + (b0 + b) * 16 + lk + 4 * r) * ldci)];
        }
#pragma unroll
        for (int b = 0; b < EB; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double v = alpha * acc[a][b0 + b][r];
            Ct[(unsigned)(di + (wc * (BN / WN) + (b0 + b) * 16 + lk + 4 * r) * ldci)] =
                (beta == 0.0) ? v : v + beta * cv[b][r];
          }
      }
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < TM; ++a) {
    const int di = wr * (BM / WM) + a * 16 + li;
    const int i = i0 + di;
#pragma unroll
    for (int b = 0; b < TN; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dj = wc * (BN / WN) + b * 16 + lk + 4 * r;
        const int j = j0 + dj;
        if (i >= m || j >= n || (mode == 1 && i < j)) continue;
        double* c = Ct + (long)di + (long)dj * ldc;
        const double v = alpha * acc[a][b][r];
        *c = (beta == 0.0) ? v : v + beta * (*c);
      }
    }
  }
}

// Tile configurations (selected by fla_hip_set_gemm_config for tuning).
int g_gemm_cfg = 0;

template <int BM, int BN, int BK, int WM, int WN>
int launch_cfg(hipStream_t s, bool ta, bool tb, int mode, int m, int n, int k, double alpha,
               const double* A, long lda, const double* B, long ldb, double beta, double* C,
               long ldc, const int* abort_flag) {
  static_assert(BM == BN, "the lower-tile enumeration assumes square tiles");
  const int Tm = (m + BM - 1) / BM, Tn = (n + BN - 1) / BN;
  const int rpc_eff = g_tile_group;  // tile-column group width (fla_tile_of)
  const long nblk = (mode == 0) ? (long)Tm * Tn : (long)Tn * Tm - (long)Tn * (Tn - 1) / 2;
  if (nblk <= 0) return FLA_HIP_OK;
  if (nblk > 0x7fffffffL) return FLA_HIP_EARG;
  dim3 grid((unsigned)nblk), block(64 * WM * WN);
#define FLA_GEMM_LAUNCH(TA_, TB_)                                                              \
  hipLaunchKernelGGL((dgemm_mfma_kernel<TA_, TB_, BM, BN, BK, WM, WN>), grid, block, 0, s,     \
                     mode, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, Tm, Tn, rpc_eff,       \
                     abort_flag)
  if (!ta && !tb)
    FLA_GEMM_LAUNCH(false, false);
  else if (!ta && tb)
    FLA_GEMM_LAUNCH(false, true);
  else if (ta && !tb)
    FLA_GEMM_LAUNCH(true, false);
  else
    FLA_GEMM_LAUNCH(true, true);
#undef FLA_GEMM_LAUNCH
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

}  // namespace

extern "C" int fla_hip_set_gemm_group(int G) {
  if (G < 1 || G > 64) return FLA_HIP_EARG;
  g_tile_group = G;
  return FLA_HIP_OK;
}

extern "C" int fla_hip_set_gemm_config(int cfg) {
  if (cfg < 0 || cfg > 7) return FLA_HIP_EARG;
  g_gemm_cfg = cfg;
  return FLA_HIP_OK;
}

int launch_dgemm_mfma(hipStream_t s, bool ta, bool tb, int mode, int m, int n, int k,
                      double alpha, const double* A, long lda, const double* B, long ldb,
                      double beta, double* C, long ldc, const int* abort_flag) {
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  if (mode == 1 && m < n) return FLA_HIP_EARG;  // lower mode expects a tall C
  switch (g_gemm_cfg) {
    case 1:
      return launch_cfg<128, 128, 32, 2, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
    case 2:
      return launch_cfg<128, 128, 16, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
    case 3:
      return launch_cfg<128, 128, 32, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
    case 4:
      return launch_cfg<128, 128, 16, 2, 4>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
    case 5:
      return launch_cfg<128, 128, 16, 4, 4>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
    case 6:
      return launch_cfg<128, 128, 8, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                           C, ldc, abort_flag);
    case 7:
      return launch_cfg<128, 128, 16, 2, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
    default:
      // Measured on MI355X (n = 16384..32768, k = 128..256): 8 waves per
      // 128x128 tile; K chunks of 16 when B is stored with j contiguous (the
      // SYRK update, ~52 TFLOP/s), of 8 when B is k-contiguous (the LU
      // update, ~39 TFLOP/s), where the smaller LDS footprint doubles occupancy.
      if (!tb)
        return launch_cfg<128, 128, 8, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                             beta, C, ldc, abort_flag);
      return launch_cfg<128, 128, 16, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
  }
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
    // C := beta*C without reading A or B (src/flablas/f2c/dgemm.c:194-214):
    // run with k = 0 (the K loop is skipped, the epilogue applies beta and
    // writes 0 where beta == 0).
    k = 0;
    alpha = 0.0;
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
  if (k == 0 || alpha == 0.0) {
    if (beta == 1.0) return FLA_HIP_OK;
    k = 0;  // C := beta*C on the triangle without reading A (dsyrk.c's quick return)
    alpha = 0.0;
  }
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
