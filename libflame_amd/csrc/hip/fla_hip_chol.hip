// Blocked Cholesky on one MI355X: diagonal-block factor, panel TRSM and the
// MFMA trailing update, scheduled with one panel of look-ahead on two streams.
//
// Reference algorithm: FLA_Chol_l_blk_var3
//   (src/lapack/dec/chol/l/flamec/FLA_Chol_l_blk_var3.c:13-66):
//     A11 := chol(A11); A21 := A21 * inv(tril(A11))'; A22 := A22 - A21*A21'
// and its leaf FLA_Chol_l_opd_var2 (src/lapack/dec/chol/l/flamec/FLA_Chol_l_opt_var2.c:136-193),
// whose failure rule (pivot <= 0 or NaN, bl1_dsqrte in
// src/base/flamec/include/blis_macro_defs.h:134-136) and 0-based failing
// column index (FLA_Chol_l_blk_var3.c:42-43) are kept.  The upper case
// (FLA_Chol_u_blk_var3.c:38-54) is the lower algorithm on A' (the triangles
// are exchanged in place before and after).
#include "fla_hip_internal.h"
#include "fla_hip_lds_tri.h"

namespace {

constexpr int PB = 128;  // panel sub-block width (diagonal block order)
constexpr int LB = 32;   // leaf width of the panel triangular solve

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int DLD = PB + 1;  // LDS leading dimension of the diagonal block

// Single-workgroup Cholesky of the nb x nb lower block at A (nb <= 128),
// blocked by 32 inside the workgroup (the reference's var3 structure one
// level down, FLA_Chol_l_blk_var3.c:13-66):
//   1. wave 0 factors the 32x32 diagonal block, lane = row, pivots broadcast
//      with readlane (no barriers): d = sqrt(a_jj) with bl1_dsqrte's failure
//      rule, l_ij = a_ij * (1/d), a_ic -= l_ij l_cj;
//   2. the rows below are solved against it (lane = row, row in VGPRs, the
//      32x32 factor read from LDS by broadcast), dtrsm's update order;
//   3. the trailing part is updated on MFMA (16x16 lower tiles, rank 32).
// The block lives in LDS (129 KB); padding beyond nb is the identity.
// INV: after a successful factor, also writes inv(L) (lower, ld 128, zero
// upper part) to W, inverted in LDS (fla_hip_lds_tri.h).
template <bool INV>
__global__ __launch_bounds__(256) void potrf_blk_kernel(int nb, double* __restrict__ A, long lda,
                                                        int col_offset, int* __restrict__ info,
                                                        double* __restrict__ W) {
  if (*info >= 0) return;
  __shared__ double S[PB * DLD];
  __shared__ int fail;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  // 64 elements per thread, 16 loads in flight at a time
  for (int q0 = 0; q0 < PB * PB / 256; q0 += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = t + (q0 + q) * 256;
      const int i = e & (PB - 1), c = e >> 7;
      v[q] = (i < nb && c < nb) ? A[(long)i + (long)c * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = t + (q0 + q) * 256;
      const int i = e & (PB - 1), c = e >> 7;
      S[c * DLD + i] = (i < c) ? 0.0 : ((i < nb) ? v[q] : (i == c ? 1.0 : 0.0));
    }
  }
  if (t == 0) fail = -1;
  __syncthreads();
  for (int b = 0; b < PB; b += LB) {
    // 1. diagonal 32x32 block: wave 0, lanes 0..31
    if (wave == 0) {
      const int r = b + (lane & 31);
      double x[LB];
#pragma unroll
      for (int c = 0; c < LB; ++c) x[c] = S[(b + c) * DLD + r];
      int f = -1;
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        if (f >= 0) continue;  // uniform: a failed column stops the sweep
        const double ajj = readlane_d(x[j], j);
        if (!(ajj > 0.0)) {  // <= 0 or NaN, as bl1_dsqrte
          f = j;
          continue;
        }
        const double d = sqrt(ajj);
        const double inv = 1.0 / d;
        if ((lane & 31) > j) x[j] *= inv;
        if ((lane & 31) == j) x[j] = d;
        // entries above the diagonal (lane < c) take garbage here and are
        // never stored or read, so the update needs no per-lane predicate
#pragma unroll
        for (int c = j + 1; c < LB; ++c) x[c] -= x[j] * readlane_d(x[j], c);
      }
      if (lane < 32) {
#pragma unroll
        for (int c = 0; c < LB; ++c)
          if (lane >= c && (f < 0 || c < f)) S[(b + c) * DLD + r] = x[c];
      }
      if (lane == 0 && f >= 0) fail = b + f;
    }
    __syncthreads();
    if (fail >= 0) break;
    // 2. rows below the diagonal block: X := X * inv(L_bb)'
    const int nrow = PB - b - LB;
    if (t < nrow) {
      const int r = b + LB + t;
      double x[LB];
#pragma unroll
      for (int c = 0; c < LB; ++c) x[c] = S[(b + c) * DLD + r];
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        x[j] *= 1.0 / S[(b + j) * DLD + b + j];
#pragma unroll
        for (int c = j + 1; c < LB; ++c) x[c] -= S[(b + j) * DLD + b + c] * x[j];
      }
#pragma unroll
      for (int c = 0; c < LB; ++c) S[(b + c) * DLD + r] = x[c];
    }
    __syncthreads();
    // 3. trailing update A(i, c) -= sum_k L(i, k) L(c, k), i >= c >= b + 32,
    //    on v_mfma_f64_16x16x4f64: 16 x 16 lower tiles spread over the waves;
    //    lane (li, lk), register q holds D(16 ti + li, 16 tj + lk + 4 q).
    {
      const int T16 = nrow / 16;
      const int ntile = T16 * (T16 + 1) / 2;
      const int i0 = b + LB;
      const int li = lane & 15, lk = lane >> 4;
      for (int tt = wave; tt < ntile; tt += 4) {
        int ti = (int)((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
        while (ti * (ti + 1) / 2 > tt) --ti;
        while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
        const int tj = tt - ti * (ti + 1) / 2;
        fla_lds::d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < LB / 4; ++ks) {
          const double* col = S + (b + 4 * ks + lk) * DLD;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(col[i0 + 16 * tj + li], col[i0 + 16 * ti + li],
                                                     acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = i0 + 16 * ti + li, cc = i0 + 16 * tj + lk + 4 * q;
          if (row >= cc) S[cc * DLD + row] -= acc[q];
        }
      }
    }
    __syncthreads();
  }
  const int ncols = (fail >= 0) ? fail : nb;
  for (int e = t; e < PB * PB; e += 256) {
    const int i = e & (PB - 1), c = e >> 7;
    if (i >= c && i < nb && c < ncols) A[(long)i + (long)c * lda] = S[c * DLD + i];
  }
  if (t == 0 && fail >= 0) *info = col_offset + fail;
  if (INV && fail < 0) {
    // the strictly upper part of S is zero and the padding the identity
    __syncthreads();
    fla_lds::trtri_lower_lds(S, t);
    for (int e = t; e < PB * PB; e += 256) {
      const int i = e & (PB - 1), c = e >> 7;
      W[i + c * PB] = (i >= c) ? S[c * DLD + i] : 0.0;
    }
  }
}

// Leaf of the panel solve: X := X * inv(L)' for an m x nb strip (nb <= 32),
// L lower non-unit.  One lane per row with the row in VGPRs, L through
// uniform loads; the update order of dtrsm_('R','L','T','N')
// (src/flablas/f2c/dtrsm.c): x_j *= 1/l_jj, then x_c -= l_cj * x_j.
template <bool FULL>
__global__ __launch_bounds__(64) void trsm_rlt_leaf_kernel(int m, int nb,
                                                           const double* __restrict__ L, long ldl,
                                                           double* __restrict__ X, long ldx,
                                                           const int* __restrict__ abort_flag) {
  if (abort_flag && *abort_flag >= 0) return;
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= m) return;
  double x[LB];
#pragma unroll
  for (int c = 0; c < LB; ++c) x[c] = (FULL || c < nb) ? X[(long)r + (long)c * ldx] : 0.0;
#pragma unroll
  for (int j = 0; j < LB; ++j) {
    if (FULL || j < nb) {
      x[j] *= 1.0 / L[(long)j + (long)j * ldl];
#pragma unroll
      for (int c = j + 1; c < LB; ++c)
        if (FULL || c < nb) x[c] -= L[(long)c + (long)j * ldl] * x[j];
    }
  }
#pragma unroll
  for (int c = 0; c < LB; ++c)
    if (FULL || c < nb) X[(long)r + (long)c * ldx] = x[c];
}

// Exchanges A(i,j) and A(j,i) for all i > j, 32x32 tile pairs through LDS.
__global__ __launch_bounds__(256) void transpose_swap_kernel(int n, double* __restrict__ A, long lda) {
  __shared__ double P[32][33];
  __shared__ double Q[32][33];
  // Lower tile (I, J), I >= J, enumerated row by row.
  const int b = blockIdx.x;
  int I = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((long)I * (I + 1) / 2 > b) --I;
  while ((long)(I + 1) * (I + 2) / 2 <= b) ++I;
  const int J = b - I * (I + 1) / 2;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty 0..7
  const int ri = I * 32, cj = J * 32;
  for (int y = ty; y < 32; y += 8) {
    const int i = ri + tx, j = cj + y;  // lower tile element (i, j)
    P[y][tx] = (i < n && j < n) ? A[(long)i + (long)j * lda] : 0.0;
    const int i2 = cj + tx, j2 = ri + y;  // upper tile element (i2, j2)
    Q[y][tx] = (i2 < n && j2 < n) ? A[(long)i2 + (long)j2 * lda] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    // new lower (ri+tx, cj+y) := old upper (cj+y, ri+tx) = Q[tx][y]
    const int i = ri + tx, j = cj + y;
    if (i < n && j < n && i > j) A[(long)i + (long)j * lda] = Q[tx][y];
    const int i2 = cj + tx, j2 = ri + y;
    if (i2 < n && j2 < n && i2 < j2) A[(long)i2 + (long)j2 * lda] = P[tx][y];
  }
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

}  // namespace

bool g_trsm_rows = true;

int launch_potrf_diag(hipStream_t s, int nb, double* A, long lda, int col_offset, int* info) {
  if (nb <= 0) return FLA_HIP_OK;
  if (nb > PB) return FLA_HIP_EARG;
  hipLaunchKernelGGL(potrf_blk_kernel<false>, dim3(1), dim3(256), 0, s, nb, A, lda, col_offset,
                     info, nullptr);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// The same, and W (128 x 128, ld 128) := inv(L) when the factor succeeds.
int launch_potrf_diag_inv(hipStream_t s, int nb, double* A, long lda, int col_offset, int* info,
                          double* W) {
  if (nb <= 0) return FLA_HIP_OK;
  if (nb > PB) return FLA_HIP_EARG;
  hipLaunchKernelGGL(potrf_blk_kernel<true>, dim3(1), dim3(256), 0, s, nb, A, lda, col_offset,
                     info, W);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// A21 := A21 * inv(L11)' for an m x nb strip, L11 lower non-unit (nb <= 128):
// one launch, one thread per row of the strip.  Columns are solved in blocks
// of 32 held in VGPRs: first the contributions of the already solved columns
// k (ascending, re-read by the thread that wrote them), then the 32 x 32
// diagonal block by substitution -- dtrsm_('R','L','T','N')'s update order
// (src/flablas/f2c/dtrsm.c), temp = 1/a_kk multiplied in.  L's columns
// below the diagonal blocks are read with wave-uniform (scalar) loads, 32
// at a time; the four 32 x 32 diagonal blocks, whose substitution would
// otherwise wait on dozens of dependent scalar loads each, are first copied
// into LDS (16.5 KB packed, small enough to share a CU with the update
// kernel).
template <bool FULL>
__global__ __launch_bounds__(256) void trsm_rlt_rows_kernel(int m, int nb,
                                                            const double* __restrict__ L,
                                                            long ldl, double* __restrict__ X,
                                                            long ldx, const int* __restrict__ abort) {
  // diagonal block b, lower triangle packed by columns: (r, c), r >= c, at
  // Ld[b][c*LB - c*(c-1)/2 + (r - c)]
  __shared__ double Ld[PB / LB][LB * (LB + 1) / 2];
  if (abort && *abort >= 0) return;
  const int nbk = FULL ? PB : nb;
  {
    // 16 elements per thread, all loads issued before the LDS stores
    double v[(PB / LB) * LB * LB / 256];
#pragma unroll
    for (int q = 0; q < (PB / LB) * LB * LB / 256; ++q) {
      const int e = threadIdx.x + 256 * q;
      const int b = e / (LB * LB), r = e % LB, c = (e / LB) % LB;
      const int gr = b * LB + r, gc = b * LB + c;
      v[q] = (gr < nbk && gc < nbk && r >= c) ? L[gr + (long)gc * ldl] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < (PB / LB) * LB * LB / 256; ++q) {
      const int e = threadIdx.x + 256 * q;
      const int b = e / (LB * LB), r = e % LB, c = (e / LB) % LB;
      if (b * LB + r < nbk && b * LB + c < nbk && r >= c)
        Ld[b][c * LB - (c * (c - 1)) / 2 + (r - c)] = v[q];
    }
  }
  __syncthreads();
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= m) return;
  double* xr = X + row;
  for (int jb = 0; jb < nbk; jb += LB) {
    const int w = FULL ? LB : ((nbk - jb) < LB ? (nbk - jb) : LB);
    double x[LB];
#pragma unroll
    for (int c = 0; c < LB; ++c) x[c] = (FULL || c < w) ? xr[(long)(jb + c) * ldx] : 0.0;
    for (int k = 0; k < jb; ++k) {
      const double xk = xr[(long)k * ldx];
      const double* lk = L + (long)k * ldl + jb;
#pragma unroll
      for (int c = 0; c < LB; ++c)
        if (FULL || c < w) x[c] -= lk[c] * xk;
    }
    const double* Lb = Ld[jb / LB];
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      if (FULL || j < w) {
        const double* lj = Lb + (j * LB - (j * (j - 1)) / 2);  // column j from its diagonal
        x[j] *= 1.0 / lj[0];
#pragma unroll
        for (int c = j + 1; c < LB; ++c)
          if (FULL || c < w) x[c] -= lj[c - j] * x[j];
      }
    }
#pragma unroll
    for (int c = 0; c < LB; ++c)
      if (FULL || c < w) xr[(long)(jb + c) * ldx] = x[c];
  }
}

int launch_trsm_rlt(hipStream_t s, int m, int nb, const double* L11, long ldl, double* A21,
                    long lda, const int* abort_flag) {
  if (m <= 0 || nb <= 0) return FLA_HIP_OK;
  if (nb > PB) return FLA_HIP_EARG;
  if (g_trsm_rows) {
    const unsigned g = (unsigned)((m + 255) / 256);
    if (nb == PB)
      hipLaunchKernelGGL(trsm_rlt_rows_kernel<true>, dim3(g), dim3(256), 0, s, m, nb, L11, ldl,
                         A21, lda, abort_flag);
    else
      hipLaunchKernelGGL(trsm_rlt_rows_kernel<false>, dim3(g), dim3(256), 0, s, m, nb, L11, ldl,
                         A21, lda, abort_flag);
    FLA_HIP_CHECK(hipGetLastError());
    return FLA_HIP_OK;
  }
  for (int b0 = 0; b0 < nb; b0 += LB) {
    const int w = (nb - b0) < LB ? (nb - b0) : LB;
    const double* Lbb = L11 + (long)b0 + (long)b0 * ldl;
    double* Xb = A21 + (long)b0 * lda;
    if (w == LB)
      hipLaunchKernelGGL(trsm_rlt_leaf_kernel<true>, dim3((m + 63) / 64), dim3(64), 0, s, m, w,
                         Lbb, ldl, Xb, lda, abort_flag);
    else
      hipLaunchKernelGGL(trsm_rlt_leaf_kernel<false>, dim3((m + 63) / 64), dim3(64), 0, s, m, w,
                         Lbb, ldl, Xb, lda, abort_flag);
    FLA_HIP_CHECK(hipGetLastError());
    const int rest = nb - b0 - w;
    if (rest > 0) {
      // X[:, b0+w:] -= X_b * L11[b0+w:, b0:b0+w]'
      FLA_TRY(launch_dgemm_mfma(s, false, true, 0, m, rest, w, -1.0, Xb, lda,
                                L11 + (long)(b0 + w) + (long)b0 * ldl, ldl, 1.0,
                                Xb + (long)w * lda, lda, abort_flag));
    }
  }
  return FLA_HIP_OK;
}

int launch_transpose_swap(hipStream_t s, int n, double* A, long lda) {
  if (n <= 1) return FLA_HIP_OK;
  const int T = (n + 31) / 32;
  const long nblk = (long)T * (T + 1) / 2;
  hipLaunchKernelGGL(transpose_swap_kernel, dim3((unsigned)nblk), dim3(256), 0, s, n, A, lda);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// Factors columns [k, k+kb) of the lower matrix (rows k..n-1): per 128-column
// sub-block, the diagonal factor, the blocked panel solve of everything
// below it, and the update of the panel's remaining columns.  All on stream s.
int chol_panel_f64(hipStream_t s, int n, double* A, long lda, int k, int kb, int* info) {
  for (int kk = k; kk < k + kb; kk += PB) {
    const int ib = (k + kb - kk) < PB ? (k + kb - kk) : PB;
    double* Akk = A + (long)kk + (long)kk * lda;
    FLA_TRY(launch_potrf_diag(s, ib, Akk, lda, kk, info));
    const int mb = n - kk - ib;
    if (mb <= 0) continue;
    FLA_TRY(launch_trsm_rlt(s, mb, ib, Akk, lda, Akk + ib, lda, info));
    const int w = k + kb - kk - ib;  // remaining columns of this panel
    if (w > 0) {
      const double* P = Akk + ib;  // mb x ib
      FLA_TRY(launch_dgemm_mfma(s, false, true, 1, mb, w, ib, -1.0, P, lda, P, lda, 1.0,
                                A + (long)(kk + ib) + (long)(kk + ib) * lda, lda, info));
    }
  }
  return FLA_HIP_OK;
}

// Same factor with fewer launches, for callers that only need a
// preconditioner-quality panel (the mixed-precision solver): each 128-block's
// panel solve multiplies by the explicit inverse of its diagonal factor,
// A21 := A21 * inv(L11)^T, one in-place MFMA GEMM instead of the blocked
// substitution (every workgroup of that GEMM owns whole 128-wide rows).
int chol_panel_f64_inv(hipStream_t s, int n, double* A, long lda, int k, int kb, int* info) {
  double* W = fla_ctx().dlinv;
  for (int kk = k; kk < k + kb; kk += PB) {
    const int ib = (k + kb - kk) < PB ? (k + kb - kk) : PB;
    double* Akk = A + (long)kk + (long)kk * lda;
    const int mb = n - kk - ib;
    if (mb <= 0) {
      FLA_TRY(launch_potrf_diag(s, ib, Akk, lda, kk, info));
      continue;
    }
    FLA_TRY(launch_potrf_diag_inv(s, ib, Akk, lda, kk, info, W));
    FLA_TRY(launch_dgemm_mfma(s, false, true, 0, mb, ib, ib, 1.0, Akk + ib, lda, W, PB, 0.0,
                              Akk + ib, lda, info));
    const int w = k + kb - kk - ib;
    if (w > 0) {
      const double* P = Akk + ib;
      FLA_TRY(launch_dgemm_mfma(s, false, true, 1, mb, w, ib, -1.0, P, lda, P, lda, 1.0,
                                A + (long)(kk + ib) + (long)(kk + ib) * lda, lda, info));
    }
  }
  return FLA_HIP_OK;
}

// Lower blocked Cholesky with look-ahead: the panel of step k+1 runs on
// s_panel while the bulk of step k's trailing update runs on s_update.
static int chol_lower(int n, double* A, long lda, int* dinfo) {
  FlaHipCtx& c = fla_ctx();
  const int nb = fla_chol_nb(n);
  hipStream_t sp = c.s_panel, su = c.s_update;
  int evi = 0;
  auto next_event = [&]() { hipEvent_t e = c.ev[evi]; evi = (evi + 1) % FlaHipCtx::kEvents; return e; };

  // Both streams start after whatever the caller queued on s_update.
  hipEvent_t start = next_event();
  FLA_HIP_CHECK(hipEventRecord(start, su));
  FLA_HIP_CHECK(hipStreamWaitEvent(sp, start, 0));

  // Panel 0.
  int kb0 = n < nb ? n : nb;
  FLA_TRY(chol_panel_f64(sp, n, A, lda, 0, kb0, dinfo));
  for (int k = 0; k < n;) {
    const int kb = (n - k) < nb ? (n - k) : nb;
    hipEvent_t panel_done = next_event();
    FLA_HIP_CHECK(hipEventRecord(panel_done, sp));
    FLA_HIP_CHECK(hipStreamWaitEvent(su, panel_done, 0));
    const int k2 = k + kb;      // first column of the trailing matrix
    const int m2 = n - k2;      // its order
    if (m2 <= 0) break;
    const double* P = A + (long)k2 + (long)k * lda;  // m2 x kb, the factored panel
    const int kb2 = m2 < nb ? m2 : nb;             // width of the next panel
    // (a) next panel's columns: rows k2.., columns [k2, k2+kb2) (lower part)
    void* tok = fla_stats_begin(su);
    FLA_TRY(launch_dgemm_mfma(su, false, true, 1, m2, kb2, kb, -1.0, P, lda, P, lda, 1.0,
                              A + (long)k2 + (long)k2 * lda, lda, dinfo));
    fla_stats_end(su, tok, 0, lower_update_flops(m2, kb2, kb));
    hipEvent_t next_ready = next_event();
    FLA_HIP_CHECK(hipEventRecord(next_ready, su));
    // (b) the rest of the trailing matrix
    const int m3 = m2 - kb2;
    if (m3 > 0) {
      const double* P3 = P + kb2;
      tok = fla_stats_begin(su);
      FLA_TRY(launch_dgemm_mfma(su, false, true, 1, m3, m3, kb, -1.0, P3, lda, P3, lda, 1.0,
                                A + (long)(k2 + kb2) + (long)(k2 + kb2) * lda, lda, dinfo));
      fla_stats_end(su, tok, 0, lower_update_flops(m3, m3, kb));
    }
    // Next panel on the panel stream, concurrent with (b).
    FLA_HIP_CHECK(hipStreamWaitEvent(sp, next_ready, 0));
    FLA_TRY(chol_panel_f64(sp, n, A, lda, k2, kb2, dinfo));
    k = k2;
  }
  // Join: s_update waits for the last panel.
  hipEvent_t fin = next_event();
  FLA_HIP_CHECK(hipEventRecord(fin, sp));
  FLA_HIP_CHECK(hipStreamWaitEvent(su, fin, 0));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dpotrf(int uplo, int n, double* A, int lda, int* info) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (uplo != kLower && uplo != kUpper) return FLA_HIP_EARG;
  if (n < 0 || lda < (n > 1 ? n : 1)) return FLA_HIP_EARG;
  *info = kSuccess;
  if (n == 0) return FLA_HIP_OK;
  FlaHipCtx& c = fla_ctx();
  int* dinfo = c.dscratch;
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, c.s_update, dinfo, kSuccess);
  FLA_HIP_CHECK(hipGetLastError());
  if (uplo == kUpper) FLA_TRY(launch_transpose_swap(c.s_update, n, A, lda));
  FLA_TRY(chol_lower(n, A, lda, dinfo));
  if (uplo == kUpper) FLA_TRY(launch_transpose_swap(c.s_update, n, A, lda));
  int h = kSuccess;
  FLA_HIP_CHECK(hipMemcpyAsync(&h, dinfo, sizeof(int), hipMemcpyDeviceToHost, c.s_update));
  FLA_HIP_CHECK(hipStreamSynchronize(c.s_update));
  fla_stats_harvest();
  *info = h;
  return FLA_HIP_OK;
}

extern "C" int fla_hip_set_panel_solve(int rows) {
  g_trsm_rows = rows != 0;
  return FLA_HIP_OK;
}
