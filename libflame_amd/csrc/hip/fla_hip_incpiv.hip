// LU with incremental pivoting over b x b tiles (FLASH_LU_incpiv) and its
// forward substitution (FLASH_FS_incpiv), on a flat column-major matrix in
// HBM whose tiles are the b x b blocks of the hierarchical object.
//
// Reference (paths relative to the libflame source tree):
//   FLASH_LU_incpiv.c:13-39 -> FLASH_LU_incpiv_opt1.c:13-42 ->
//   zaux/flamec/flash/FLASH_LU_incpiv_var2.c:13-114 (per diagonal tile k:
//   LU_piv of A_kk; FLASH_Trsm_piv on the tile row; for each tile i below,
//   FLASH_SA_LU of [A_kk; A_ik] and FLASH_SA_FS of tile rows k and i),
//   leaves zaux/flamec/fla/FLA_SA_LU_blk.c, FLA_SA_LU_unb.c, FLA_SA_FS_blk.c,
//   FLA_SA_Apply_pivots.c; FLASH_FS_incpiv_aux1.c / _aux2.c for the solve.
//
// GPU mapping.  The tile row k right of the diagonal tile is one strided
// b x (n - (k+1) b) matrix of the flat copy, so every per-tile operation on
// it (interchanges, the unit-lower solve, the rank-bs update of SA_FS) is one
// launch over the whole row instead of one task per tile:
//   * A_kk: the library's LU with partial pivoting (pivots within the tile);
//   * tile row k: device interchanges + the library's dtrsm;
//   * SA_LU(i): for each nb_alg-column block, sa_lu_unb_kernel (one
//     workgroup: the sequential structure-aware LU of [U11; D1]), the
//     U12 / D2 interchanges, a unit-lower substitution on U12, an MFMA GEMM
//     D2 -= D1 U12;
//   * SA_FS(i): per block, the C1 / E interchanges, the substitution on C1
//     and an MFMA GEMM E -= D1 C1 across the whole tile row.
// The chain over i is sequential (each SA_LU updates A_kk's upper triangle),
// as in the reference.
#include <utility>

#include "fla_hip_internal.h"

namespace {

constexpr int kUnbThreads = 256;

// idamax's choice (src/flablas/f2c/idamax.c:40-46, "<=" skip) over a
// column: the scan restarts after every NaN, so the result is the first
// maximum of |x| after the last NaN (the last NaN itself when it ends the
// column); without NaN, the first maximum.  Reduced over the workgroup.
__device__ int block_idamax(const double* __restrict__ x, int n, int* s_int, double* s_dbl) {
  const int t = threadIdx.x;
  int last_nan = -1;
  for (int i = t; i < n; i += kUnbThreads)
    if (x[i] != x[i]) last_nan = i;
  s_int[t] = last_nan;
  __syncthreads();
  for (int w = kUnbThreads / 2; w > 0; w >>= 1) {
    if (t < w && s_int[t + w] > s_int[t]) s_int[t] = s_int[t + w];
    __syncthreads();
  }
  const int lo = s_int[0] + 1;  // scan start
  __syncthreads();
  if (lo >= n) return n - 1;  // the column ends with a NaN
  double best = -1.0;
  int bi = n;
  for (int i = lo + t; i < n; i += kUnbThreads) {
    const double v = fabs(x[i]);
    if (v > best) {
      best = v;
      bi = i;
    }
  }
  s_dbl[t] = best;
  s_int[t] = bi;
  __syncthreads();
  for (int w = kUnbThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      const double o = s_dbl[t + w];
      const int oi = s_int[t + w];
      if (o > s_dbl[t] || (o == s_dbl[t] && oi < s_int[t])) {
        s_dbl[t] = o;
        s_int[t] = oi;
      }
    }
    __syncthreads();
  }
  const int r = s_int[0];
  __syncthreads();
  return r;
}

// FLA_SA_LU_unb (double branch) on U (bs x bs, upper part), D (mD x bs),
// pivots p[0..bs) and L (bs x bs): L := triu(U); per column j the row of D
// with the largest |D(:, j)| replaces row j of L when strictly larger in
// magnitude than L(j, j) (p[j] = ipiv + bs - j, else 0), D(:, j) is scaled by
// 1 / L(j, j), D(:, j+1:) -= D(:, j) L(j, j+1:) (no fused multiply-add, as
// dger_), and row j of U gets row j of L from the diagonal on.
__global__ __launch_bounds__(kUnbThreads) void sa_lu_unb_kernel(int bs, double* __restrict__ U,
                                                                long ldu, int mD,
                                                                double* __restrict__ D, long ldd,
                                                                int* __restrict__ p,
                                                                double* __restrict__ L,
                                                                long ldl) {
  __shared__ int s_int[kUnbThreads];
  __shared__ double s_dbl[kUnbThreads];
  const int t = threadIdx.x;
  for (int e = t; e < bs * bs; e += kUnbThreads) {
    const int i = e % bs, c = e / bs;
    L[i + c * ldl] = (i <= c) ? U[i + c * ldu] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < bs; ++j) {
    const int ipiv = block_idamax(D + (long)j * ldd, mD, s_int, s_dbl);
    const double ltmp = L[j + j * ldl];
    const double dtmp = D[ipiv + (long)j * ldd];
    const bool swap = fabs(ltmp) < fabs(dtmp);
    __syncthreads();  // every thread has read L(j, j) and D(ipiv, j)
    if (swap) {
      for (int c = t; c < bs; c += kUnbThreads) {
        const double a = L[j + c * ldl];
        L[j + c * ldl] = D[ipiv + (long)c * ldd];
        D[ipiv + (long)c * ldd] = a;
      }
    }
    if (t == 0) p[j] = swap ? ipiv + bs - j : 0;
    __syncthreads();
    const double dinv = 1.0 / L[j + j * ldl];
    for (int i = t; i < mD; i += kUnbThreads) D[i + (long)j * ldd] = dinv * D[i + (long)j * ldd];
    __syncthreads();
    // dger_: for each column c > j with L(j, c) != 0, D(:, c) += D(:, j) * (-L(j, c))
    const int nr = bs - j - 1;
    for (long e = t; e < (long)nr * mD; e += kUnbThreads) {
      const int i = (int)(e % mD), c = j + 1 + (int)(e / mD);
      const double y = L[j + c * ldl];
      if (y != 0.0) {
        const double temp = -1.0 * y;
        D[i + (long)c * ldd] = __dadd_rn(D[i + (long)c * ldd], __dmul_rn(D[i + (long)j * ldd], temp));
      }
    }
    for (int c = j + t; c < bs; c += kUnbThreads) U[j + c * ldu] = L[j + c * ldl];
    __syncthreads();
  }
}

// The same leaf with D held in registers: thread t owns row t of D (mD <=
// 256, bs <= BSMAX), L lives in LDS (row-major), and each column costs two
// barriers: one after the per-wave pivot candidates are posted, one after
// the interchange.  The operations and their order per element are those of
// sa_lu_unb_kernel (the same idamax rule, the reciprocal scaling, dger_'s
// separate multiply and add, skipped for L(j, c) == 0), so the results are
// identical.
struct SaCand {
  int nan;     // last NaN row of the wave, -1 if none
  int idx;     // first maximum of |D(i, j)| over the wave's rows after its last NaN, -1 if none
  double best; // that maximum
  double val;  // D(idx, j) (signed)
};

__device__ __forceinline__ double sa_wave_max(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  return x;
}

// Column J of the register leaf (compile-time J, so that every index into
// the row array is a constant and the array stays in VGPRs).
template <int BSMAX, int J>
__device__ __forceinline__ void sa_reg_column(double (&x)[BSMAX], int bs, int t, bool act,
                                              double* Ls, SaCand* cand, double* __restrict__ U,
                                              long ldu, int* __restrict__ p) {
  constexpr int LS = BSMAX + 1;
  if (J >= bs) return;  // uniform
  const int lane = t & 63, wave = t >> 6;
  // L(J, J) before the barrier: row J of L changes only in this column's
  // interchange, which the pivot row's owner makes after the barrier.
  const double ltmp = Ls[J * LS + J];
  // ---- per-wave idamax candidates of column J ----
  {
    const double xj = x[J];
    const bool isnan = act && xj != xj;
    const unsigned long long nb = __ballot(isnan);
    const int wnan = nb ? 63 - __builtin_clzll(nb) : -1;  // lane of the wave's last NaN
    const double key = (act && !isnan && lane > wnan) ? fabs(xj) : -1.0;
    const double m = sa_wave_max(key);
    const unsigned long long bm = __ballot(m >= 0.0 && key == m);
    const int l = bm ? __builtin_ctzll(bm) : 0;
    const double v = __shfl(xj, l, 64);
    if (lane == 0) {
      cand[wave].nan = wnan >= 0 ? wave * 64 + wnan : -1;
      cand[wave].idx = bm ? wave * 64 + l : -1;
      cand[wave].best = m;
      cand[wave].val = v;
    }
  }
  __syncthreads();
  // ---- combine the waves in row order: a NaN restarts the scan ----
  int nanr = -1, ipiv = -1;
  double best = -1.0, dtmp = 0.0;
#pragma unroll
  for (int w = 0; w < kUnbThreads / 64; ++w) {
    const SaCand cw = cand[w];
    if (cw.nan >= 0) {
      nanr = cw.nan;
      ipiv = cw.idx;
      best = cw.best;
      dtmp = cw.val;
    } else if (cw.idx >= 0 && (ipiv < 0 || cw.best > best)) {
      ipiv = cw.idx;
      best = cw.best;
      dtmp = cw.val;
    }
  }
  if (ipiv < 0) {  // the column ends with a NaN: it is the choice
    ipiv = nanr;
    dtmp = __builtin_nan("");
  }
  const bool swap = fabs(ltmp) < fabs(dtmp);
  if (swap && t == ipiv) {
#pragma unroll
    for (int c = 0; c < BSMAX; ++c) {
      if (c < bs) {
        const double a = Ls[J * LS + c];
        Ls[J * LS + c] = x[c];
        x[c] = a;
      }
    }
  }
  if (t == 0) p[J] = swap ? ipiv + bs - J : 0;
  __syncthreads();
  // ---- scale column J, rank-1 update of the columns right of it ----
  const double dinv = 1.0 / Ls[J * LS + J];
  if (act) {
    x[J] = dinv * x[J];
#pragma unroll
    for (int c = J + 1; c < BSMAX; ++c) {
      if (c < bs) {
        const double y = Ls[J * LS + c];
        if (y != 0.0) x[c] = __dadd_rn(x[c], __dmul_rn(x[J], -1.0 * y));
      }
    }
  }
  if (t >= J && t < bs) U[J + (long)t * ldu] = Ls[J * LS + t];
}

template <int BSMAX, int... J>
__device__ __forceinline__ void sa_reg_columns(std::integer_sequence<int, J...>,
                                               double (&x)[BSMAX], int bs, int t, bool act,
                                               double* Ls, SaCand* cand, double* __restrict__ U,
                                               long ldu, int* __restrict__ p) {
  (sa_reg_column<BSMAX, J>(x, bs, t, act, Ls, cand, U, ldu, p), ...);
}

template <int BSMAX>
__global__ __launch_bounds__(kUnbThreads) void sa_lu_reg_kernel(int bs, double* __restrict__ U,
                                                                long ldu, int mD,
                                                                double* __restrict__ D, long ldd,
                                                                int* __restrict__ p,
                                                                double* __restrict__ L,
                                                                long ldl) {
  constexpr int LS = BSMAX + 1;
  __shared__ double Ls[BSMAX * LS];  // Ls[i * LS + c] = L(i, c)
  __shared__ SaCand cand[kUnbThreads / 64];
  const int t = threadIdx.x;
  const bool act = t < mD;
  double x[BSMAX];
#pragma unroll
  for (int c = 0; c < BSMAX; ++c) x[c] = (act && c < bs) ? D[t + (long)c * ldd] : 0.0;
  for (int e = t; e < bs * bs; e += kUnbThreads) {
    const int i = e % bs, c = e / bs;
    Ls[i * LS + c] = (i <= c) ? U[i + c * ldu] : 0.0;
  }
  __syncthreads();
  sa_reg_columns<BSMAX>(std::make_integer_sequence<int, BSMAX>{}, x, bs, t, act, Ls, cand, U, ldu,
                        p);
  if (act) {
#pragma unroll
    for (int c = 0; c < BSMAX; ++c)
      if (c < bs) D[t + (long)c * ldd] = x[c];
  }
  __syncthreads();
  for (int e = t; e < bs * bs; e += kUnbThreads) {
    const int i = e % bs, c = e / bs;
    L[i + c * ldl] = Ls[i * LS + c];
  }
}

int launch_sa_lu_unb(hipStream_t s, int bs, double* U, long ldu, int mD, double* D, long ldd,
                     int* p, double* L, long ldl) {
  if (mD <= kUnbThreads && bs <= 32)
    hipLaunchKernelGGL(sa_lu_reg_kernel<32>, dim3(1), dim3(kUnbThreads), 0, s, bs, U, ldu, mD, D,
                       ldd, p, L, ldl);
  else if (mD <= kUnbThreads && bs <= 64)
    hipLaunchKernelGGL(sa_lu_reg_kernel<64>, dim3(1), dim3(kUnbThreads), 0, s, bs, U, ldu, mD, D,
                       ldd, p, L, ldl);
  else
    hipLaunchKernelGGL(sa_lu_unb_kernel, dim3(1), dim3(kUnbThreads), 0, s, bs, U, ldu, mD, D, ldd,
                       p, L, ldl);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// FLA_SA_Apply_pivots: for i in order, row i of C (bs rows) <-> row
// p[i] - (bs - i) of E when p[i] != 0; one thread per column.
__global__ void sa_swap_kernel(int bs, int ncols, double* __restrict__ C, long ldc,
                               double* __restrict__ E, long lde, const int* __restrict__ p) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  double* c = C + (long)col * ldc;
  double* e = E + (long)col * lde;
  for (int i = 0; i < bs; ++i) {
    const int pi = p[i];
    if (pi != 0) {
      const int r = pi - (bs - i);
      const double a = c[i];
      c[i] = e[r];
      e[r] = a;
    }
  }
}

int launch_sa_swap(hipStream_t s, int bs, int ncols, double* C, long ldc, double* E, long lde,
                   const int* p) {
  if (ncols <= 0 || bs <= 0) return FLA_HIP_OK;
  hipLaunchKernelGGL(sa_swap_kernel, dim3((ncols + 255) / 256), dim3(256), 0, s, bs, ncols, C,
                     ldc, E, lde, p);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// X := inv(trilu(L11)) X for the bs x bs unit-lower L11 (its strictly lower
// part) and bs x nc X: one lane per column up to bs = 128, the general dtrsm
// beyond.
int unit_lower_solve(hipStream_t s, int bs, int nc, const double* L11, long ldl, double* X,
                     long ldx) {
  if (bs <= 128) return launch_trsm_llnu_subst(s, nc, bs, L11, ldl, X, ldx);
  return fla_hip_dtrsm(kLeft, kLower, kNoTrans, kUnit, bs, nc, 1.0, L11, (int)ldl, X, (int)ldx);
}

// FLA_SA_FS_blk: the transformations of one SA_LU (L tile b x nb_alg at L,
// D = A_ik, pivots p) applied to C (b x nc) over E (b x nc).
int sa_fs(hipStream_t s, int b, int nb_alg, const double* L, long ldl, const double* D,
          long ldd, const int* p, int nc, double* C, long ldc, double* E, long lde) {
  if (nc <= 0) return FLA_HIP_OK;
  for (int s0 = 0; s0 < b; s0 += nb_alg) {
    const int bs = (b - s0) < nb_alg ? (b - s0) : nb_alg;
    FLA_TRY(launch_sa_swap(s, bs, nc, C + s0, ldc, E, lde, p + s0));
    FLA_TRY(unit_lower_solve(s, bs, nc, L + s0, ldl, C + s0, ldc));
    FLA_TRY(launch_dgemm_mfma(s, false, false, 0, b, nc, bs, -1.0, D + (long)s0 * ldd, ldd,
                              C + s0, ldc, 1.0, E, lde, nullptr));
  }
  return FLA_HIP_OK;
}

// FLA_SA_LU_blk on U = A_kk (b x b, upper part), D = A_ik (b x b).
int sa_lu(hipStream_t s, int b, int nb_alg, double* U, long ldu, double* D, long ldd, int* p,
          double* L, long ldl) {
  for (int s0 = 0; s0 < b; s0 += nb_alg) {
    const int bs = (b - s0) < nb_alg ? (b - s0) : nb_alg;
    const int nrest = b - s0 - bs;
    double* U11 = U + s0 + (long)s0 * ldu;
    double* U12 = U + s0 + (long)(s0 + bs) * ldu;
    double* D1 = D + (long)s0 * ldd;
    double* D2 = D + (long)(s0 + bs) * ldd;
    FLA_TRY(launch_sa_lu_unb(s, bs, U11, ldu, b, D1, ldd, p + s0, L + s0, ldl));
    if (nrest > 0) {
      FLA_TRY(launch_sa_swap(s, bs, nrest, U12, ldu, D2, ldd, p + s0));
      FLA_TRY(unit_lower_solve(s, bs, nrest, L + s0, ldl, U12, ldu));
      FLA_TRY(launch_dgemm_mfma(s, false, false, 0, b, nrest, bs, -1.0, D1, ldd, U12, ldu, 1.0,
                                D2, ldd, nullptr));
    }
  }
  return FLA_HIP_OK;
}

int incpiv_args(int m, int n, int lda, int b, int nb_alg, int ldp, int ldl) {
  if (m < 0 || n < 0 || b <= 0 || nb_alg <= 0) return FLA_HIP_EARG;
  if (m % b != 0 || n % b != 0) {
    std::fprintf(stderr,
                 "libflame_amd: LU with incremental pivoting needs m and n multiples of the "
                 "tile size %d (got %d x %d)\n",
                 b, m, n);
    return FLA_HIP_EARG;
  }
  if (lda < (m > 1 ? m : 1) || ldp < (m > 1 ? m : 1) || ldl < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  return FLA_HIP_OK;
}

}  // namespace

extern "C" int fla_hip_dgetrf_incpiv(int m, int n, double* A, int lda, int b, int nb_alg, int* p,
                                     int ldp, double* L, int ldl) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  FLA_TRY(incpiv_args(m, n, lda, b, nb_alg, ldp, ldl));
  const int mt = m / b, nt = n / b, kt = mt < nt ? mt : nt;
  hipStream_t s = fla_ctx().s_update;
  for (int k = 0; k < kt; ++k) {
    double* Akk = A + (long)k * b * (lda + 1);
    int* pkk = p + (long)k * b + (long)k * ldp;
    const int nr = n - (k + 1) * b;
    double* Crow = A + (long)k * b + (long)(k + 1) * b * lda;  // tile row k, right part
    int info = 0;
    // LU_piv_copy_task: the tile's LU (pivots within the tile; the result is
    // FLA_SUCCESS for FLASH_LU_incpiv whatever the tile's null pivots)
    FLA_TRY(fla_hip_dgetrf(b, b, Akk, lda, pkk, &info));
    if (nr > 0) {  // FLASH_Trsm_piv: pivots, then the unit-lower solve
      FLA_TRY(fla_hip_dlaswp_rows(b, nr, Crow, lda, 0, b - 1, pkk, 1));
      FLA_TRY(fla_hip_dtrsm(kLeft, kLower, kNoTrans, kUnit, b, nr, 1.0, Akk, lda, Crow, lda));
    }
    for (int i = k + 1; i < mt; ++i) {
      double* Aik = A + (long)i * b + (long)k * b * lda;
      int* pik = p + (long)i * b + (long)k * ldp;
      double* Lik = L + (long)i * b + (long)k * nb_alg * ldl;
      FLA_TRY(sa_lu(s, b, nb_alg, Akk, lda, Aik, lda, pik, Lik, ldl));
      if (nr > 0)
        FLA_TRY(sa_fs(s, b, nb_alg, Lik, ldl, Aik, lda, pik, nr, Crow, lda,
                      A + (long)i * b + (long)(k + 1) * b * lda, lda));
    }
  }
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dfs_incpiv(int m, int n, const double* A, int lda, int b, int nb_alg,
                                  const int* p, int ldp, const double* L, int ldl, int nrhs,
                                  double* X, int ldx) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  FLA_TRY(incpiv_args(m, n, lda, b, nb_alg, ldp, ldl));
  if (nrhs < 0 || ldx < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  if (nrhs == 0) return FLA_HIP_OK;
  const int mt = m / b, nt = n / b, kt = mt < nt ? mt : nt;
  hipStream_t s = fla_ctx().s_update;
  for (int k = 0; k < kt; ++k) {
    const double* Akk = A + (long)k * b * (lda + 1);
    const int* pkk = p + (long)k * b + (long)k * ldp;
    double* Xk = X + (long)k * b;
    FLA_TRY(fla_hip_dlaswp_rows(b, nrhs, Xk, ldx, 0, b - 1, pkk, 1));
    FLA_TRY(fla_hip_dtrsm(kLeft, kLower, kNoTrans, kUnit, b, nrhs, 1.0, Akk, lda, Xk, ldx));
    for (int i = k + 1; i < mt; ++i)
      FLA_TRY(sa_fs(s, b, nb_alg, L + (long)i * b + (long)k * nb_alg * ldl, ldl,
                    A + (long)i * b + (long)k * b * lda, lda, p + (long)i * b + (long)k * ldp,
                    nrhs, Xk, ldx, X + (long)i * b, ldx));
  }
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  return FLA_HIP_OK;
}
