// Triangular solves on the GPU.
//
// General dtrsm (all side/uplo/trans/diag combinations of dtrsm_,
// src/flablas/f2c/dtrsm.c, used by FLA_Trsm_external and the solve paths
// FLA_Chol_solve / FLA_LU_piv_solve / dpotrs_ / dgetrs_), blocked by 128:
//   X_k := inv(op(T_kk)) * (B_k - sum_{j<k} op(T)_kj X_j)
// The diagonal-block solves substitute in dtrsm_'s order (block staged in
// LDS, one lane per right-hand side); the updates are MFMA GEMMs.  With at
// most 32 right-hand sides on the left, one persistent launch per 16 columns
// (fla_hip_trsv.hip) does the whole solve; it multiplies by the inverted
// diagonal blocks (inverted in parallel, one workgroup per block) where their
// condition number is at most kInvCondMax and substitutes through the rest.
//
// LU panel: A12 := inv(tril_unit(L11)) * A12 (FLA_LU_piv_blk_var5.c:88-91,
// dtrsm_('L','L','N','U')): for a full 128-column panel, a blocked
// substitution on MFMA (one wave per 16 columns); for a narrower last panel,
// one lane per column with the column in VGPRs and L11 read through uniform
// loads.
#include "fla_hip_internal.h"
#include "fla_hip_lds_tri.h"

namespace {

constexpr int TB = 128;       // diagonal block order
constexpr int TLD = TB + 1;   // LDS leading dimension
// Largest condition number of a diagonal block that the few-right-hand-side
// solve multiplies by its inverse (the block order: the inverse's error is
// then of the order of a substitution's); beyond it the solve substitutes.
constexpr double kInvCondMax = 128.0;

// Inverts the nbk x nbk diagonal block k of a triangular matrix (lower or
// upper storage, unit or non-unit diagonal) into W (ld TB per block, same
// triangle orientation as A; the other triangle of W is zeroed).
// The block is loaded as a lower triangle L (transposed for upper storage)
// and inverted in LDS by recursive 2x2 blocking,
//   inv([A 0; B C]) = [inv(A) 0; -inv(C) B inv(A)  inv(C)],
// with the four 32x32 diagonal blocks solved by one wave each (lane = column
// of the inverse, forward substitution, L read by broadcast) and the
// off-diagonal blocks at sizes 64 and 128 formed as register-held products.
// Abs sums of the 128 x 128 lower block in S: thread t < 128 sums column t,
// thread 128 + r sums row r (out[t]).
__device__ __forceinline__ void block_abs_sums(const double* S, int t, double* out) {
  double sum = 0.0;
  if (t < TB) {
    for (int i = 0; i < TB; ++i) sum += fabs(S[t * TLD + i]);
  } else {
    const int r = t - TB;
    for (int c = 0; c < TB; ++c) sum += fabs(S[c * TLD + r]);
  }
  out[t] = sum;
}

// When ill != nullptr, ill[k] = 1 if the condition number of block k,
// max(||T_kk||_1 ||inv(T_kk)||_1, ||T_kk||_inf ||inv(T_kk)||_inf), exceeds
// kInvCondMax (or is not finite), else 0: a solve that multiplies by the
// inverse has a backward error that grows with that number, so the solves
// substitute through such blocks instead.
__global__ __launch_bounds__(256) void trtri_blocks_kernel(int n, const double* __restrict__ A,
                                                           long lda, int upper, int unit,
                                                           double* __restrict__ W,
                                                           int* __restrict__ ill) {
  __shared__ double S[TB * TLD];
  __shared__ double nrm[2][2 * TB];
  const int k = blockIdx.x;
  const int r0 = k * TB;
  const int nbk = (n - r0) < TB ? (n - r0) : TB;
  const int t = threadIdx.x;
  // 64 elements per thread, loads issued 16 at a time; consecutive threads
  // read consecutive addresses of the stored triangle.
  for (int q0 = 0; q0 < TB * TB / 256; q0 += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = t + (q0 + q) * 256;
      const int f = idx % TB, g = idx / TB;  // stored element (f, g)
      v[q] = (f < nbk && g < nbk) ? A[(long)(r0 + f) + (long)(r0 + g) * lda] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = t + (q0 + q) * 256;
      const int f = idx % TB, g = idx / TB;
      const int i = upper ? g : f, j = upper ? f : g;  // S(i, j), lower
      double x = (i >= j && i < nbk) ? v[q] : 0.0;
      if (i == j && (unit || i >= nbk)) x = 1.0;  // unit diagonal / identity padding
      S[j * TLD + i] = x;
    }
  }
  __syncthreads();
  if (ill) block_abs_sums(S, t, nrm[0]);
  fla_lds::trtri_lower_lds(S, t);  // synchronizes before reading S
  if (ill) {
    block_abs_sums(S, t, nrm[1]);
    __syncthreads();
    if (t == 0) {
      double a1 = 0.0, ai = 0.0, b1 = 0.0, bi = 0.0;
      bool finite = true;  // an overflowed inverse holds inf or NaN (fmax would drop NaN)
      for (int q = 0; q < TB; ++q) {
        a1 = fmax(a1, nrm[0][q]);
        ai = fmax(ai, nrm[0][TB + q]);
        b1 = fmax(b1, nrm[1][q]);
        bi = fmax(bi, nrm[1][TB + q]);
        finite = finite && isfinite(nrm[0][q]) && isfinite(nrm[0][TB + q]) &&
                 isfinite(nrm[1][q]) && isfinite(nrm[1][TB + q]);
      }
      const double kappa = fmax(a1 * b1, ai * bi);
      ill[k] = (finite && kappa <= kInvCondMax) ? 0 : 1;
    }
  }
  double* Wk = W + (long)k * TB * TB;
  for (int idx = t; idx < TB * TB; idx += 256) {
    const int i = idx % TB, j = idx / TB;  // W(i,j), ld TB
    double v;
    if (upper)
      v = (i <= j) ? S[i * TLD + j] : 0.0;  // inv(U) = inv(L)' with L = U'
    else
      v = (i >= j) ? S[j * TLD + i] : 0.0;
    Wk[(long)i + (long)j * TB] = v;
  }
}

// A12 := inv(L11) * A12 with L11 unit lower nb x nb (nb <= 128, ld ldl):
// lane = column, forward substitution in dtrsm_('L','L','N','U') order.
// FULL (nb == 128) drops the uniform bounds tests of the partial last panel.
template <bool FULL>
__global__ __launch_bounds__(64) void trsm_llnu_kernel(int ncols, int nb,
                                                       const double* __restrict__ L, long ldl,
                                                       double* __restrict__ B, long ldb) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= ncols) return;
  double* col = B + (long)j * ldb;
  double x[TB];
#pragma unroll
  for (int i = 0; i < TB; ++i) x[i] = (FULL || i < nb) ? col[i] : 0.0;
#pragma unroll
  for (int c = 0; c < TB; ++c) {
    if (FULL || c < nb) {
#pragma unroll
      for (int i = c + 1; i < TB; ++i)
        if (FULL || i < nb) x[i] -= x[c] * L[(long)i + (long)c * ldl];
    }
  }
#pragma unroll
  for (int i = 0; i < TB; ++i)
    if (FULL || i < nb) col[i] = x[i];
}

// Substitution solves with the nb x nb (nb <= 128) triangle staged in LDS
// (read by broadcast) and one lane per right-hand side, 4 waves per
// workgroup, in dtrsm_'s order (src/flablas/f2c/dtrsm.c):
//   LLNU: A12 := inv(trilu(L11)) A12, lane = column: for c, for i > c:
//         x[i] -= x[c] L(i, c);
//   RUNN: A21 := A21 inv(triu(U11)), lane = row (coalesced): for c, for
//         k < c: x[c] -= U(k, c) x[k]; then x[c] *= 1 / U(c, c).
// (dtrsm_ skips a zero multiplier; here it is applied, which differs only
// when the solution holds an infinity or NaN.)
template <bool FULL, bool RIGHT>
__global__ __launch_bounds__(256) void trsm_subst_lds_kernel(int nrhs, int nb,
                                                             const double* __restrict__ T,
                                                             long ldt, double* __restrict__ B,
                                                             long ldb) {
  __shared__ double Ts[TB * TB];  // Ts[c * TB + i] = T(i, c) on the used triangle, else 0
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int i = e % TB, c = e / TB;
    const bool in = RIGHT ? (i <= c) : (i > c);
    Ts[e] = (in && i < nb && c < nb) ? T[(long)i + (long)c * ldt] : 0.0;
  }
  __syncthreads();
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nrhs) return;
  double x[TB];
  if (RIGHT) {
#pragma unroll
    for (int c = 0; c < TB; ++c) x[c] = (FULL || c < nb) ? B[(long)q + (long)c * ldb] : 0.0;
#pragma unroll
    for (int c = 0; c < TB; ++c) {
      if (FULL || c < nb) {
#pragma unroll
        for (int k = 0; k < c; ++k) x[c] -= Ts[c * TB + k] * x[k];
        x[c] *= 1.0 / Ts[c * TB + c];
      }
    }
#pragma unroll
    for (int c = 0; c < TB; ++c)
      if (FULL || c < nb) B[(long)q + (long)c * ldb] = x[c];
  } else {
    double* col = B + (long)q * ldb;
#pragma unroll
    for (int i = 0; i < TB; ++i) x[i] = (FULL || i < nb) ? col[i] : 0.0;
#pragma unroll
    for (int c = 0; c < TB; ++c) {
      if (FULL || c < nb) {
#pragma unroll
        for (int i = c + 1; i < TB; ++i) x[i] -= x[c] * Ts[c * TB + i];
      }
    }
#pragma unroll
    for (int i = 0; i < TB; ++i)
      if (FULL || i < nb) col[i] = x[i];
  }
}

// Diagonal-block solve of the blocked dtrsm, by substitution with the block
// staged in LDS and one lane per right-hand side:
//   left side:  M = op(T_kk), the right-hand sides are the columns of B_k;
//   right side: M = op(T_kk)', the right-hand sides are the rows of B_k
//               (x op(T) = b  <=>  op(T)' x' = b').
// M(i, c) = flip ? T(c, i) : T(i, c).  M lower: forward in dtrsm_'s column
// order (x_c := x_c / M_cc, then x_i -= x_c M_ic for i > c); M upper:
// backward.  Left-side solves divide by the diagonal and right-side solves
// multiply by its reciprocal, as dtrsm_ does (src/flablas/f2c/dtrsm.c).
template <bool MLOWER, bool ROWS>
__global__ __launch_bounds__(256) void trsm_diag_subst_kernel(int nrhs, int nb,
                                                              const double* __restrict__ T,
                                                              long ldt, int flip, int unit,
                                                              double* __restrict__ B, long ldb) {
  __shared__ double Ms[TB * TB];  // Ms[c * TB + i] = M(i, c) on the strict triangle, else 0
  __shared__ double dg[TB];       // M(c, c); 1 for a unit diagonal and beyond nb
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int i = e % TB, c = e / TB;
    const bool in = MLOWER ? (i > c) : (i < c);
    double v = 0.0;
    if (in && i < nb && c < nb) v = flip ? T[(long)c + (long)i * ldt] : T[(long)i + (long)c * ldt];
    Ms[e] = v;
  }
  for (int c = threadIdx.x; c < TB; c += 256)
    dg[c] = (unit || c >= nb) ? 1.0 : T[(long)c * (ldt + 1)];
  __syncthreads();
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= nrhs) return;
  double x[TB];
#pragma unroll
  for (int c = 0; c < TB; ++c)
    x[c] = (c < nb) ? (ROWS ? B[(long)q + (long)c * ldb] : B[(long)c + (long)q * ldb]) : 0.0;
#pragma unroll
  for (int cc = 0; cc < TB; ++cc) {
    const int c = MLOWER ? cc : TB - 1 - cc;
    if (c < nb) {
      if (!unit) x[c] = ROWS ? x[c] * (1.0 / dg[c]) : x[c] / dg[c];
      if (MLOWER) {
#pragma unroll
        for (int i = c + 1; i < TB; ++i) x[i] -= x[c] * Ms[c * TB + i];
      } else {
#pragma unroll
        for (int i = 0; i < c; ++i) x[i] -= x[c] * Ms[c * TB + i];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < TB; ++c)
    if (c < nb) {
      if (ROWS)
        B[(long)q + (long)c * ldb] = x[c];
      else
        B[(long)c + (long)q * ldb] = x[c];
    }
}

// A12 := inv(L11) * A12 for L11 unit lower of order 128 (ld ldl), in place:
// one wave per 16 columns, 16-row blocks,
//   X_i = inv(L_ii) (B_i - sum_{k<i} L_ik X_k),
// both products on v_mfma_f64_16x16x4f64 with L as the first operand.  The
// product D(m, n) lands in lane 16*(m%4) + n, register m/4, which is also the
// second operand's layout for contraction rows 4s + lane/16 in register s, so
// a finished block feeds the next products straight from its registers.  The
// eight inv(L_ii) (16 x 16, unit lower) are formed in LDS by every workgroup.
typedef double v4d_t __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void trsm_llnu128_kernel(int ncols,
                                                           const double* __restrict__ L, long ldl,
                                                           double* __restrict__ B, long ldb) {
  __shared__ double Ld[8][16][17];  // strictly lower parts of the diagonal blocks
  __shared__ double Li[8][16][17];  // their unit lower inverses
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int e = t; e < 8 * 256; e += 256) {
    const int i = e >> 8, c = (e >> 4) & 15, r = e & 15;
    Ld[i][r][c] = (r > c) ? L[(long)(16 * i + r) + (long)(16 * i + c) * ldl] : 0.0;
  }
  __syncthreads();
  if (t < 128) {  // column c of inv(L_ii) by forward substitution
    const int i = t >> 4, c = t & 15;
    double y[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < r; ++q) acc = fma(Ld[i][r][q], y[q], acc);
      y[r] = (r < c) ? 0.0 : ((r == c) ? 1.0 : -acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) Li[i][r][c] = y[r];
  }
  __syncthreads();
  const int c0 = (blockIdx.x * 4 + wave) * 16;
  if (c0 >= ncols) return;
  const int ln = lane & 15, lq = lane >> 4;
  const int col = c0 + ln;
  const bool cv = col < ncols;
  double* bcol = B + (long)col * ldb;
  v4d_t x[8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) x[i][r] = cv ? bcol[16 * i + 4 * r + lq] : 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    v4d_t acc = x[i];
#pragma unroll
    for (int kb = 0; kb < i; ++kb) {
      double p[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) p[s] = -L[(long)(16 * i + ln) + (long)(16 * kb + 4 * s + lq) * ldl];
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(p[s], x[kb][s], acc, 0, 0, 0);
    }
    v4d_t y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
      y = __builtin_amdgcn_mfma_f64_16x16x4f64(Li[i][ln][4 * s + lq], acc[s], y, 0, 0, 0);
    x[i] = y;
  }
  if (cv) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) bcol[16 * i + 4 * r + lq] = x[i][r];
  }
}

}  // namespace

int launch_trsm_llnu(hipStream_t s, int ncols, int nb, const double* L11, long ldl, double* A12,
                     long lda) {
  if (ncols <= 0 || nb <= 0) return FLA_HIP_OK;
  if (nb > TB) return FLA_HIP_EARG;
  if (nb == TB)
    hipLaunchKernelGGL(trsm_llnu128_kernel, dim3((ncols + 63) / 64), dim3(256), 0, s, ncols, L11,
                       ldl, A12, lda);
  else
    hipLaunchKernelGGL(trsm_llnu_kernel<false>, dim3((ncols + 63) / 64), dim3(64), 0, s, ncols,
                       nb, L11, ldl, A12, lda);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

template <bool RIGHT>
static int launch_subst(hipStream_t s, int nrhs, int nb, const double* T, long ldt, double* B,
                        long ldb) {
  if (nrhs <= 0 || nb <= 0) return FLA_HIP_OK;
  if (nb > TB) return FLA_HIP_EARG;
  const dim3 grid((nrhs + 255) / 256);
  if (nb == TB)
    hipLaunchKernelGGL((trsm_subst_lds_kernel<true, RIGHT>), grid, dim3(256), 0, s, nrhs, nb, T,
                       ldt, B, ldb);
  else
    hipLaunchKernelGGL((trsm_subst_lds_kernel<false, RIGHT>), grid, dim3(256), 0, s, nrhs, nb, T,
                       ldt, B, ldb);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

int launch_trsm_llnu_subst(hipStream_t s, int ncols, int nb, const double* L11, long ldl,
                           double* A12, long lda) {
  return launch_subst<false>(s, ncols, nb, L11, ldl, A12, lda);
}

int launch_trsm_runn_rows(hipStream_t s, int nrows, int nb, const double* U11, long ldu,
                          double* A21, long lda) {
  return launch_subst<true>(s, nrows, nb, U11, ldu, A21, lda);
}

int launch_trtri_blocks(hipStream_t s, int na, const double* A, long lda, bool upper, bool unit,
                        double* W, int* ill) {
  if (na <= 0) return FLA_HIP_OK;
  const int nblk = (na + TB - 1) / TB;
  hipLaunchKernelGGL(trtri_blocks_kernel, dim3(nblk), dim3(256), 0, s, na, A, lda, upper ? 1 : 0,
                     unit ? 1 : 0, W, ill);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// Solves the nb x nb diagonal block T (ld ldt) of a triangle against nrhs
// right-hand sides of B in place (left: B is nb x nrhs; right: nrhs x nb).
static int launch_diag_subst(hipStream_t s, bool left, bool upper, bool trans, bool unit, int nb,
                             const double* T, long ldt, double* B, long ldb, int nrhs) {
  if (nb <= 0 || nrhs <= 0) return FLA_HIP_OK;
  const int flip = (left == trans) ? 1 : 0;
  const bool mlower = (!upper && !flip) || (upper && flip);
  const dim3 grid((unsigned)((nrhs + 255) / 256));
  const int u = unit ? 1 : 0;
  if (mlower && left)
    hipLaunchKernelGGL((trsm_diag_subst_kernel<true, false>), grid, dim3(256), 0, s, nrhs, nb, T,
                       ldt, flip, u, B, ldb);
  else if (mlower)
    hipLaunchKernelGGL((trsm_diag_subst_kernel<true, true>), grid, dim3(256), 0, s, nrhs, nb, T,
                       ldt, flip, u, B, ldb);
  else if (left)
    hipLaunchKernelGGL((trsm_diag_subst_kernel<false, false>), grid, dim3(256), 0, s, nrhs, nb, T,
                       ldt, flip, u, B, ldb);
  else
    hipLaunchKernelGGL((trsm_diag_subst_kernel<false, true>), grid, dim3(256), 0, s, nrhs, nb, T,
                       ldt, flip, u, B, ldb);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

static bool is_trans(int t) { return t == kTrans || t == kConjTrans; }

// Scales B by alpha (m x n) when alpha != 1: one MFMA-free elementwise pass.
__global__ void dscal_matrix_kernel(int m, int n, double alpha, double* B, long ldb) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)m * n;
  if (idx >= total) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  double* p = B + (long)i + (long)j * ldb;
  *p = (alpha == 0.0) ? 0.0 : alpha * (*p);
}

static int dtrsm_args_ok(int side, int uplo, int m, int n, int lda, int ldb) {
  if ((side != kLeft && side != kRight) || (uplo != kLower && uplo != kUpper)) return 0;
  if (m < 0 || n < 0) return 0;
  if (m == 0 || n == 0) return 1;
  const int na = (side == kLeft) ? m : n;  // order of A
  return (lda >= na && ldb >= m) ? 1 : 0;
}

// The few-right-hand-side path (the only one that needs a workspace).
static bool dtrsm_narrow(int side, int n, int na) {
  return side == kLeft && n <= kNarrowRhs && na <= kUcSolveRows;
}

// Doubles of workspace dtrsm_body needs for a triangle of order na: the
// inverted 128-blocks, then int words: the wait counter, the wait-failure
// flag and one condition flag per block.
static size_t dtrsm_workspace_doubles(int na) {
  const size_t nblk = (size_t)((na + TB - 1) / TB);
  return nblk * TB * TB + 32 + nblk;
}

// dtrsm on checked, non-empty arguments with the caller's workspace W
// (dtrsm_workspace_doubles(na) doubles of device memory).
static int dtrsm_body(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                      const double* A, int lda, double* B, int ldb, double* W) {
  const int na = (side == kLeft) ? m : n;  // order of A
  FlaHipCtx& c = fla_ctx();
  hipStream_t s = c.s_update;
  if (alpha != 1.0) {
    const long total = (long)m * n;
    hipLaunchKernelGGL(dscal_matrix_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       m, n, alpha, B, (long)ldb);
    FLA_HIP_CHECK(hipGetLastError());
    if (alpha == 0.0) return FLA_HIP_OK;
  }
  const bool tr = is_trans(trans);
  const bool upper = (uplo == kUpper);
  const bool unit = (diag == kUnit);
  const int nblk = (na + TB - 1) / TB;
  // op(A) is lower when (lower, N) or (upper, T): forward substitution.
  const bool op_lower = (!upper && !tr) || (upper && tr);
  if (dtrsm_narrow(side, n, na)) {
    // Few right-hand sides: one persistent launch per 16 columns that
    // multiplies by the inverted diagonal blocks where their condition
    // number allows and substitutes through the others.
    int* flags = reinterpret_cast<int*>(W + (long)nblk * TB * TB);
    int* ill = flags + 2;
    FLA_TRY(launch_trtri_blocks(s, na, A, lda, upper, unit, W, ill));
    FLA_TRY(launch_trsm_left_narrow(s, op_lower, tr, na, n, A, lda, W, B, ldb, flags, flags + 1,
                                    ill, unit));
    int herr = 0;
    FLA_HIP_CHECK(hipMemcpyAsync(&herr, flags + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    FLA_HIP_CHECK(hipStreamSynchronize(s));
    return herr ? FLA_HIP_ESPIN : FLA_HIP_OK;
  }
  auto blk = [&](int idx) { return idx * TB; };
  if (side == kLeft) {
    // op(A) X = B.  Forward when op(A) is lower.
    for (int q = 0; q < nblk; ++q) {
      const int kb = op_lower ? q : (nblk - 1 - q);
      const int r0 = blk(kb);
      const int nbk = (na - r0) < TB ? (na - r0) : TB;
      // X_k := inv(op(T_kk)) * B_k by substitution, in place
      FLA_TRY(launch_diag_subst(s, true, upper, tr, unit, nbk, A + (long)r0 * (lda + 1), lda,
                                B + r0, ldb, n));
      // B_rest -= op(A)[rest, k] * X_k
      if (op_lower) {
        const int rr = r0 + nbk, mr = na - rr;
        if (mr > 0) {
          // op(A)(i, c) for i in rest, c in block: N -> A[i + c*lda]; T -> A[c + i*lda]
          const double* Ap = tr ? A + (long)r0 + (long)rr * lda : A + (long)rr + (long)r0 * lda;
          FLA_TRY(launch_dgemm_mfma(s, tr, false, 0, mr, n, nbk, -1.0, Ap, lda, B + r0, ldb, 1.0,
                                    B + rr, ldb, nullptr));
        }
      } else {
        const int mr = r0;  // rows above
        if (mr > 0) {
          const double* Ap = tr ? A + (long)r0 : A + (long)r0 * lda;
          FLA_TRY(launch_dgemm_mfma(s, tr, false, 0, mr, n, nbk, -1.0, Ap, lda, B + r0, ldb, 1.0,
                                    B, ldb, nullptr));
        }
      }
    }
  } else {
    // X op(A) = B.  Forward over columns when op(A) is upper.
    const bool fwd = !op_lower;
    for (int q = 0; q < nblk; ++q) {
      const int kb = fwd ? q : (nblk - 1 - q);
      const int c0 = blk(kb);
      const int nbk = (na - c0) < TB ? (na - c0) : TB;
      double* Bk = B + (long)c0 * ldb;
      // X_k := B_k * inv(op(T_kk)) by substitution, in place
      FLA_TRY(launch_diag_subst(s, false, upper, tr, unit, nbk, A + (long)c0 * (lda + 1), lda, Bk,
                                ldb, m));
      if (fwd) {
        const int cc = c0 + nbk, nr = na - cc;
        if (nr > 0) {
          // B[:, rest] -= X_k * op(A)[k, rest];  op(A)(c, j): N -> A[c + j*lda], T -> A[j + c*lda]
          const double* Ap = tr ? A + (long)cc + (long)c0 * lda : A + (long)c0 + (long)cc * lda;
          FLA_TRY(launch_dgemm_mfma(s, false, tr, 0, m, nr, nbk, -1.0, Bk, ldb, Ap, lda, 1.0,
                                    B + (long)cc * ldb, ldb, nullptr));
        }
      } else {
        const int nr = c0;
        if (nr > 0) {
          const double* Ap = tr ? A + (long)c0 * lda : A + (long)c0;
          FLA_TRY(launch_dgemm_mfma(s, false, tr, 0, m, nr, nbk, -1.0, Bk, ldb, Ap, lda, 1.0, B,
                                    ldb, nullptr));
        }
      }
    }
  }
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                             const double* A, int lda, double* B, int ldb) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (!dtrsm_args_ok(side, uplo, m, n, lda, ldb)) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const int na = (side == kLeft) ? m : n;
  void* w = nullptr;
  if (dtrsm_narrow(side, n, na))
    FLA_TRY(fla_workspace(dtrsm_workspace_doubles(na) * sizeof(double), &w));
  return dtrsm_body(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb,
                    static_cast<double*>(w));
}

// ---- dtrsm without allocation ---------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): fla_hip_dtrsm_noalloc runs the same kernels
// as fla_hip_dtrsm with its workspace in `scratch`, a file-scope device array
// sized for triangles of order <= kNoAllocMaxOrder (8 MB; every GPU the
// library runs on holds its own copy until the process ends).  Host code
// never passes `scratch` itself to a kernel (in host code the name is a
// host-side stand-in): fla_noalloc_setup, called by fla_hip_init with the
// device current, resolves the array's device address once per device, and
// only that address is used.  Every call on one GPU shares the array: calls
// that overlap on a GPU (from several host threads) race.  The wait words
// are cleared by a kernel at the start of each solve.
namespace {
constexpr int kNoAllocMaxOrder = 8192;
constexpr size_t kNoAllocDoubles = (size_t)(kNoAllocMaxOrder / TB) * TB * TB + 32 +
                                   kNoAllocMaxOrder / TB;
__device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kNoAllocDevices = 64;
double* g_scratch_ptr[kNoAllocDevices] = {};
}  // namespace

int fla_noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocDevices) return FLA_HIP_EARG;
  if (g_scratch_ptr[device]) return FLA_HIP_OK;
  void* p = nullptr;
  FLA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return FLA_HIP_EARG;
  g_scratch_ptr[device] = static_cast<double*>(p);
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dtrsm_noalloc(int side, int uplo, int trans, int diag, int m, int n,
                                     double alpha, const double* A, int lda, double* B,
                                     int ldb) {
  FLA_USER_STREAM_SCOPE;
  FlaHipCtx& c = fla_ctx();
  // no implicit initialization here: fla_hip_init allocates streams and buffers
  if (!c.ready || c.device < 0 || c.device >= kNoAllocDevices || !g_scratch_ptr[c.device]) {
    std::fprintf(stderr,
                 "libflame_amd: fla_hip_dtrsm_noalloc needs FLA_Init / fla_hip_init first\n");
    return FLA_HIP_ENOINIT;
  }
  if (!dtrsm_args_ok(side, uplo, m, n, lda, ldb)) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const int na = (side == kLeft) ? m : n;
  if (na > kNoAllocMaxOrder) {
    std::fprintf(stderr,
                 "libflame_amd: fla_hip_dtrsm_noalloc: triangle of order %d exceeds its scratch "
                 "limit of %d\n",
                 na, kNoAllocMaxOrder);
    return FLA_HIP_ESCRATCH;
  }
  return dtrsm_body(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb,
                    g_scratch_ptr[c.device]);
}
