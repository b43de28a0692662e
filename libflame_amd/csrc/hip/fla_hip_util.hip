// Device-side utilities: counter-based input generators (the same integer
// hash as the CPU generators, so uniform inputs are bit-identical on host and
// device), Gram-matrix construction, FP64 residual products, pivot shifts and
// the general row-interchange kernel behind FLA_Apply_pivots / dlaswp_.
#include "fla_hip_internal.h"

#include <vector>

namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long rand_u64(unsigned long long seed,
                                                       unsigned long long idx) {
  return mix64(mix64(seed + 0x9E3779B97F4A7C15ULL) ^
               (idx * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
}

__global__ void gen_kernel(unsigned long long seed, int m, int n, double* A, long lda,
                           int normal) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)m * n;
  if (idx >= total) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  const unsigned long long e = (unsigned long long)j * (unsigned long long)m + (unsigned long long)i;
  double v;
  if (!normal) {
    v = (double)(rand_u64(seed, e) >> 11) * 0x1.0p-52 - 1.0;
  } else {
    const double u1 = ((double)(rand_u64(seed, 2 * e) >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(rand_u64(seed, 2 * e + 1) >> 11) * 0x1.0p-53;
    v = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  A[(long)i + (long)j * lda] = v;
}

__global__ void add_diag_kernel(int n, double shift, double* A, long lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(long)i + (long)i * lda] += shift;
}

// A(i, j) *= s[i]
__global__ void scale_rows_kernel(int m, int n, const double* __restrict__ s, double* A,
                                  long lda) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)m * n;
  if (idx >= total) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  A[(long)i + (long)j * lda] *= s[i];
}

__global__ void shift_pivots_kernel(int to_lapack, int m, int* p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] += to_lapack ? (i + 1) : -(i + 1);
}

// ---- row interchanges (FLA_Apply_pivots / dlaswp) on the device ----------
// The interchange sequence: step q = 0..K-1 (in application order) swaps rows
// r(q) and t(q); forward order r(q) = k1 + q, reverse order r(q) = k2 - q, and
// t(q) = r(q) + p[(r(q) - k1) * stride] (p points at row k1's pivot, so the
// caller's vector holds row i's pivot at i * stride, as
// FLA_Apply_pivots_ln_opt_var1.c:313-330 reads it).  A step
// whose t lies outside [0, rows) is skipped, so no row outside the matrix is
// ever addressed.
struct SwapSeq {
  const int* p;
  int k1, k2, stride, rev, rows;
  __device__ __forceinline__ void step(int q, int& r, int& t) const {
    r = rev ? k2 - q : k1 + q;
    t = r + p[(long)(r - k1) * stride];
    if (t < 0 || t >= rows) t = r;
  }
};

constexpr int kSwapChunk = 1024;

// Final source row of every touched row, without walking the sequence on one
// thread: slot q < K resolves row r(q); slot K + q resolves t(q) when it lies
// outside [k1, k2] (rows inside are resolved by their own slot).  Each slot
// follows its row backwards through the whole sequence (x = t at a step whose
// r is x, x = r at a step whose t is x); the steps are staged through LDS in
// chunks and read as LDS broadcasts.  pos[slot] = -1 marks an unused slot.
__global__ __launch_bounds__(256) void laswp_resolve_kernel(SwapSeq S, int K, int* __restrict__ pos,
                                                            int* __restrict__ src) {
  __shared__ int sr[kSwapChunk], st[kSwapChunk];
  const int slot = blockIdx.x * 256 + threadIdx.x;
  int x0 = -1;
  if (slot < 2 * K) {
    const int q = slot < K ? slot : slot - K;
    int r, t;
    S.step(q, r, t);
    if (slot < K)
      x0 = r;
    else
      x0 = (t == r || (t >= S.k1 && t <= S.k2)) ? -1 : t;
  }
  int x = x0;
  for (int c0 = ((K - 1) / kSwapChunk) * kSwapChunk; c0 >= 0; c0 -= kSwapChunk) {
    const int cn = min(kSwapChunk, K - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < cn; e += 256) {
      int r, t;
      S.step(c0 + e, r, t);
      sr[e] = r;
      st[e] = t;
    }
    __syncthreads();
    if (x0 >= 0) {
      for (int e = cn - 1; e >= 0; --e) {
        const int r = sr[e], t = st[e];
        x = (x == r) ? t : ((x == t) ? r : x);
      }
    }
  }
  if (slot < 2 * K) {
    pos[slot] = x0;
    src[slot] = x;
  }
}

// tmp(v, j) := A(src[v], j) for every slot whose row changes.
__global__ __launch_bounds__(256) void laswp_gather_kernel(int nslot, int ncols,
                                                           const int* __restrict__ pos,
                                                           const int* __restrict__ src,
                                                           const double* __restrict__ A, long lda,
                                                           double* __restrict__ tmp) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nslot) return;
  const int x = pos[v], y = src[v];
  if (x < 0 || x == y) return;
  for (int j = blockIdx.y; j < ncols; j += gridDim.y)
    tmp[(long)v + (long)j * nslot] = A[(long)y + (long)j * lda];
}

// A(pos[v], j) := tmp(v, j).  Two slots naming one row carry identical data.
__global__ __launch_bounds__(256) void laswp_scatter_kernel(int nslot, int ncols,
                                                            const int* __restrict__ pos,
                                                            const int* __restrict__ src,
                                                            const double* __restrict__ tmp,
                                                            double* __restrict__ A, long lda) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nslot) return;
  const int x = pos[v], y = src[v];
  if (x < 0 || x == y) return;
  for (int j = blockIdx.y; j < ncols; j += gridDim.y)
    A[(long)x + (long)j * lda] = tmp[(long)v + (long)j * nslot];
}

// ---- FP64 reference products for residual checks -----------------------
// op(M)(i, j) for kind 0 (full), 1 (symmetric from the uplo triangle),
// 2 (triangular uplo, unit or not).
struct MatSpec {
  int kind, lower, trans, unit;
  const double* M;
  long ld;
  __device__ __forceinline__ double get(int i, int j) const {
    if (trans && kind != 1) {
      const int t = i;
      i = j;
      j = t;
    }
    if (kind == 0) return M[(long)i + (long)j * ld];
    if (kind == 1) {
      const bool in = lower ? (i >= j) : (i <= j);
      return in ? M[(long)i + (long)j * ld] : M[(long)j + (long)i * ld];
    }
    const bool in = lower ? (i >= j) : (i <= j);
    if (!in) return 0.0;
    if (i == j && unit) return 1.0;
    return M[(long)i + (long)j * ld];
  }
};

constexpr int CT = 64;  // tile order of the check product

// Y(i0:i0+64, :) = beta*Y + op(M)(i0:, :) * X, X with nrhs <= 8 columns.
__global__ __launch_bounds__(256) void matmul_check_kernel(MatSpec S, int n, int nrhs,
                                                           const double* __restrict__ X, long ldx,
                                                           double beta, double* __restrict__ Y,
                                                           long ldy) {
  __shared__ double T[CT][CT + 1];
  __shared__ double Xs[CT][8];
  const int i0 = blockIdx.x * CT;
  const int t = threadIdx.x;
  const int ri = t & (CT - 1);  // row within tile
  const int rg = t >> 6;        // 0..3: rhs group (2 rhs each)
  double acc[2] = {0.0, 0.0};
  for (int j0 = 0; j0 < n; j0 += CT) {
    // Load the tile op(M)(i0.., j0..) choosing a coalesced orientation.
    for (int e = t; e < CT * CT; e += 256) {
      const int a = e & (CT - 1), b = e >> 6;
      // Try "i fastest" (a = row) when the source is M[i + j*ld], else "j fastest".
      int i = i0 + a, j = j0 + b;
      double v = 0.0;
      if (i < n && j < n) v = S.get(i, j);
      T[a][b] = v;
    }
    for (int e = t; e < CT * 8; e += 256) {
      const int a = e & (CT - 1), r = e >> 6;
      Xs[a][r] = (j0 + a < n && r < nrhs) ? X[(long)(j0 + a) + (long)r * ldx] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int b = 0; b < CT; ++b) {
      const double mv = T[ri][b];
      acc[0] += mv * Xs[b][2 * rg];
      acc[1] += mv * Xs[b][2 * rg + 1];
    }
    __syncthreads();
  }
  const int i = i0 + ri;
  if (i < n) {
    for (int q = 0; q < 2; ++q) {
      const int r = 2 * rg + q;
      if (r < nrhs) {
        double* y = Y + (long)i + (long)r * ldy;
        *y = (beta == 0.0) ? acc[q] : acc[q] + beta * (*y);
      }
    }
  }
}

// ---- A = Q diag(s) Q^T with Q a product of Householder reflectors ----------

// A := diag(kappa^(i/(n-1))) (full matrix).
__global__ void logdiag_kernel(int n, double kappa, double* A, long lda) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  double v = 0.0;
  if (i == j) v = (n > 1) ? pow(kappa, (double)i / (double)(n - 1)) : 1.0;
  A[(long)i + (long)j * lda] = v;
}

// out[0] = v.w, out[1] = v.v (one workgroup, fixed order).
__global__ void dot2_kernel(int n, const double* __restrict__ v, const double* __restrict__ w,
                            double* __restrict__ out) {
  __shared__ double r0[256], r1[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    a += v[i] * w[i];
    b += v[i] * v[i];
  }
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r0[threadIdx.x] += r0[threadIdx.x + st];
      r1[threadIdx.x] += r1[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r0[0];
    out[1] = r1[0];
  }
}

// z = tau w - (tau^2 alpha / 2) v with tau = 2 / (v.v), alpha = v.w.
__global__ void refl_z_kernel(int n, const double* __restrict__ v, const double* __restrict__ w,
                              const double* __restrict__ d, double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tau = 2.0 / d[1];
  z[i] = tau * w[i] - 0.5 * tau * tau * d[0] * v[i];
}

// A -= tau x y^T with tau = 2 / d[1] (d[1] = v.v of the reflector).
__global__ void rank1_kernel(int n, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ d, double* __restrict__ A, long lda) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  A[(long)i + (long)j * lda] -= (2.0 / d[1]) * x[i] * y[j];
}

// w[j] = sum_i A(i, j) v[i] (one workgroup per column, fixed order).
__global__ void gemv_t_kernel(int n, const double* __restrict__ A, long lda,
                              const double* __restrict__ v, double* __restrict__ w) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += A[(long)i + (long)j * lda] * v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) w[j] = red[0];
}

// A -= v z^T + z v^T
__global__ void rank2_kernel(int n, const double* __restrict__ v, const double* __restrict__ z,
                             double* __restrict__ A, long lda) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  A[(long)i + (long)j * lda] -= v[i] * z[j] + z[i] * v[j];
}

}  // namespace

extern "C" int fla_hip_gen_uniform(unsigned long long seed, int m, int n, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(gen_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, seed, m, n, A, (long)lda, 0);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_gen_normal(unsigned long long seed, int m, int n, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(gen_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, seed, m, n, A, (long)lda, 1);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_gram_lower(int n, int k, const double* B, int ldb, double scale,
                                  double shift, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n <= 0) return FLA_HIP_OK;
  hipStream_t s = fla_ctx().s_update;
  FLA_TRY(launch_dgemm_mfma(s, false, true, 1, n, n, k, scale, B, ldb, B, ldb, 0.0, A, lda,
                            nullptr));
  hipLaunchKernelGGL(add_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, shift, A,
                     (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// The row scales 10^(-i/max(m,n)) come from the host C library's pow, so the
// generated matrix has the same bytes as a host generator's (the device pow
// rounds differently in the last bit for some rows).
extern "C" int fla_hip_scale_rows_logspace(int m, int n, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  std::vector<double> h((size_t)m);
  const double mx = (double)(m > n ? m : n);
  for (int i = 0; i < m; ++i) h[(size_t)i] = std::pow(10.0, -(double)i / mx);
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(3, (size_t)m * sizeof(double), &w));
  hipStream_t s = fla_ctx().s_update;
  FLA_HIP_CHECK(hipMemcpyAsync(w, h.data(), (size_t)m * sizeof(double), hipMemcpyHostToDevice, s));
  const long total = (long)m * n;
  hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, m,
                     n, static_cast<const double*>(w), A, (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  FLA_HIP_CHECK(hipStreamSynchronize(s));  // h is released on return
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dmatmul_check(int kind, int uplo, int trans, int diag, int n, int nrhs,
                                     const double* M, int ldm, const double* X, int ldx,
                                     double beta, double* Y, int ldy) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n <= 0 || nrhs <= 0) return FLA_HIP_OK;
  if (nrhs > 8 || kind < 0 || kind > 2) return FLA_HIP_EARG;
  MatSpec S;
  S.kind = kind;
  S.lower = (uplo != kUpper);
  S.trans = (trans == kTrans || trans == kConjTrans);
  S.unit = (diag == kUnit);
  S.M = M;
  S.ld = ldm;
  hipLaunchKernelGGL(matmul_check_kernel, dim3((n + CT - 1) / CT), dim3(256), 0,
                     fla_ctx().s_update, S, n, nrhs, X, (long)ldx, beta, Y, (long)ldy);
  FLA_HIP_CHECK(hipGetLastError());
  return fla_hip_synchronize();
}

extern "C" int fla_hip_shift_pivots(int ptype, int m, int* p) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m <= 0) return FLA_HIP_OK;
  if (ptype != kNativePivots && ptype != kLapackPivots) return FLA_HIP_EARG;
  hipLaunchKernelGGL(shift_pivots_kernel, dim3((m + 255) / 256), dim3(256), 0, fla_ctx().s_update,
                     ptype == kLapackPivots ? 1 : 0, m, p);
  FLA_HIP_CHECK(hipGetLastError());
  return fla_hip_synchronize();
}

// Row interchanges i <-> i + p[i * |incp|] for i = k1..k2 (reverse
// order for incp < 0) on the ncols columns of A, rows limited to [0, rows):
// the final source of each touched row is resolved on the device and only
// those rows move (gather into a workspace, then scatter), asynchronously on
// stream s.  Replaces the walk of FLA_Apply_pivots_ln_opt_var1
// (src/lapack/util/app/piv/ln/flamec/FLA_Apply_pivots_ln_opt_var1.c:313-330).
int launch_laswp(hipStream_t s, int rows, int ncols, double* A, long lda, int k1, int k2,
                 const int* p, int incp) {
  if (ncols <= 0 || k2 < k1 || incp == 0 || rows <= 0) return FLA_HIP_OK;
  if (k1 < 0 || k2 >= rows) return FLA_HIP_EARG;
  const int K = k2 - k1 + 1;
  if (K > (1 << 29)) return FLA_HIP_EARG;
  const int nslot = 2 * K;
  const size_t idx_bytes = ((size_t)nslot * sizeof(int) + 255) & ~(size_t)255;
  void* w = nullptr;
  FLA_TRY(fla_workspace(2 * idx_bytes + (size_t)nslot * ncols * sizeof(double), &w));
  int* pos = static_cast<int*>(w);
  int* src = reinterpret_cast<int*>(static_cast<char*>(w) + idx_bytes);
  double* tmp = reinterpret_cast<double*>(static_cast<char*>(w) + 2 * idx_bytes);
  SwapSeq S;
  S.p = p + (long)k1 * (incp > 0 ? incp : -incp);
  S.k1 = k1;
  S.k2 = k2;
  S.stride = incp > 0 ? incp : -incp;
  S.rev = incp < 0;
  S.rows = rows;
  const unsigned gs = (unsigned)((nslot + 255) / 256);
  hipLaunchKernelGGL(laswp_resolve_kernel, dim3(gs), dim3(256), 0, s, S, K, pos, src);
  const unsigned gy = (unsigned)(ncols < 64 ? ncols : 64);
  hipLaunchKernelGGL(laswp_gather_kernel, dim3(gs, gy), dim3(256), 0, s, nslot, ncols, pos, src,
                     A, lda, tmp);
  hipLaunchKernelGGL(laswp_scatter_kernel, dim3(gs, gy), dim3(256), 0, s, nslot, ncols, pos, src,
                     tmp, A, lda);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dlaswp(int ncols, double* A, int lda, int k1, int k2, const int* p,
                              int incp) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  // Every row index must lie inside the leading dimension.
  return launch_laswp(fla_ctx().s_update, lda, ncols, A, lda, k1, k2, p, incp);
}

extern "C" int fla_hip_dlaswp_rows(int m, int ncols, double* A, int lda, int k1, int k2,
                                   const int* p, int incp) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m > lda) return FLA_HIP_EARG;
  return launch_laswp(fla_ctx().s_update, m, ncols, A, lda, k1, k2, p, incp);
}

// Symmetric positive definite A = Q diag(s) Q^T, s_i = kappa^(i/(n-1)),
// Q = H_1 ... H_nrefl with H_r = I - 2 v v^T / (v^T v), v ~ N(0,1) drawn by
// fla_hip_gen_normal(seed + r + 1, n, 1, ...).  Each reflector is applied as
// A := H A H = A - v z^T - z v^T.  Both triangles are written.
extern "C" int fla_hip_gen_spd_reflectors(unsigned long long seed, int n, int nrefl, double kappa,
                                          double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n <= 0) return FLA_HIP_OK;
  if (nrefl < 0 || lda < n || !(kappa >= 1.0)) return FLA_HIP_EARG;
  hipStream_t s = fla_ctx().s_update;
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(3, sizeof(double) * (3 * (size_t)n + 8), &w));
  double* v = static_cast<double*>(w);
  double* wv = v + n;
  double* z = wv + n;
  double* d = z + n;
  const long nn = (long)n * n;
  hipLaunchKernelGGL(logdiag_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, n,
                     kappa, A, (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  for (int r = 0; r < nrefl; ++r) {
    hipLaunchKernelGGL(gen_kernel, dim3((n + 255) / 256), dim3(256), 0, s, seed + r + 1, n, 1, v,
                       (long)n, 1);
    FLA_TRY(launch_dgemm_mfma(s, false, false, 0, n, 1, n, 1.0, A, lda, v, n, 0.0, wv, n,
                              nullptr));
    hipLaunchKernelGGL(dot2_kernel, dim3(1), dim3(256), 0, s, n, v, wv, d);
    hipLaunchKernelGGL(refl_z_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, v, wv, d, z);
    hipLaunchKernelGGL(rank2_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, n, v, z,
                       A, (long)lda);
    FLA_HIP_CHECK(hipGetLastError());
  }
  return FLA_HIP_OK;
}

// General A = Q1 diag(s) Q2^T: A := diag(s), then for each reflector pair
// A := H_r A (A -= tau v (A^T v)^T) and A := A G_r (A -= tau (A u) u^T), with
// v from seed + r + 1 and u from seed2 + r + 1 (fla_hip_gen_normal).
extern "C" int fla_hip_gen_ge_reflectors(unsigned long long seed, unsigned long long seed2, int n,
                                         int nrefl, double kappa, double* A, int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n <= 0) return FLA_HIP_OK;
  if (nrefl < 0 || lda < n || !(kappa >= 1.0)) return FLA_HIP_EARG;
  hipStream_t s = fla_ctx().s_update;
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(3, sizeof(double) * (2 * (size_t)n + 8), &w));
  double* v = static_cast<double*>(w);
  double* wv = v + n;
  double* d = wv + n;
  const long nn = (long)n * n;
  const unsigned gnn = (unsigned)((nn + 255) / 256);
  hipLaunchKernelGGL(logdiag_kernel, dim3(gnn), dim3(256), 0, s, n, kappa, A, (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  for (int r = 0; r < nrefl; ++r) {
    // left: A := (I - tau v v^T) A
    hipLaunchKernelGGL(gen_kernel, dim3((n + 255) / 256), dim3(256), 0, s, seed + r + 1, n, 1, v,
                       (long)n, 1);
    hipLaunchKernelGGL(gemv_t_kernel, dim3(n), dim3(256), 0, s, n, A, (long)lda, v, wv);
    hipLaunchKernelGGL(dot2_kernel, dim3(1), dim3(256), 0, s, n, v, v, d);
    hipLaunchKernelGGL(rank1_kernel, dim3(gnn), dim3(256), 0, s, n, v, wv, d, A, (long)lda);
    // right: A := A (I - tau u u^T)
    hipLaunchKernelGGL(gen_kernel, dim3((n + 255) / 256), dim3(256), 0, s, seed2 + r + 1, n, 1, v,
                       (long)n, 1);
    FLA_TRY(launch_dgemm_mfma(s, false, false, 0, n, 1, n, 1.0, A, lda, v, n, 0.0, wv, n,
                              nullptr));
    hipLaunchKernelGGL(dot2_kernel, dim3(1), dim3(256), 0, s, n, v, v, d);
    hipLaunchKernelGGL(rank1_kernel, dim3(gnn), dim3(256), 0, s, n, wv, v, d, A, (long)lda);
    FLA_HIP_CHECK(hipGetLastError());
  }
  return FLA_HIP_OK;
}

namespace {
__global__ void d2s_kernel(int m, int n, const double* __restrict__ A, long lda,
                           float* __restrict__ S, long lds) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * n) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  S[(long)i + (long)j * lds] = (float)A[(long)i + (long)j * lda];
}
}  // namespace

// S := float(A) for an m x n matrix (round to nearest).  A and S must not
// overlap.
extern "C" int fla_hip_convert_d2s(int m, int n, const double* A, int lda, float* S, int lds) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || lds < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(d2s_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, m, n, A, (long)lda, S, (long)lds);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// ---- LAPACK auxiliaries on device matrices: dlange / dlansy norms, dlacpy,
// dlaset (netlib semantics: src/map/common/lapacksrc/f2c/dlange.c,
// dlansy.c, dlacpy.c, dlaset.c).  The norms reduce per column (or per row
// for the row sums of a stored triangle) and then over columns; NaN
// propagates.
namespace {

// Row range of column j that a reduction covers: 0 = all m rows,
// 1 = rows j..m-1 (lower triangle with the diagonal), 2 = rows 0..j (upper).
__device__ __forceinline__ void col_range(int mode, int m, int j, int& r0, int& r1) {
  r0 = mode == 1 ? j : 0;
  r1 = mode == 2 ? (j + 1 < m ? j + 1 : m) : m;
}

__device__ __forceinline__ double nanmax(double a, double b) {
  return (b > a || b != b) ? b : a;
}

// op 0: sum |a|, 1: max |a|, 2: sum a^2, over the covered rows of column j.
__global__ __launch_bounds__(256) void col_reduce_kernel(int op, int mode, int m, int n,
                                                         const double* __restrict__ A, long lda,
                                                         double* __restrict__ out) {
  const int j = blockIdx.x;
  if (j >= n) return;
  int r0, r1;
  col_range(mode, m, j, r0, r1);
  double acc = 0.0;
  for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) {
    const double a = A[(long)i + (long)j * lda];
    acc = op == 0 ? acc + fabs(a) : (op == 1 ? nanmax(acc, fabs(a)) : acc + a * a);
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double b = red[threadIdx.x + s];
      red[threadIdx.x] = op == 1 ? nanmax(red[threadIdx.x], b) : red[threadIdx.x] + b;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// Sum of |A(i, c)| over the columns c of row i: mode 0 all n columns,
// 1 columns 0..i-1 (strictly lower part of the row), 2 columns i+1..n-1.
__global__ __launch_bounds__(256) void row_abssum_kernel(int mode, int m, int n,
                                                         const double* __restrict__ A, long lda,
                                                         double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int c0 = mode == 2 ? i + 1 : 0;
  const int c1 = mode == 1 ? (i < n ? i : n) : n;
  double acc = 0.0;
  for (int c = c0; c < c1; ++c) acc += fabs(A[(long)i + (long)c * lda]);
  out[i] = acc;
}

// out[0] = op over k < cnt of (a[k] + (b ? b[k] : 0)): op 1 max, 0 sum.
__global__ __launch_bounds__(256) void final_reduce_kernel(int op, int cnt,
                                                           const double* __restrict__ a,
                                                           const double* __restrict__ b,
                                                           double* __restrict__ out) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < cnt; k += 256) {
    const double v = a[k] + (b ? b[k] : 0.0);
    acc = op == 1 ? nanmax(acc, v) : acc + v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double c = red[threadIdx.x + s];
      red[threadIdx.x] = op == 1 ? nanmax(red[threadIdx.x], c) : red[threadIdx.x] + c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// mode 0 full, 1 lower trapezoid (i >= j), 2 upper trapezoid (i <= j).
__global__ void lacpy_kernel(int mode, int m, int n, const double* __restrict__ A, long lda,
                             double* __restrict__ B, long ldb) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * n) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  if ((mode == 1 && i < j) || (mode == 2 && i > j)) return;
  B[(long)i + (long)j * ldb] = A[(long)i + (long)j * lda];
}

// dlaset: off-diagonal entries of the selected part := alpha, diagonal := beta.
__global__ void laset_kernel(int mode, int m, int n, double alpha, double beta,
                             double* __restrict__ A, long lda) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)m * n) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  if (i == j)
    A[(long)i + (long)j * lda] = beta;
  else if (mode == 0 || (mode == 1 && i > j) || (mode == 2 && i < j))
    A[(long)i + (long)j * lda] = alpha;
}

int norm_kind(int norm) {
  switch (norm) {
    case '1': case 'O': case 'o': return 1;
    case 'I': case 'i': return 2;
    case 'M': case 'm': return 3;
    case 'F': case 'f': case 'E': case 'e': return 4;
    default: return 0;
  }
}

int uplo_mode(int uplo) { return uplo == kLower ? 1 : (uplo == kUpper ? 2 : 0); }

// Shared driver of dlange / dlansy: m x n, sym = 0 general, else the
// symmetric matrix stored in the uplo triangle (m == n).
int norm_driver(int kind, int sym_mode, int m, int n, const double* A, long lda, double* value) {
  hipStream_t s = fla_ctx().s_update;
  void* w = nullptr;
  const size_t len = (size_t)(m > n ? m : n);
  FLA_TRY(fla_workspace_slot(3, (2 * len + 8) * sizeof(double), &w));
  double* colv = static_cast<double*>(w);
  double* rowv = colv + len;
  double* res = rowv + len;
  const bool onecol = (kind == 1) || (sym_mode != 0 && kind == 2);  // symmetric: ||.||_1 == ||.||_inf
  if (onecol) {
    hipLaunchKernelGGL(col_reduce_kernel, dim3(n), dim3(256), 0, s, 0, sym_mode, m, n, A, lda,
                       colv);
    if (sym_mode != 0)
      hipLaunchKernelGGL(row_abssum_kernel, dim3((m + 255) / 256), dim3(256), 0, s, sym_mode, m,
                         n, A, lda, rowv);
    hipLaunchKernelGGL(final_reduce_kernel, dim3(1), dim3(256), 0, s, 1, n, colv,
                       sym_mode != 0 ? rowv : nullptr, res);
  } else if (kind == 2) {
    hipLaunchKernelGGL(row_abssum_kernel, dim3((m + 255) / 256), dim3(256), 0, s, 0, m, n, A, lda,
                       rowv);
    hipLaunchKernelGGL(final_reduce_kernel, dim3(1), dim3(256), 0, s, 1, m, rowv, nullptr, res);
  } else if (kind == 3) {
    hipLaunchKernelGGL(col_reduce_kernel, dim3(n), dim3(256), 0, s, 1, sym_mode, m, n, A, lda,
                       colv);
    hipLaunchKernelGGL(final_reduce_kernel, dim3(1), dim3(256), 0, s, 1, n, colv, nullptr, res);
  } else {
    // Frobenius: a symmetric matrix counts each stored off-diagonal entry twice
    hipLaunchKernelGGL(col_reduce_kernel, dim3(n), dim3(256), 0, s, 2, sym_mode, m, n, A, lda,
                       colv);
    hipLaunchKernelGGL(final_reduce_kernel, dim3(1), dim3(256), 0, s, 0, n, colv, nullptr, res);
    if (sym_mode != 0) {
      hipLaunchKernelGGL(col_reduce_kernel, dim3(n), dim3(256), 0, s, 2, 0, 1, n, A, lda + 1,
                         rowv);  // the diagonal: a 1 x n matrix with stride lda + 1
      hipLaunchKernelGGL(final_reduce_kernel, dim3(1), dim3(256), 0, s, 0, n, rowv, nullptr,
                         res + 1);
    }
  }
  FLA_HIP_CHECK(hipGetLastError());
  double h[2] = {0.0, 0.0};
  FLA_HIP_CHECK(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  if (kind == 4) *value = sym_mode != 0 ? sqrt(2.0 * h[0] - h[1]) : sqrt(h[0]);
  else *value = h[0];
  return FLA_HIP_OK;
}

}  // namespace

extern "C" int fla_hip_dlange(int norm, int m, int n, const double* A, int lda, double* value) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  const int kind = norm_kind(norm);
  if (!value || kind == 0 || m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  *value = 0.0;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  return norm_driver(kind, 0, m, n, A, lda, value);
}

extern "C" int fla_hip_dlansy(int norm, int uplo, int n, const double* A, int lda,
                              double* value) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  const int kind = norm_kind(norm);
  if (!value || kind == 0 || (uplo != kLower && uplo != kUpper) || n < 0 ||
      lda < (n > 1 ? n : 1))
    return FLA_HIP_EARG;
  *value = 0.0;
  if (n == 0) return FLA_HIP_OK;
  return norm_driver(kind, uplo_mode(uplo), n, n, A, lda, value);
}

extern "C" int fla_hip_dlacpy(int uplo, int m, int n, const double* A, int lda, double* B,
                              int ldb) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || ldb < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(lacpy_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, uplo_mode(uplo), m, n, A, (long)lda, B, (long)ldb);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dlaset(int uplo, int m, int n, double alpha, double beta, double* A,
                              int lda) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(laset_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, uplo_mode(uplo), m, n, alpha, beta, A, (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}
