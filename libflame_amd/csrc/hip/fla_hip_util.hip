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

__global__ void scale_rows_kernel(int m, int n, double* A, long lda) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)m * n;
  if (idx >= total) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  const double s = pow(10.0, -(double)i / (double)(m > n ? m : n));
  A[(long)i + (long)j * lda] *= s;
}

__global__ void shift_pivots_kernel(int to_lapack, int m, int* p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] += to_lapack ? (i + 1) : -(i + 1);
}

// dst(q, :) := src(perm[q], :), perm[q] being the source row of final row q
// of the interchanges i <-> i + p[i*incp] (fla_hip_dlaswp).
__global__ void gather_rows_kernel(int rows, int ncols, const int* __restrict__ perm,
                                   const double* __restrict__ src, long lds,
                                   double* __restrict__ dst, long ldd) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)rows * ncols;
  if (idx >= total) return;
  const int q = (int)(idx % rows), j = (int)(idx / rows);
  dst[(long)q + (long)j * ldd] = src[(long)perm[q] + (long)j * lds];
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
  FLA_TRY(fla_ensure_ready());
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(gen_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, seed, m, n, A, (long)lda, 0);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_gen_normal(unsigned long long seed, int m, int n, double* A, int lda) {
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

extern "C" int fla_hip_scale_rows_logspace(int m, int n, double* A, int lda) {
  FLA_TRY(fla_ensure_ready());
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, m, n, A, (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dmatmul_check(int kind, int uplo, int trans, int diag, int n, int nrhs,
                                     const double* M, int ldm, const double* X, int ldx,
                                     double beta, double* Y, int ldy) {
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
  FLA_TRY(fla_ensure_ready());
  if (m <= 0) return FLA_HIP_OK;
  if (ptype != kNativePivots && ptype != kLapackPivots) return FLA_HIP_EARG;
  hipLaunchKernelGGL(shift_pivots_kernel, dim3((m + 255) / 256), dim3(256), 0, fla_ctx().s_update,
                     ptype == kLapackPivots ? 1 : 0, m, p);
  FLA_HIP_CHECK(hipGetLastError());
  return fla_hip_synchronize();
}

extern "C" int fla_hip_dlaswp(int ncols, double* A, int lda, int k1, int k2, const int* p,
                              int incp) {
  FLA_TRY(fla_ensure_ready());
  if (ncols <= 0 || k2 < k1 || incp == 0) return FLA_HIP_OK;
  hipStream_t s = fla_ctx().s_update;
  // The interchanges touch rows up to max(i + p[i]); the permutation is built
  // over rows [0, lda) (every row index must lie inside the leading dimension).
  const int rows = lda;
  // The interchange sequence is inherently sequential: walk it on the host
  // (one small copy each way) rather than in a single GPU thread.
  const int pstride = incp > 0 ? incp : -incp;
  const long plen = (long)(k2 - k1) * pstride + 1;
  std::vector<int> hp((size_t)plen), perm((size_t)rows);
  FLA_HIP_CHECK(hipMemcpyAsync(hp.data(), p + (long)k1 * pstride, sizeof(int) * plen,
                               hipMemcpyDeviceToHost, s));
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  for (int q = 0; q < rows; ++q) perm[q] = q;
  const int step = incp > 0 ? 1 : -1;
  const int ib = incp > 0 ? k1 : k2;
  const int ie = incp > 0 ? k2 + 1 : k1 - 1;
  for (int i = ib; i != ie; i += step) {
    if (i < 0 || i >= rows) break;
    const int t = i + hp[(size_t)(i - k1) * pstride];
    if (t < 0 || t >= rows) break;  // pivot outside the matrix: leave the rest alone
    const int a = perm[i];
    perm[i] = perm[t];
    perm[t] = a;
  }
  void* w = nullptr;
  const size_t perm_bytes = ((size_t)rows * sizeof(int) + 255) & ~(size_t)255;
  FLA_TRY(fla_workspace(perm_bytes + (size_t)rows * ncols * sizeof(double), &w));
  int* dperm = static_cast<int*>(w);
  double* tmp = reinterpret_cast<double*>(static_cast<char*>(w) + perm_bytes);
  FLA_HIP_CHECK(hipMemcpyAsync(dperm, perm.data(), sizeof(int) * rows, hipMemcpyHostToDevice, s));
  // tmp := A (rows x ncols), then A(q, :) := tmp(perm[q], :)
  FLA_HIP_CHECK(hipMemcpy2DAsync(tmp, (size_t)rows * sizeof(double), A, (size_t)lda * sizeof(double),
                                 (size_t)rows * sizeof(double), (size_t)ncols,
                                 hipMemcpyDeviceToDevice, s));
  const long total = (long)rows * ncols;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     rows, ncols, dperm, tmp, (long)rows, A, (long)lda);
  FLA_HIP_CHECK(hipGetLastError());
  // perm (host) must outlive the upload
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  return FLA_HIP_OK;
}

// Symmetric positive definite A = Q diag(s) Q^T, s_i = kappa^(i/(n-1)),
// Q = H_1 ... H_nrefl with H_r = I - 2 v v^T / (v^T v), v ~ N(0,1) drawn by
// fla_hip_gen_normal(seed + r + 1, n, 1, ...).  Each reflector is applied as
// A := H A H = A - v z^T - z v^T.  Both triangles are written.
extern "C" int fla_hip_gen_spd_reflectors(unsigned long long seed, int n, int nrefl, double kappa,
                                          double* A, int lda) {
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
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1) || lds < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const long total = (long)m * n;
  hipLaunchKernelGGL(d2s_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     fla_ctx().s_update, m, n, A, (long)lda, S, (long)lds);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}
