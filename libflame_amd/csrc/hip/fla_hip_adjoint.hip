// Reverse-mode adjoints of the Cholesky factorization and of the Cholesky
// solve (no counterpart in the reference; built from the same TRSM / GEMM
// kernels as FLA_Trsm / FLA_Gemm).
//
// For A = L L^T and a scalar function with gradient Lbar with respect to the
// lower factor, dL = L Phi(L^-1 dA L^-T) (Phi: lower triangle, diagonal
// halved) gives df = <Phi(L^T Lbar), L^-1 dA L^-T> = <L^-T P L^-1, dA> with
// P = Phi(L^T Lbar).  The returned Abar is the symmetric gradient
// (G + G^T)/2 with G = L^-T P L^-1: df = <Abar, dA> for every symmetric dA.
// Cost: the lower product L^T Lbar (n^3/3 multiply-adds on MFMA) and two
// n x n triangular solves.
//
// For X = A^-1 B: Bbar = A^-1 Xbar and the symmetric Abar contribution is
// -(Bbar X^T + X Bbar^T)/2.
#include "fla_hip_internal.h"

namespace {

// D := lower(S) with zero strictly upper part (the inputs' upper parts are
// not referenced).
__global__ void copy_lower_kernel(int n, const double* __restrict__ S, long lds,
                                  double* __restrict__ D, long ldd) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  D[(long)i + (long)j * ldd] = (i >= j) ? S[(long)i + (long)j * lds] : 0.0;
}

// P := Phi(M): strictly upper part zeroed, diagonal halved.
__global__ void phi_kernel(int n, double* __restrict__ M, long ldm) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  double* p = M + (long)i + (long)j * ldm;
  if (i < j)
    *p = 0.0;
  else if (i == j)
    *p *= 0.5;
}

// A := (A + A^T)/2 (32 x 32 tile pairs through LDS).
__global__ void symmetrize_kernel(int n, double* __restrict__ A, long lda) {
  __shared__ double tl[32][33], tu[32][33];
  const int T = (n + 31) / 32;
  int bj = 0, rem = blockIdx.x;
  while (rem >= T - bj) {
    rem -= T - bj;
    ++bj;
  }
  const int bi = bj + rem;  // tile (bi, bj), bi >= bj, and its mirror (bj, bi)
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + r;
    tl[r][tx] = (i < n && j < n) ? A[(long)i + (long)j * lda] : 0.0;
    const int i2 = bj * 32 + tx, j2 = bi * 32 + r;
    tu[r][tx] = (i2 < n && j2 < n) ? A[(long)i2 + (long)j2 * lda] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    // element (bi*32 + tx, bj*32 + r) pairs with (bj*32 + r, bi*32 + tx) = tu[tx][r]
    const int i = bi * 32 + tx, j = bj * 32 + r;
    const double v = 0.5 * (tl[r][tx] + tu[tx][r]);
    if (i < n && j < n) A[(long)i + (long)j * lda] = v;
    const int i2 = bj * 32 + tx, j2 = bi * 32 + r;
    const double v2 = 0.5 * (tu[r][tx] + tl[tx][r]);
    if (i2 < n && j2 < n) A[(long)i2 + (long)j2 * lda] = v2;
  }
}

inline unsigned nblk(long total) { return (unsigned)((total + 255) / 256); }

}  // namespace

extern "C" int fla_hip_dchol_adjoint(int n, const double* L, int ldl, const double* Lbar,
                                     int ldlb, double* Abar, int ldab) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n < 0 || ldl < (n > 1 ? n : 1) || ldlb < (n > 1 ? n : 1) || ldab < (n > 1 ? n : 1))
    return FLA_HIP_EARG;
  if (n == 0) return FLA_HIP_OK;
  FlaHipCtx& c = fla_ctx();
  hipStream_t s = c.s_update;
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(3, sizeof(double) * 2 * (size_t)n * n + 256, &w));
  double* Lc = static_cast<double*>(w);  // clean lower L
  double* Lb = Lc + (size_t)n * n;       // clean lower Lbar
  const long nn = (long)n * n;
  hipLaunchKernelGGL(copy_lower_kernel, dim3(nblk(nn)), dim3(256), 0, s, n, L, (long)ldl, Lc,
                     (long)n);
  hipLaunchKernelGGL(copy_lower_kernel, dim3(nblk(nn)), dim3(256), 0, s, n, Lbar, (long)ldlb, Lb,
                     (long)n);
  FLA_HIP_CHECK(hipGetLastError());
  // Abar (lower) := L^T Lbar ; only the lower triangle is needed
  FLA_TRY(launch_dgemm_mfma(s, true, false, 1, n, n, n, 1.0, Lc, n, Lb, n, 0.0, Abar, ldab,
                            nullptr));
  hipLaunchKernelGGL(phi_kernel, dim3(nblk(nn)), dim3(256), 0, s, n, Abar, (long)ldab);
  FLA_HIP_CHECK(hipGetLastError());
  // G = L^-T P L^-1
  FLA_TRY(fla_hip_dtrsm(kLeft, kLower, kTrans, kNonUnit, n, n, 1.0, Lc, n, Abar, ldab));
  FLA_TRY(fla_hip_dtrsm(kRight, kLower, kNoTrans, kNonUnit, n, n, 1.0, Lc, n, Abar, ldab));
  const int T = (n + 31) / 32;
  hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((long)T * (T + 1) / 2)), dim3(256), 0, s,
                     n, Abar, (long)ldab);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dchol_solve_adjoint(int n, int nrhs, const double* L, int ldl,
                                           const double* X, int ldx, const double* Xbar,
                                           int ldxb, double* Bbar, int ldbb, double* Abar,
                                           int ldab) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  const int m1 = n > 1 ? n : 1;
  if (n < 0 || nrhs < 0 || ldl < m1 || ldx < m1 || ldxb < m1 || ldbb < m1 || ldab < m1)
    return FLA_HIP_EARG;
  if (n == 0 || nrhs == 0) return FLA_HIP_OK;
  hipStream_t s = fla_ctx().s_update;
  FLA_HIP_CHECK(hipMemcpy2DAsync(Bbar, sizeof(double) * ldbb, Xbar, sizeof(double) * ldxb,
                                 sizeof(double) * n, nrhs, hipMemcpyDeviceToDevice, s));
  FLA_TRY(fla_hip_dtrsm(kLeft, kLower, kNoTrans, kNonUnit, n, nrhs, 1.0, L, ldl, Bbar, ldbb));
  FLA_TRY(fla_hip_dtrsm(kLeft, kLower, kTrans, kNonUnit, n, nrhs, 1.0, L, ldl, Bbar, ldbb));
  // Abar -= (Bbar X^T + X Bbar^T) / 2
  FLA_TRY(launch_dgemm_mfma(s, false, true, 0, n, n, nrhs, -0.5, Bbar, ldbb, X, ldx, 1.0, Abar,
                            ldab, nullptr));
  FLA_TRY(launch_dgemm_mfma(s, false, true, 0, n, n, nrhs, -0.5, X, ldx, Bbar, ldbb, 1.0, Abar,
                            ldab, nullptr));
  return FLA_HIP_OK;
}

// ---- Gaussian-process helpers (RBF kernel, log marginal likelihood) -------

namespace {

__device__ __forceinline__ double sqdist(const double* x, long ldx, int d, int i, int j) {
  double s = 0.0;
  for (int c = 0; c < d; ++c) {
    const double t = x[(long)i + (long)c * ldx] - x[(long)j + (long)c * ldx];
    s += t * t;
  }
  return s;
}

__global__ void rbf_kernel(int n, int d, const double* __restrict__ x, long ldx, double l, double s,
                           double noise, double* __restrict__ K, long ldk) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  double v = s * s * exp(-sqdist(x, ldx, d, i, j) / (2.0 * l * l));
  if (i == j) v += noise;
  K[(long)i + (long)j * ldk] = v;
}

// Per-block partial sums of Kbar .* dK/dl and Kbar .* dK/ds.
__global__ void rbf_grad_partial(int n, int d, const double* __restrict__ x, long ldx, double l,
                                 double s, const double* __restrict__ Kb, long ldkb,
                                 double* __restrict__ part) {
  __shared__ double r0[256], r1[256];
  double gl = 0.0, gs = 0.0;
  const long total = (long)n * n;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const double d2 = sqdist(x, ldx, d, i, j);
    const double e = exp(-d2 / (2.0 * l * l));
    const double kb = Kb[(long)i + (long)j * ldkb];
    gl += kb * s * s * e * d2 / (l * l * l);
    gs += kb * 2.0 * s * e;
  }
  r0[threadIdx.x] = gl;
  r1[threadIdx.x] = gs;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      r0[threadIdx.x] += r0[threadIdx.x + st];
      r1[threadIdx.x] += r1[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

// out[0] = sum_i part[2i], out[1] = sum_i part[2i+1] (fixed order).
__global__ void sum2_kernel(int np, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < np; ++i) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  out[0] = a;
  out[1] = b;
}

// out[0] = y . alpha, out[1] = sum_i log L_ii (one workgroup, fixed order).
__global__ void lml_terms_kernel(int n, const double* __restrict__ L, long ldl,
                                 const double* __restrict__ y, const double* __restrict__ alpha,
                                 double* __restrict__ out) {
  __shared__ double r0[256], r1[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    a += y[i] * alpha[i];
    b += log(L[(long)i + (long)i * ldl]);
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

__global__ void scale_vec_kernel(int n, double a, double* __restrict__ v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] *= a;
}

// Lbar := diag(-1 / L_ii) (full matrix, zero off the diagonal)
__global__ void logdet_bar_kernel(int n, const double* __restrict__ L, long ldl,
                                  double* __restrict__ Lb, long ldlb) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  Lb[(long)i + (long)j * ldlb] = (i == j) ? -1.0 / L[(long)i + (long)i * ldl] : 0.0;
}

}  // namespace

extern "C" int fla_hip_gp_rbf_kernel(int n, int d, const double* x, int ldx, double l, double s,
                                     double noise, double* K, int ldk) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n < 0 || d < 1 || ldx < (n > 1 ? n : 1) || ldk < (n > 1 ? n : 1) || !(l > 0.0))
    return FLA_HIP_EARG;
  if (n == 0) return FLA_HIP_OK;
  hipLaunchKernelGGL(rbf_kernel, dim3(nblk((long)n * n)), dim3(256), 0, fla_ctx().s_update, n, d,
                     x, (long)ldx, l, s, noise, K, (long)ldk);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_gp_rbf_grad(int n, int d, const double* x, int ldx, double l, double s,
                                   const double* Kbar, int ldkb, double* g) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n < 1 || d < 1 || ldx < n || ldkb < n || !(l > 0.0) || !g) return FLA_HIP_EARG;
  hipStream_t st = fla_ctx().s_update;
  const int np = 1024;
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(0, sizeof(double) * (2 * np + 2), &w));
  double* part = static_cast<double*>(w);
  hipLaunchKernelGGL(rbf_grad_partial, dim3(np), dim3(256), 0, st, n, d, x, (long)ldx, l, s, Kbar,
                     (long)ldkb, part);
  hipLaunchKernelGGL(sum2_kernel, dim3(1), dim3(64), 0, st, np, part, part + 2 * np);
  FLA_HIP_CHECK(hipGetLastError());
  FLA_HIP_CHECK(hipMemcpyAsync(g, part + 2 * np, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  FLA_HIP_CHECK(hipStreamSynchronize(st));
  return FLA_HIP_OK;
}

// Log marginal likelihood of y under N(0, K) and its gradient with respect
// to the RBF parameters (l, s), for K = s^2 exp(-|x-x'|^2/(2 l^2)) + noise*I:
//   lml = -y^T K^-1 y / 2 - sum log L_ii - n log(2 pi) / 2,
//   Kbar = (alpha alpha^T - K^-1) / 2 assembled from the Cholesky adjoint
//   (Lbar = -diag(1/L_ii)) and the Cholesky-solve adjoint (Xbar = -y/2),
//   grad = (<Kbar, dK/dl>, <Kbar, dK/ds>).
// Workspace: K, L, Lbar, Kbar (n x n each) are allocated per call.
extern "C" int fla_hip_gp_rbf_lml_grad(int n, int d, const double* x, int ldx, const double* y,
                                       double l, double s, double noise, double* lml,
                                       double* grad, int* info) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n < 1 || d < 1 || ldx < n || !lml || !grad || !info) return FLA_HIP_EARG;
  hipStream_t st = fla_ctx().s_update;
  const size_t nn = (size_t)n * n;
  double* buf = nullptr;
  FLA_TRY(fla_hip_malloc(reinterpret_cast<void**>(&buf), sizeof(double) * (3 * nn + 4 * (size_t)n + 8)));
  double* L = buf;
  double* Lb = L + nn;
  double* Kb = Lb + nn;
  double* alpha = Kb + nn;
  double* Bb = alpha + n;
  double* ybar = Bb + n;
  double* terms = ybar + n + n;
  int rc = FLA_HIP_OK;
  do {
    if ((rc = fla_hip_gp_rbf_kernel(n, d, x, ldx, l, s, noise, L, n)) != FLA_HIP_OK) break;
    if ((rc = fla_hip_dpotrf(kLower, n, L, n, info)) != FLA_HIP_OK) break;
    if (*info != kSuccess) break;
    // alpha = K^-1 y
    FLA_HIP_CHECK(hipMemcpyAsync(alpha, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if ((rc = fla_hip_dtrsm(kLeft, kLower, kNoTrans, kNonUnit, n, 1, 1.0, L, n, alpha, n)) != FLA_HIP_OK) break;
    if ((rc = fla_hip_dtrsm(kLeft, kLower, kTrans, kNonUnit, n, 1, 1.0, L, n, alpha, n)) != FLA_HIP_OK) break;
    hipLaunchKernelGGL(lml_terms_kernel, dim3(1), dim3(256), 0, st, n, L, (long)n, y, alpha, terms);
    // Kbar from the log-determinant: Cholesky adjoint with Lbar = -diag(1/L_ii)
    hipLaunchKernelGGL(logdet_bar_kernel, dim3(nblk((long)nn)), dim3(256), 0, st, n, L, (long)n, Lb,
                       (long)n);
    FLA_HIP_CHECK(hipGetLastError());
    if ((rc = fla_hip_dchol_adjoint(n, L, n, Lb, n, Kb, n)) != FLA_HIP_OK) break;
    // + the data-fit term: X = alpha, Xbar = -y/2
    FLA_HIP_CHECK(hipMemcpyAsync(ybar, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(scale_vec_kernel, dim3(nblk(n)), dim3(256), 0, st, n, -0.5, ybar);
    FLA_HIP_CHECK(hipGetLastError());
    if ((rc = fla_hip_dchol_solve_adjoint(n, 1, L, n, alpha, n, ybar, n, Bb, n, Kb, n)) != FLA_HIP_OK) break;
    double t[2] = {0.0, 0.0};
    FLA_HIP_CHECK(hipMemcpyAsync(t, terms, sizeof(t), hipMemcpyDeviceToHost, st));
    FLA_HIP_CHECK(hipStreamSynchronize(st));
    *lml = -0.5 * t[0] - t[1] - 0.5 * n * 1.8378770664093453;  // log(2 pi)
    if ((rc = fla_hip_gp_rbf_grad(n, d, x, ldx, l, s, Kb, n, grad)) != FLA_HIP_OK) break;
  } while (0);
  FLA_HIP_CHECK(hipStreamSynchronize(st));
  FLA_TRY(fla_hip_free(buf));
  return rc;
}
