// Batched FP32 adjoints for many small problems (order n <= 256), one
// workgroup of 256 threads per problem:
//
//   fla_hip_schol_adjoint_batched   Abar_i = Phi-sym(L_i^-T Phi(L_i^T Lbar_i) L_i^-1)
//   fla_hip_sgp_rbf_lml_grad_batched
//       per problem: RBF kernel K_i (length scale l_i), batched FP32 Cholesky
//       and solve alpha_i = K_i^-1 y_i (fla_hip_sposv_batched), the log
//       marginal likelihood, Kbar_i from the Cholesky adjoint (Lbar =
//       -diag(1/L_jj), the log-determinant) plus the Cholesky-solve adjoint
//       (Xbar = -y/2, which adds alpha alpha^T / 2), and the gradient
//       (<Kbar, dK/dl>, <Kbar, dK/ds>).
//
// The adjoint is the one of fla_hip_dchol_adjoint (the reference has no
// adjoint; the formula is Murray's, "Differentiation of the Cholesky
// decomposition", 2016): P = Phi(L^T Lbar) (lower part, diagonal halved),
// Y = L^-T P and Z = Y L^-1 by blocked triangular solves, Abar = (Z + Z^T)/2.
// Each workgroup keeps one problem's intermediate matrices in global memory
// (L2-resident at these sizes) and stages 32 x 32 blocks of L through LDS:
//   step 1  P block (I, J) = sum_{K >= I} L_KI^T Lbar_KJ  (thread = row of the block, 4 columns)
//   step 2  Y, block row I from the bottom: thread = column of Y, 32 rows in
//           VGPRs, the contributions of the solved block rows K > I folded in
//           with L_KI broadcast from LDS, then the 32 x 32 back substitution
//   step 3  Z, block column J from the right: thread = row of Z, likewise
//   step 4  in-place symmetrization (thread r owns the pairs (r, j), j < r)
// Global accesses of steps 2 and 3 are coalesced where they read across the
// threads (step 3, P/Y/Z rows) or are one 128-byte line per thread (step 2).
#include "fla_hip_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr int TN = 256;  // threads per workgroup = maximum order
constexpr int TB = 32;   // block order
constexpr int SLD = TB + 1;

// S[k * SLD + r] = M(r0 + k, c0 + r) for the 32 x 32 block at (r0, c0) of the
// column-major n x n M (zero outside the matrix; lower part only when
// `lower`, i.e. zero where row < column).  When `diag` is given, M is the
// diagonal matrix diag(diag) instead.
__device__ __forceinline__ void load_block(float* S, const float* M, const float* diag, int n,
                                           int r0, int c0, bool lower, int t) {
#pragma unroll
  for (int u = 0; u < TB * TB / TN; ++u) {
    const int e = t + TN * u;
    const int k = e & (TB - 1), r = e >> 5;
    const int row = r0 + k, col = c0 + r;
    float v = 0.0f;
    if (row < n && col < n && (!lower || row >= col))
      v = diag ? (row == col ? diag[row] : 0.0f) : M[row + (long)col * n];
    S[k * SLD + r] = v;
  }
}

// Steps 1-4 for one problem: Abar := Phi-sym(L^-T Phi(L^T Lbar) L^-1).
// Lbar is the full lower matrix Lb, or diag(dLb) when Lb is null.
// sA, sB: 32 x 33 staging blocks; sX: 32 x 256 per-thread solve columns.
__device__ void chol_adjoint_wg(int n, const float* __restrict__ L, const float* __restrict__ Lb,
                                const float* __restrict__ dLb, float* __restrict__ Ab, float* sA,
                                float* sB, float* sX) {
  const int t = threadIdx.x;
  const int nb = (n + TB - 1) / TB;
  // ---- step 1: P = Phi(tril(L^T Lbar)) into Ab (zero above the diagonal) ----
  {
    const int r = t & (TB - 1), cq = t >> 5;  // row r, columns 4cq .. 4cq+3 of the block
    for (int J = 0; J < nb; ++J) {
      for (int I = 0; I < nb; ++I) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (I >= J) {
          for (int K = I; K < nb; ++K) {
            __syncthreads();
            load_block(sA, L, nullptr, n, K * TB, I * TB, true, t);
            load_block(sB, Lb, Lb ? nullptr : dLb, n, K * TB, J * TB, true, t);
            __syncthreads();
#pragma unroll 8
            for (int k = 0; k < TB; ++k) {
              const float a = sA[k * SLD + r];
#pragma unroll
              for (int q = 0; q < 4; ++q) acc[q] += a * sB[k * SLD + 4 * cq + q];
            }
          }
        }
        const int row = I * TB + r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int col = J * TB + 4 * cq + q;
          if (row >= n || col >= n) continue;
          float v = acc[q];
          if (row < col) v = 0.0f;
          if (row == col) v *= 0.5f;
          Ab[row + (long)col * n] = v;
        }
      }
    }
  }
  __syncthreads();
  // ---- step 2: Y = L^-T P in place (thread = column c) ----
  {
    const int c = t;
    for (int I = nb - 1; I >= 0; --I) {
      const int i0 = I * TB;
      float acc[TB];
#pragma unroll
      for (int r = 0; r < TB; ++r)
        acc[r] = (c < n && i0 + r < n) ? Ab[i0 + r + (long)c * n] : 0.0f;
      for (int K = I + 1; K < nb; ++K) {
        __syncthreads();
        load_block(sA, L, nullptr, n, K * TB, i0, false, t);  // sA[k][r] = L(K*32+k, i0+r)
        __syncthreads();
        const float* yc = Ab + K * TB + (long)c * n;
#pragma unroll 1
        for (int k = 0; k < TB; ++k) {
          const float yk = (c < n && K * TB + k < n) ? yc[k] : 0.0f;
#pragma unroll
          for (int r = 0; r < TB; ++r) acc[r] -= sA[k * SLD + r] * yk;
        }
      }
      __syncthreads();
      load_block(sA, L, nullptr, n, i0, i0, true, t);  // diagonal block
      __syncthreads();
      // L_II^T y = acc: y_r = (acc_r - sum_{k>r} L(i0+k, i0+r) y_k) / L(i0+r, i0+r),
      // on this thread's column of sX
      float* x = sX + t;
#pragma unroll
      for (int r = 0; r < TB; ++r) x[r * TN] = acc[r];
      for (int r = TB - 1; r >= 0; --r) {
        float v = x[r * TN];
        for (int k = r + 1; k < TB; ++k) v -= sA[k * SLD + r] * x[k * TN];
        x[r * TN] = (i0 + r < n) ? v / sA[r * SLD + r] : 0.0f;
      }
      if (c < n) {
#pragma unroll
        for (int r = 0; r < TB; ++r)
          if (i0 + r < n) Ab[i0 + r + (long)c * n] = x[r * TN];
      }
    }
  }
  __syncthreads();
  __threadfence_block();
  // ---- step 3: Z = Y L^-1 in place (thread = row r) ----
  {
    const int r = t;
    for (int J = nb - 1; J >= 0; --J) {
      const int j0 = J * TB;
      float acc[TB];
#pragma unroll
      for (int c = 0; c < TB; ++c)
        acc[c] = (r < n && j0 + c < n) ? Ab[r + (long)(j0 + c) * n] : 0.0f;
      for (int K = J + 1; K < nb; ++K) {
        __syncthreads();
        load_block(sA, L, nullptr, n, K * TB, j0, false, t);  // sA[k][c] = L(K*32+k, j0+c)
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < TB; ++k) {
          const float zk = (r < n && K * TB + k < n) ? Ab[r + (long)(K * TB + k) * n] : 0.0f;
#pragma unroll
          for (int c = 0; c < TB; ++c) acc[c] -= zk * sA[k * SLD + c];
        }
      }
      __syncthreads();
      load_block(sA, L, nullptr, n, j0, j0, true, t);
      __syncthreads();
      // z L_JJ = acc: z_c = (acc_c - sum_{k>c} z_k L(j0+k, j0+c)) / L(j0+c, j0+c)
      float* z = sX + t;
#pragma unroll
      for (int c = 0; c < TB; ++c) z[c * TN] = acc[c];
      for (int c = TB - 1; c >= 0; --c) {
        float v = z[c * TN];
        for (int k = c + 1; k < TB; ++k) v -= z[k * TN] * sA[k * SLD + c];
        z[c * TN] = (j0 + c < n) ? v / sA[c * SLD + c] : 0.0f;
      }
      if (r < n) {
#pragma unroll
        for (int c = 0; c < TB; ++c)
          if (j0 + c < n) Ab[r + (long)(j0 + c) * n] = z[c * TN];
      }
    }
  }
  __syncthreads();
  __threadfence_block();
  // ---- step 4: Abar = (Z + Z^T) / 2 ----
  if (t < n) {
    const int r = t;
    for (int j = 0; j < r; ++j) {
      const float v = 0.5f * (Ab[r + (long)j * n] + Ab[j + (long)r * n]);
      Ab[r + (long)j * n] = v;
      Ab[j + (long)r * n] = v;
    }
  }
  __syncthreads();
  __threadfence_block();
}

__global__ __launch_bounds__(TN) void chol_adjoint_batched_kernel(int n,
                                                                 const float* __restrict__ L,
                                                                 const float* __restrict__ Lb,
                                                                 float* __restrict__ Ab) {
  __shared__ float sA[TB * SLD], sB[TB * SLD];
  __shared__ float sX[TB * TN];
  const long off = (long)blockIdx.x * n * n;
  chol_adjoint_wg(n, L + off, Lb + off, nullptr, Ab + off, sA, sB, sX);
}

// K_i(r, c) = s^2 exp(-|x_r - x_c|^2 / (2 l_i^2)) + noise [r == c] (full n x n).
__global__ void gp_build_kernel(int n, int d, const float* __restrict__ x,
                                const float* __restrict__ l, float s, float noise,
                                float* __restrict__ K) {
  const int i = blockIdx.y;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)n * n) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const float* xi = x + (long)i * n * d;
  float d2 = 0.0f;
  for (int q = 0; q < d; ++q) {
    const float u = xi[r + (long)q * n] - xi[c + (long)q * n];
    d2 += u * u;
  }
  const float li = l[i];
  K[(long)i * n * n + e] = s * s * expf(-d2 / (2.0f * li * li)) + (r == c ? noise : 0.0f);
}

__device__ __forceinline__ float block_sum(float v, float* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int st = TN / 2; st > 0; st >>= 1) {
    if (t < st) red[t] += red[t + st];
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

// Per problem, after the batched Cholesky and solve: lml, Kbar (into Kb) and
// the gradient.
__global__ __launch_bounds__(TN) void gp_grad_batched_kernel(
    int n, int d, const float* __restrict__ x, const float* __restrict__ y,
    const float* __restrict__ l, float s, const float* __restrict__ Lall,
    const float* __restrict__ alpha_all, const int* __restrict__ info, float* __restrict__ Kball,
    float* __restrict__ lml, float* __restrict__ grad) {
  __shared__ float sA[TB * SLD], sB[TB * SLD];
  __shared__ float sX[TB * TN];
  __shared__ float red[TN];
  __shared__ float dLb[TN];
  __shared__ float xs[TN * 8];
  __shared__ float al[TN];
  const int i = blockIdx.x, t = threadIdx.x;
  if (info[i] != kSuccess) {
    if (t == 0) {
      lml[i] = __int_as_float(0x7fc00000);
      grad[2 * i] = grad[2 * i + 1] = __int_as_float(0x7fc00000);
    }
    return;
  }
  const float* L = Lall + (long)i * n * n;
  const float* alpha = alpha_all + (long)i * n;
  const float* yi = y + (long)i * n;
  const float* xi = x + (long)i * n * d;
  float* Kb = Kball + (long)i * n * n;
  // lml = -y'alpha/2 - sum log L_jj - n log(2 pi)/2; Lbar = -diag(1/L_jj)
  float fit = 0.0f, logdet = 0.0f;
  if (t < n) {
    const float ljj = L[t + (long)t * n];
    fit = yi[t] * alpha[t];
    logdet = logf(ljj);
    dLb[t] = -1.0f / ljj;
    al[t] = alpha[t];
    for (int q = 0; q < d; ++q) xs[t * 8 + q] = xi[t + (long)q * n];
  }
  fit = block_sum(fit, red);
  logdet = block_sum(logdet, red);
  if (t == 0) lml[i] = -0.5f * fit - logdet - 0.5f * n * 1.8378770664093453f;
  // Kbar = Phi-sym(L^-T Phi(L^T Lbar) L^-1) + alpha alpha^T / 2
  chol_adjoint_wg(n, L, nullptr, dLb, Kb, sA, sB, sX);
  const float li = l[i];
  float gl = 0.0f, gs = 0.0f;
  if (t < n) {
    const int r = t;
    for (int c = 0; c < n; ++c) {
      float d2 = 0.0f;
      for (int q = 0; q < d; ++q) {
        const float u = xs[r * 8 + q] - xs[c * 8 + q];
        d2 += u * u;
      }
      const float E = expf(-d2 / (2.0f * li * li));
      const float kb = Kb[r + (long)c * n] + 0.5f * al[r] * al[c];
      gl += kb * s * s * E * d2 / (li * li * li);
      gs += kb * 2.0f * s * E;
    }
  }
  gl = block_sum(gl, red);
  gs = block_sum(gs, red);
  if (t == 0) {
    grad[2 * i] = gl;
    grad[2 * i + 1] = gs;
  }
}

__global__ void copy_rows_kernel(int count, int n, const float* __restrict__ y,
                                 float* __restrict__ out) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < (long)count * n) out[e] = y[e];
}

}  // namespace

extern "C" int fla_hip_schol_adjoint_batched(int count, int n, const float* L, const float* Lbar,
                                             float* Abar) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (count < 0 || n < 0 || n > TN) return FLA_HIP_EARG;
  if (count == 0 || n == 0) return FLA_HIP_OK;
  if (!L || !Lbar || !Abar) return FLA_HIP_EARG;
  hipLaunchKernelGGL(chol_adjoint_batched_kernel, dim3(count), dim3(TN), 0, fla_ctx().s_update, n,
                     L, Lbar, Abar);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

extern "C" int fla_hip_sgp_rbf_lml_grad_batched(int count, int n, int d, const float* x,
                                                const float* y, const float* l, float s,
                                                float noise, float* lml, float* grad, int* info) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (count < 0 || n < 1 || n > TN || d < 1 || d > 8) return FLA_HIP_EARG;
  if (count == 0) return FLA_HIP_OK;
  if (!x || !y || !l || !lml || !grad || !info) return FLA_HIP_EARG;
  hipStream_t st = fla_ctx().s_update;
  const size_t nn = (size_t)n * n;
  // slot 3: K (becomes L) | Kbar | alpha
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(3, sizeof(float) * ((size_t)count * (2 * nn + n)) + 256, &w));
  float* K = static_cast<float*>(w);
  float* Kb = K + (size_t)count * nn;
  float* alpha = Kb + (size_t)count * nn;
  hipLaunchKernelGGL(gp_build_kernel, dim3((unsigned)((nn + 255) / 256), count), dim3(256), 0, st,
                     n, d, x, l, s, noise, K);
  hipLaunchKernelGGL(copy_rows_kernel, dim3((unsigned)(((size_t)count * n + 255) / 256)),
                     dim3(256), 0, st, count, n, y, alpha);
  FLA_HIP_CHECK(hipGetLastError());
  std::vector<int> nv(count, n), ld(count, n), one(count, 1);
  std::vector<float*> Ap(count), Bp(count);
  for (int i = 0; i < count; ++i) {
    Ap[i] = K + (size_t)i * nn;
    Bp[i] = alpha + (size_t)i * n;
  }
  FLA_TRY(fla_hip_sposv_batched(count, nv.data(), Ap.data(), ld.data(), Bp.data(), ld.data(),
                                one.data(), info));
  hipLaunchKernelGGL(gp_grad_batched_kernel, dim3(count), dim3(TN), 0, st, n, d, x, y, l, s, K,
                     alpha, info, Kb, lml, grad);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}
