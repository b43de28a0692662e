// Left triangular solve with few right-hand sides (the solves of
// FLA_Chol_solve / FLA_LU_piv_solve / dpotrs_ / dgetrs_, dtrsm_ with side 'L'
// and nrhs <= 32): X := inv(op(T)) * X, one persistent launch per 16 columns.
//
// Workgroup I owns block row I (128 rows).  It keeps its rows of the
// right-hand side in MFMA accumulators, folds in -op(T)_IJ X_J for every block
// J it depends on as soon as block J is final, then multiplies by the inverted
// diagonal block (staged in LDS at launch), writes X_I and publishes it.
// Blocks finish in dependency order, so one monotonic counter says which are
// final (block s of the processing order is final when done > s).  The T tile
// of the next dependency is loaded into registers before waiting, so the wait
// hides its HBM latency; the triangle is read from HBM once per solve.
// Every wait is bounded by the 100 MHz wall clock; an expired wait sets *err
// and the kernel drains.
//
// Products use v_mfma_f64_16x16x4f64 with the right-hand side as the MFMA A
// operand and the triangle as B, so D(col, row) lands in lane
// 16*(col%4) + row%16, register col/4.  Inside a 128-block the contraction
// index of step ks and lane group lk is k = 32*lk + ks; LDS copies of X and of
// the inverse are stored in that fragment order, so every fragment read is 64
// consecutive doubles.
#include "fla_hip_internal.h"

namespace {

constexpr int NB = 128;
constexpr int RHS = 16;
constexpr unsigned long long kWaitTicks = 200000000ULL;  // 2 s

typedef double d4 __attribute__((ext_vector_type(4)));

template <bool FWD, bool TRANS>
__global__ __launch_bounds__(256) void trsm_left_narrow_kernel(int n, int nrhs,
                                                               const double* __restrict__ T,
                                                               long ldt,
                                                               const double* __restrict__ W,
                                                               double* X, long ldx, int* done,
                                                               int* err) {
  __shared__ double Wf[NB * NB];   // op(inv(T_II)), fragment order
  __shared__ double Xf[NB * RHS];  // X_J (negated) or Y_I, fragment order
  __shared__ int bail;
  const int nb = (n + NB - 1) / NB;
  const int I = FWD ? (int)blockIdx.x : nb - 1 - (int)blockIdx.x;
  const int r0 = I * NB;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;

  // op(W_I)(row, k) -> Wf[((row/16)*32 + k%32)*64 + (k/32)*16 + row%16]
  const double* Wi = W + (long)I * NB * NB;
  for (int e = t; e < NB * NB; e += 256) {
    const int p = e & (NB - 1), q = e >> 7;  // stored W(p, q)
    const int row = TRANS ? q : p, k = TRANS ? p : q;
    Wf[((row >> 4) * 32 + (k & 31)) * 64 + (k >> 5) * 16 + (row & 15)] = Wi[e];
  }
  if (t == 0) bail = 0;

  // acc[a] holds X(r0 + 32w + 16a + li, lk + 4r) in component r.
  d4 acc[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int row = r0 + 32 * w + 16 * a + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = lk + 4 * r;
      acc[a][r] = (row < n && col < nrhs) ? X[(long)row + (long)col * ldx] : 0.0;
    }
  }

  const int ndep = FWD ? I : nb - 1 - I;
  for (int s = 0; s < ndep; ++s) {
    const int J = FWD ? s : nb - 1 - s;
    const int c0 = J * NB;
    // op(T)_IJ fragments: tf[a][ks] = op(T)(r0 + 32w + 16a + li, c0 + 32lk + ks).
    // They do not depend on other workgroups: load them before waiting.
    double tf[2][32];
    const bool full = (r0 + NB <= n) && (c0 + NB <= n);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int row = r0 + 32 * w + 16 * a + li;
#pragma unroll
      for (int ks = 0; ks < 32; ++ks) {
        const int k = c0 + 32 * lk + ks;
        const long off = TRANS ? (long)k + (long)row * ldt : (long)row + (long)k * ldt;
        tf[a][ks] = (full || (row < n && k < n)) ? T[off] : 0.0;
      }
    }
    if (t == 0) {
      const unsigned long long t0 = wall_clock64();
      while (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= s) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            wall_clock64() - t0 > kWaitTicks) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          bail = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    if (bail) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int e = t; e < NB * RHS; e += 256) {
      const int k = e & (NB - 1), col = e >> 7;
      const double v = (c0 + k < n && col < nrhs) ? X[(long)(c0 + k) + (long)col * ldx] : 0.0;
      Xf[(k & 31) * 64 + (k >> 5) * 16 + col] = -v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 32; ++ks) {
      const double xa = Xf[ks * 64 + lane];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, tf[a][ks], acc[a], 0, 0, 0);
    }
    __syncthreads();
  }

  // Y_I -> Xf as the contraction operand (its row is the contraction index).
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int r = 0; r < 4; ++r) Xf[(16 * a + li) * 64 + w * 16 + lk + 4 * r] = acc[a][r];
  __syncthreads();
  d4 out[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) out[a] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 32; ++ks) {
    const double ya = Xf[ks * 64 + lane];
#pragma unroll
    for (int a = 0; a < 2; ++a)
      out[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(ya, Wf[((2 * w + a) * 32 + ks) * 64 + lane],
                                                    out[a], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int row = r0 + 32 * w + 16 * a + li;
    if (row < n) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = lk + 4 * r;
        if (col < nrhs) X[(long)row + (long)col * ldx] = out[a][r];
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (t == 0) __hip_atomic_fetch_add(done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void reset2_kernel(int* a, int* b) {
  *a = 0;
  *b = 0;
}

template <bool FWD, bool TRANS>
void launch_one(hipStream_t s, int nb, int na, int nr, const double* T, long ldt,
                const double* W, double* X, long ldx, int* done, int* err) {
  hipLaunchKernelGGL((trsm_left_narrow_kernel<FWD, TRANS>), dim3(nb), dim3(256), 0, s, na, nr, T,
                     ldt, W, X, ldx, done, err);
}

}  // namespace

// X (na x nrhs, any nrhs) := inv(op(T)) X with W holding the inverted
// diagonal blocks (stored orientation, ld 128).  op(T) lower -> forward.
int launch_trsm_left_narrow(hipStream_t s, bool op_lower, bool trans, int na, int nrhs,
                            const double* T, long ldt, const double* W, double* X, long ldx,
                            int* done, int* err) {
  const int nb = (na + NB - 1) / NB;
  for (int c = 0; c < nrhs; c += RHS) {
    const int nr = (nrhs - c) < RHS ? (nrhs - c) : RHS;
    double* Xc = X + (long)c * ldx;
    hipLaunchKernelGGL(reset2_kernel, dim3(1), dim3(1), 0, s, done, err);
    if (op_lower && !trans)
      launch_one<true, false>(s, nb, na, nr, T, ldt, W, Xc, ldx, done, err);
    else if (op_lower)
      launch_one<true, true>(s, nb, na, nr, T, ldt, W, Xc, ldx, done, err);
    else if (!trans)
      launch_one<false, false>(s, nb, na, nr, T, ldt, W, Xc, ldx, done, err);
    else
      launch_one<false, true>(s, nb, na, nr, T, ldt, W, Xc, ldx, done, err);
    FLA_HIP_CHECK(hipGetLastError());
  }
  return FLA_HIP_OK;
}
