// Left triangular solve with few right-hand sides (the solves of
// FLA_Chol_solve / FLA_LU_piv_solve / dpotrs_ / dgetrs_, dtrsm_ with side 'L'
// and nrhs <= 32): X := inv(op(T)) * X, one persistent launch per 16 columns.
//
// Workgroup I owns block row I (128 rows).  It keeps its rows of the
// right-hand side in MFMA accumulators, folds in -op(T)_IJ X_J for every block
// J it depends on as soon as block J is final, then multiplies by the inverted
// diagonal block (staged in LDS at launch), writes X_I and publishes it.
// Blocks finish in dependency order, so one monotonic counter says which are
// final (block s of the processing order is final when done > s).  Every
// wait is bounded by the 100 MHz wall clock; an expired wait sets *err and
// the kernel drains.
//
// Twelve waves per workgroup: eight compute waves own 16 rows each and hold
// the op(T)_IJ fragments of the current dependency and of the next one
// (loaded one step ahead, so the triangle streams from HBM while the chain
// waits); four loader waves poll the counter and stage X_J in LDS.  The
// roles matter because a wave's vector-memory counter completes in order:
// X_J loads issued behind a wave's in-flight triangle loads would wait for
// all of them.  The triangle is read from HBM once per solve.  Finished
// blocks are exchanged through an uncached copy Xm (rows x 16), so neither
// side needs an L2 write-back or invalidate across the XCDs: the writer waits
// for its stores to be acknowledged before bumping the counter.
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
constexpr int MLD = NB + 1;  // row stride of the staged triangle (substitution path)
constexpr int RHS = 16;
constexpr int NCOMP = 8;                 // compute waves (16 rows each)
constexpr int NTHREADS = (NCOMP + 4) * 64;  // + four loader waves
constexpr unsigned long long kWaitTicks = 200000000ULL;  // 2 s

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// op(T)(row, c0 + 32*lk + ks) for ks = 0..31 of this lane's row.
template <bool TRANS>
__device__ __forceinline__ void load_frag(double (&f)[32], const double* __restrict__ T, long ldt,
                                          int n, int row, int c0, int lk, bool full) {
  const int kb = c0 + 32 * lk;
  if (TRANS) {
    // the lane's 32 elements are contiguous: 16-byte loads when aligned
    const double* p = T + (long)kb + (long)row * ldt;
    if (full && ((reinterpret_cast<unsigned long>(p) & 15) == 0)) {
#pragma unroll
      for (int ks = 0; ks < 32; ks += 2) {
        const d2 v = *reinterpret_cast<const d2*>(p + ks);
        f[ks] = v[0];
        f[ks + 1] = v[1];
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < 32; ++ks) f[ks] = (row < n && kb + ks < n) ? p[ks] : 0.0;
    }
  } else {
#pragma unroll
    for (int ks = 0; ks < 32; ++ks)
      f[ks] = (full || (row < n && kb + ks < n)) ? T[(long)row + (long)(kb + ks) * ldt] : 0.0;
  }
}

template <bool FWD, bool TRANS>
__global__ __launch_bounds__(NTHREADS) void trsm_left_narrow_kernel(int n, int nrhs,
                                                                    const double* __restrict__ T,
                                                                    long ldt,
                                                                    const double* __restrict__ W,
                                                                    const int* __restrict__ ill,
                                                                    int unit, double* X, long ldx,
                                                                    double* __restrict__ Xm,
                                                                    int* done, int* err) {
  // Inverse path: op(inv(T_II)) in fragment order (NB*NB doubles).
  // Substitution path: M = op(T_II) row-major with ld MLD on its strict
  // triangle (lower when FWD), zero elsewhere, and the diagonal's reciprocals.
  __shared__ double Wf[NB * MLD];
  __shared__ double Xf[NB * RHS];  // X_J (negated) or Y_I, fragment order
  __shared__ double rdiag[NB];
  __shared__ int bail;
  const int nb = (n + NB - 1) / NB;
  const int I = FWD ? (int)blockIdx.x : nb - 1 - (int)blockIdx.x;
  const int r0 = I * NB;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const bool comp = w < NCOMP;
  const int lt = t - NCOMP * 64;  // loader thread index (0..255) when !comp
  const bool subst = ill != nullptr && ill[I] != 0;  // uniform per workgroup

  if (!subst) {
    // op(W_I)(row, k) -> Wf[((row/16)*32 + k%32)*64 + (k/32)*16 + row%16]
    // eight loads in flight per thread before the LDS writes
    const double* Wi = W + (long)I * NB * NB;
    for (int e0 = t; e0 < NB * NB; e0 += 8 * NTHREADS) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * NTHREADS;
        v[u] = e < NB * NB ? Wi[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * NTHREADS;
        if (e < NB * NB) {
          const int p = e & (NB - 1), q = e >> 7;  // stored W(p, q)
          const int row = TRANS ? q : p, k = TRANS ? p : q;
          Wf[((row >> 4) * 32 + (k & 31)) * 64 + (k >> 5) * 16 + (row & 15)] = v[u];
        }
      }
    }
  } else {
    // M(i, c) = op(T)(r0 + i, r0 + c) = TRANS ? T(r0 + c, r0 + i) : T(r0 + i, r0 + c)
    for (int e = t; e < NB * NB; e += NTHREADS) {
      const int p = e & (NB - 1), q = e >> 7;  // stored T(r0 + p, r0 + q), p contiguous
      const int i = TRANS ? q : p, c = TRANS ? p : q;
      const bool in = FWD ? (i > c) : (i < c);
      Wf[i * MLD + c] =
          (in && r0 + p < n && r0 + q < n) ? T[(long)(r0 + p) + (long)(r0 + q) * ldt] : 0.0;
    }
    for (int i = t; i < NB; i += NTHREADS)
      rdiag[i] = (unit || r0 + i >= n) ? 1.0 : 1.0 / T[(long)(r0 + i) * (ldt + 1)];
  }
  if (t == 0) bail = 0;

  // compute wave w: acc holds X(r0 + 16w + li, lk + 4r) in component r.
  const int row = r0 + 16 * w + li;
  d4 acc = d4{0.0, 0.0, 0.0, 0.0};
  if (comp) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = lk + 4 * r;
      acc[r] = (row < n && col < nrhs) ? X[(long)row + (long)col * ldx] : 0.0;
    }
  }

  const int ndep = FWD ? I : nb - 1 - I;
  auto dep_col = [&](int s) { return (FWD ? s : nb - 1 - s) * NB; };
  auto full_tile = [&](int c0) { return (r0 + NB <= n) && (c0 + NB <= n); };
  double fa[32];  // op(T)_IJ fragments of the next dependency

  // Inverse path with dependencies: the last dependency J (the block solved
  // just before this one) is folded through V = op(W_I) op(T)_IJ, formed
  // here at launch, so that after X_J arrives one 128 x 128 x 16 product
  // remains instead of two (X_I = op(W_I) Y_I - V X_J, where Y_I has every
  // other dependency folded in).  V(row, k) for this lane's row sits in
  // vf[4t + r] with k = 16t + 4r + lk: the MFMA result layout of V^T's
  // 16-column tile t, used directly as the B operand.  The fragment buffer
  // is single: vf takes the registers of a second one.
  const bool pre = !subst && ndep > 0;
  const int nfold = pre ? ndep - 1 : ndep;  // dependencies folded into Y_I
  double vf[32];
  if (pre) {
    // Tile tt of op(T)_IJ (rows r0.., columns cl + 16 tt ..) goes through Xf
    // in the X_J staging order: the loader waves fetch tile tt + 1 while the
    // compute waves multiply tile tt.
    const int cl = dep_col(ndep - 1);
    double tv[8];
    auto fetch = [&](int tt) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = lt + 256 * q;
        const int m = TRANS ? (e >> 4) : (e & (NB - 1)), i = TRANS ? (e & 15) : (e >> 7);
        const int r = r0 + m, c = cl + 16 * tt + i;
        tv[q] = (r < n && c < n) ? (TRANS ? T[(long)c + (long)r * ldt] : T[(long)r + (long)c * ldt])
                                 : 0.0;
      }
    };
    auto put = [&]() {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = lt + 256 * q;
        const int m = TRANS ? (e >> 4) : (e & (NB - 1)), i = TRANS ? (e & 15) : (e >> 7);
        Xf[(m & 31) * 64 + (m >> 5) * 16 + i] = tv[q];
      }
    };
    if (!comp) {
      fetch(0);
      put();
    }
    __syncthreads();  // Wf and tile 0 staged
#pragma unroll
    for (int tt = 0; tt < 8; ++tt) {
      if (!comp && tt + 1 < 8) fetch(tt + 1);
      if (comp) {
        // V^T tile: sum over m of op(T)(r0 + m, cl + 16 tt + col) op(W_I)(16 w + row, m)
        d4 v = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 32; ++ks)
          v = __builtin_amdgcn_mfma_f64_16x16x4f64(Xf[ks * 64 + lane],
                                                   Wf[(w * 32 + ks) * 64 + lane], v, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) vf[4 * tt + r] = v[r];
      }
      __syncthreads();
      if (tt + 1 < 8) {
        if (!comp) put();
        __syncthreads();
      }
    }
  }
  if (comp && nfold > 0)
    load_frag<TRANS>(fa, T, ldt, n, row, dep_col(0), lk, full_tile(dep_col(0)));

  // Wait until block s of the processing order is final; false when the wait failed.
  auto wait_final = [&](int s) -> bool {
    if (lt == 0) {
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
    return bail == 0;
  };
  // Loader waves: -X_J -> Xf, its row k placed as the contraction index of
  // MFMA step ks, lane group lk: k = 32 lk + ks against op(T) fragments,
  // k = 16 (ks / 4) + 4 (ks % 4) + lk against vf.
  // Thread lt stages row k = lt % 128 of columns lt / 128 + 2q, q = 0..7:
  // the eight loads are all issued before the first LDS write, so their
  // round trips to the uncached copy overlap.
  auto stage_dep = [&](int c0, bool vlayout) {
    if (!comp) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // Xm is uncached
      const int k = lt & (NB - 1), col0 = lt >> 7;
      const bool rin = c0 + k < n;
      // one per-lane base; the column steps 2 q n are uniform
      double* const xb = Xm + (long)(c0 + k) + (long)col0 * n;
      double v[NB * RHS / 256];
#pragma unroll
      for (int q = 0; q < NB * RHS / 256; ++q) {
        v[q] = (rin && col0 + 2 * q < nrhs)
                   ? __hip_atomic_load(xb + (long)(2 * q) * n, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT)
                   : 0.0;
      }
      const int slot = vlayout ? (4 * (k >> 4) + ((k >> 2) & 3)) * 64 + (k & 3) * 16
                               : (k & 31) * 64 + (k >> 5) * 16;
#pragma unroll
      for (int q = 0; q < NB * RHS / 256; ++q) Xf[slot + col0 + 2 * q] = -v[q];
    }
    __syncthreads();
  };

  // One dependency: wait for X_J, stage it, fold it in, then start loading
  // the next dependency's fragments (they arrive while the next wait runs).
  for (int s = 0; s < nfold; ++s) {
    if (!wait_final(s)) return;
    stage_dep(dep_col(s), false);
    if (comp) {
#pragma unroll
      for (int ks = 0; ks < 32; ++ks)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Xf[ks * 64 + lane], fa[ks], acc, 0, 0, 0);
      if (s + 1 < nfold) {
        const int c1 = dep_col(s + 1);
        load_frag<TRANS>(fa, T, ldt, n, row, c1, lk, full_tile(c1));
      }
    }
    __syncthreads();
  }

  d4 out = d4{0.0, 0.0, 0.0, 0.0};
  if (pre) {
    // out = op(W_I) Y_I while the loader polls for X_J, then out -= V X_J.
    if (comp) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Xf[(16 * (w & 1) + li) * 64 + (w >> 1) * 16 + lk + 4 * r] = acc[r];
    }
    __syncthreads();
    if (comp) {
#pragma unroll
      for (int ks = 0; ks < 32; ++ks)
        out = __builtin_amdgcn_mfma_f64_16x16x4f64(Xf[ks * 64 + lane],
                                                   Wf[(w * 32 + ks) * 64 + lane], out, 0, 0, 0);
    }
    if (!wait_final(ndep - 1)) return;  // its barrier also ends the reads of Xf
    stage_dep(dep_col(ndep - 1), true);
    if (comp) {
#pragma unroll
      for (int ks = 0; ks < 32; ++ks)
        out = __builtin_amdgcn_mfma_f64_16x16x4f64(Xf[ks * 64 + lane], vf[ks], out, 0, 0, 0);
    }
  } else if (!subst) {
    // Y_I -> Xf as the contraction operand (its row k = 16w + li is the
    // contraction index), then X_I = op(W_I) Y_I.
    if (comp) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Xf[(16 * (w & 1) + li) * 64 + (w >> 1) * 16 + lk + 4 * r] = acc[r];
    }
    __syncthreads();
    if (comp) {
#pragma unroll
      for (int ks = 0; ks < 32; ++ks)
        out = __builtin_amdgcn_mfma_f64_16x16x4f64(Xf[ks * 64 + lane],
                                                   Wf[(w * 32 + ks) * 64 + lane], out, 0, 0, 0);
    }
  } else {
    // Blocked substitution over the eight 16-row sub-blocks (wave s owns
    // sub-block s), in processing order: wave s substitutes through its
    // 16 x 16 diagonal block in registers (x_j broadcast from lane j of each
    // 16-lane group, dtrsm_'s column order), publishes X_s, and every later
    // wave folds -M_ws X_s into its rows on the MFMA.
    // The staged triangle and diagonal were written by every thread at
    // launch; a block without dependencies reaches this point with no
    // barrier in between.
    __syncthreads();
    for (int q = 0; q < NCOMP; ++q) {
      const int sb = FWD ? q : NCOMP - 1 - q;
      if (comp && w == sb) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
          const int j = FWD ? jj : 15 - jj;
          const double rd = rdiag[16 * sb + j];
          const double m = Wf[(16 * sb + li) * MLD + 16 * sb + j];
          const bool below = FWD ? (li > j) : (li < j);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double xj = __shfl(acc[r], j + 16 * lk, 64) * rd;
            if (li == j)
              acc[r] = xj;
            else if (below)
              acc[r] -= m * xj;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) Xf[li * 16 + lk + 4 * r] = -acc[r];
      }
      __syncthreads();
      if (comp && (FWD ? w > sb : w < sb)) {
        // acc(col, row) += sum_k (-X_s)(k, col) M(16w + row, 16 sb + k)
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
              Xf[(4 * q2 + lk) * 16 + li], Wf[(16 * w + li) * MLD + 16 * sb + 4 * q2 + lk], acc,
              0, 0, 0);
      }
      __syncthreads();
    }
    out = acc;
  }
  if (comp) {
    if (row < n) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = lk + 4 * r;
        if (col < nrhs) {
          X[(long)row + (long)col * ldx] = out[r];
          Xm[(long)row + (long)col * n] = out[r];
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // the Xm stores are acknowledged
  }
  __syncthreads();
  if (t == 0) __hip_atomic_fetch_add(done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void reset2_kernel(int* a, int* b) {
  *a = 0;
  *b = 0;
}

template <bool FWD, bool TRANS>
void launch_one(hipStream_t s, int nb, int na, int nr, const double* T, long ldt,
                const double* W, const int* ill, int unit, double* X, long ldx, double* Xm,
                int* done, int* err) {
  hipLaunchKernelGGL((trsm_left_narrow_kernel<FWD, TRANS>), dim3(nb), dim3(NTHREADS), 0, s, na, nr,
                     T, ldt, W, ill, unit, X, ldx, Xm, done, err);
}

}  // namespace

// X (na x nrhs, any nrhs) := inv(op(T)) X with W holding the inverted
// diagonal blocks (stored orientation, ld 128).  op(T) lower -> forward.
int launch_trsm_left_narrow(hipStream_t s, bool op_lower, bool trans, int na, int nrhs,
                            const double* T, long ldt, const double* W, double* X, long ldx,
                            int* done, int* err, const int* ill, bool unit) {
  const int nb = (na + NB - 1) / NB;
  if (na > kUcSolveRows) {
    std::fprintf(stderr, "libflame_amd: the narrow triangular solve supports n <= %ld\n",
                 kUcSolveRows);
    return FLA_HIP_EARG;
  }
  // the uncached block copy; launches on one stream run one after another
  double* Xm = reinterpret_cast<double*>(fla_ctx().uc + kUcLuBytes);
  const int u = unit ? 1 : 0;
  for (int c = 0; c < nrhs; c += RHS) {
    const int nr = (nrhs - c) < RHS ? (nrhs - c) : RHS;
    double* Xc = X + (long)c * ldx;
    hipLaunchKernelGGL(reset2_kernel, dim3(1), dim3(1), 0, s, done, err);
    if (op_lower && !trans)
      launch_one<true, false>(s, nb, na, nr, T, ldt, W, ill, u, Xc, ldx, Xm, done, err);
    else if (op_lower)
      launch_one<true, true>(s, nb, na, nr, T, ldt, W, ill, u, Xc, ldx, Xm, done, err);
    else if (!trans)
      launch_one<false, false>(s, nb, na, nr, T, ldt, W, ill, u, Xc, ldx, Xm, done, err);
    else
      launch_one<false, true>(s, nb, na, nr, T, ldt, W, ill, u, Xc, ldx, Xm, done, err);
    FLA_HIP_CHECK(hipGetLastError());
  }
  return FLA_HIP_OK;
}
