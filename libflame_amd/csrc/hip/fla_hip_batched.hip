// Batched factorizations and solves of many small independent matrices of
// uneven sizes (n <= 512) in one persistent launch.
//
// Every workgroup pulls matrix indices from a device work counter (the batch
// is ordered by decreasing n, so the large matrices start first and the
// small ones fill the tail), factors its matrix alone and, when right-hand
// sides are given, solves with them before taking the next index.  The
// algorithms follow the right-looking blocked forms of FLA_Chol_l_blk_var3
// (src/lapack/dec/chol/l/flamec/FLA_Chol_l_blk_var3.c) and FLA_LU_piv_blk_var5
// (src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_blk_var5.c) with a panel of
// NB columns held in LDS:
//   Cholesky: unblocked right-looking factor of the m x NB panel in LDS
//     (dpotf2 order: sqrt, invscal, rank-1 update), then the trailing update
//     A22 -= L21 L21^T on MFMA (lower 32 x 32 tiles, read-modify-write).
//   LU: per column of the panel the pivot is the first maximal |a| of the
//     column (idamax_, src/flablas/f2c/idamax.c:40-46, NaN rule included),
//     rows are swapped across the whole matrix (FLAME-native relative
//     pivots), the column scaled by the reciprocal of the pivot and the
//     panel updated; then A12 := inv(L11) A12 and A22 -= L21 A12 on MFMA.
// The forward substitution of the solve runs inside the factorization,
// panel by panel, while L's panel is in LDS; the backward substitution reads
// the factor once more from memory, with the right-hand sides staged in LDS.
// Both solve 16 rows at a time: one wave sweeps the diagonal block with the
// right-hand sides in registers (lanes exchange values by shuffles, no
// barrier per column), then every thread updates the remaining rows.
//
// Element types: double (v_mfma_f64_16x16x4f64) and float
// (v_mfma_f32_16x16x4f32).  Info per matrix: FLA_SUCCESS, or the 0-based
// failing column (Cholesky) / first null pivot (LU, factorization continues
// as FLA_LU_piv does).
#include "fla_hip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

constexpr int MAXN = 512;

struct BatchItem {
  void* A;
  void* B;
  int* ipiv;
  long lda, ldb;
  int n, nrhs, index;
  int pad;
};

// A matrix pointer read from the item array as a global-address-space
// value: the accesses derived from it compile to global loads and stores.
// Read as a plain pointer it is generic, and flat accesses share the LDS
// counter, so every LDS read waits for the outstanding memory loads (the
// LU's interchanges, which read their slot lists from LDS between loads,
// ran twice as long that way).
typedef __attribute__((address_space(1))) void* global_void_ptr;
template <typename T>
__device__ __forceinline__ T* global_field(void* const* field) {
  const global_void_ptr g = *reinterpret_cast<const global_void_ptr*>(field);
  return (T*)g;
}

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// MFMA 16x16x4 with the column fragment as the A operand and the row
// fragment as B: lane (li = l%16, lk = l/16), register r holds
// C(row0 + li, col0 + dj(lk, r)).
__device__ __forceinline__ d4 mfma4(double colf, double rowf, d4 acc) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(colf, rowf, acc, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma4(float colf, float rowf, f4 acc) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(colf, rowf, acc, 0, 0, 0);
}
template <typename T>
struct Acc;
template <>
struct Acc<double> {
  typedef d4 type;
  static constexpr __device__ __host__ int dj(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Acc<float> {
  typedef f4 type;
  static constexpr __device__ __host__ int dj(int lk, int r) { return 4 * lk + r; }
};

__device__ __forceinline__ double fabs_t(double x) { return fabs(x); }
__device__ __forceinline__ float fabs_t(float x) { return fabsf(x); }

// idamax_ summary of a contiguous run of a column: the last NaN position and
// the first maximal |x| after it.
struct AmaxSum {
  double best;
  int bidx;  // -1: no element after the last NaN
  int nidx;  // -1: no NaN
};
__device__ __forceinline__ AmaxSum amax_combine(const AmaxSum& L, const AmaxSum& R) {
  if (R.nidx >= 0) return R;
  AmaxSum o = L;
  if (R.bidx >= 0 && (o.bidx < 0 || R.best > o.best)) {
    o.best = R.best;
    o.bidx = R.bidx;
  }
  return o;
}

// idamax_ summary of one element at position row.
__device__ __forceinline__ AmaxSum amax_of(double a, int row) {
  AmaxSum s{0.0, -1, -1};
  if (a != a)
    s.nidx = row;
  else {
    s.best = a;
    s.bidx = row;
  }
  return s;
}

// Combines the summaries of a wave's lanes in lane order (lane l holds a
// run of rows that precedes lane l + 1's); every lane gets the result.
__device__ __forceinline__ AmaxSum amax_wave(AmaxSum v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    AmaxSum o;
    o.best = __shfl_xor(v.best, off, 64);
    o.bidx = __shfl_xor(v.bidx, off, 64);
    o.nidx = __shfl_xor(v.nidx, off, 64);
    v = (lane & off) ? amax_combine(o, v) : amax_combine(v, o);
  }
  return v;
}

// Maximum of v over the wave by DPP moves (quad exchanges, row mirrors,
// row broadcasts), returned uniform.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, ROWMASK,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, ROWMASK,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_max_f64(double v) {
  v = fmax(v, dpp_f64<0xB1, 0xF>(v));   // quad_perm [1, 0, 3, 2]
  v = fmax(v, dpp_f64<0x4E, 0xF>(v));   // quad_perm [2, 3, 0, 1]
  v = fmax(v, dpp_f64<0x141, 0xF>(v));  // row_half_mirror
  v = fmax(v, dpp_f64<0x140, 0xF>(v));  // row_mirror
  v = fmax(v, dpp_f64<0x142, 0xA>(v));  // row_bcast:15 into rows 1 and 3
  v = fmax(v, dpp_f64<0x143, 0xC>(v));  // row_bcast:31 into rows 2 and 3
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// One wave's pivot candidates: its largest |a| and the first of its rows
// holding it among rows 0..255 (i1) and 256..511 (i2); nan if any of its
// candidates is a NaN.
struct PivRed {
  double best;
  int i1, i2, nan, pad;
};

// Trailing update C := C - Lr * Lc^T on MFMA, 32 x 32 tiles (lower tiles
// only when LOWER) of the m2 x n2 block C (ld ldc), K = kb <= NB.  Row
// operand: LDS panel rows (element (row, kk) at Pr[row + kk*MAXN]); column
// operand element (col, kk) at Pc[col*pcs + kk*pks] (the LDS panel for the
// Cholesky, U12 in memory for the LU: PC_GLOBAL).  Each wave walks its
// tiles with the next tile's C (and, from memory, its column fragments)
// loaded while the current tile's MFMAs run, so the global latency of one
// tile overlaps the work of the previous one.
template <typename T, bool LOWER, int NB, bool PC_GLOBAL>
__device__ void tile_update(int m2, int n2, int kb, const T* Pr, const T* Pc, long pcs, long pks,
                            T* C, long ldc, int t) {
  typedef typename Acc<T>::type AT;
  constexpr int NK = NB / 4;  // k-steps of 4
  const int w = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const int T1 = (m2 + 31) / 32, T2 = (n2 + 31) / 32;
  const int ntiles = LOWER ? T1 * (T1 + 1) / 2 : T1 * T2;
  auto coords = [&](int tile, int& ti, int& tj) {
    if (LOWER) {
      tj = 0;
      int rem = tile;
      while (rem >= T1 - tj) {
        rem -= T1 - tj;
        ++tj;
      }
      ti = tj + rem;
    } else {
      ti = tile % T1;
      tj = tile / T1;
    }
  };
  auto in_c = [&](int i, int j) { return i < m2 && j < n2 && (!LOWER || i >= j); };
  // Tiles wholly inside C (and, when LOWER, below the diagonal tiles) take
  // unpredicated accesses: element (a, b, r) of a lane sits at
  // cp[a*16 + (b*16 + r*DSTEP)*ldc] from the lane's base cp.
  constexpr int DSTEP = Acc<T>::dj(0, 1);
  auto interior = [&](int ti, int tj) {
    return (ti + 1) * 32 <= m2 && (tj + 1) * 32 <= n2 && (!LOWER || ti > tj);
  };
  auto cbase = [&](int ti, int tj) {
    return C + (ti * 32 + li) + ((long)tj * 32 + Acc<T>::dj(lk, 0)) * ldc;
  };
  // loads C and (PC_GLOBAL) the column fragments of one tile
  auto load = [&](int ti, int tj, T (&cv)[2][2][4], T (&pf)[NK][2], bool frags) {
    if (interior(ti, tj)) {
      const T* cp = cbase(ti, tj);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) cv[a][b][r] = cp[a * 16 + (long)(b * 16 + r * DSTEP) * ldc];
    } else {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = ti * 32 + a * 16 + li, j = tj * 32 + b * 16 + Acc<T>::dj(lk, r);
            cv[a][b][r] = in_c(i, j) ? C[i + (long)j * ldc] : (T)0;
          }
    }
    if (PC_GLOBAL && frags) {
      if (kb == NB && (tj + 1) * 32 <= n2) {
        const T* pp = Pc + (long)(tj * 32 + li) * pcs + (long)lk * pks;
#pragma unroll
        for (int q = 0; q < NK; ++q)
#pragma unroll
          for (int b = 0; b < 2; ++b) pf[q][b] = pp[(long)(b * 16) * pcs + (long)(4 * q) * pks];
      } else {
#pragma unroll
        for (int q = 0; q < NK; ++q) {
          const int kk = 4 * q + lk;
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const int col = tj * 32 + b * 16 + li;
            pf[q][b] = (col < n2 && kk < kb) ? Pc[(long)col * pcs + (long)kk * pks] : (T)0;
          }
        }
      }
    }
  };
  int tile = w;
  if (tile >= ntiles) return;
  int ti, tj;
  coords(tile, ti, tj);
  T cn[2][2][4], pn[NK][2];
  load(ti, tj, cn, pn, true);
  for (; tile < ntiles; tile += 4) {
    T cv[2][2][4], pf[NK][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[a][b][r] = cn[a][b][r];
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      pf[q][0] = pn[q][0];
      pf[q][1] = pn[q][1];
    }
    const int ci = ti, cj = tj;
    if (tile + 4 < ntiles) {
      coords(tile + 4, ti, tj);
      // a wave's next tile is mostly in the same tile column: its column
      // fragments (from memory) are then the ones already held
      load(ti, tj, cn, pn, tj != cj);
    }
    AT acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] = (T)0;
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      if (4 * q < kb) {
        const int kk = 4 * q + lk;
        T rf[2], cf[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int row = ci * 32 + a * 16 + li;
          rf[a] = (row < m2 && kk < kb) ? Pr[row + kk * MAXN] : (T)0;
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const int col = cj * 32 + b * 16 + li;
          cf[b] = PC_GLOBAL ? pf[q][b]
                            : ((col < n2 && kk < kb) ? Pc[(long)col * pcs + (long)kk * pks] : (T)0);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) acc[a][b] = mfma4(cf[b], rf[a], acc[a][b]);
      }
    }
    if (interior(ci, cj)) {
      T* cp = cbase(ci, cj);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            cp[a * 16 + (long)(b * 16 + r * DSTEP) * ldc] = cv[a][b][r] - acc[a][b][r];
    } else {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = ci * 32 + a * 16 + li;
            const int j = cj * 32 + b * 16 + Acc<T>::dj(lk, r);
            if (in_c(i, j)) C[i + (long)j * ldc] = cv[a][b][r] - acc[a][b][r];
          }
    }
  }
}

// One wave solves a kb x kb (kb <= 16) triangular block for up to 8
// right-hand sides held in registers: lane r + 16 g holds row r of the
// columns g and g + 4 (x[0], x[1]).  cf[j] is the coefficient of row r
// against column j of the block and dg the diagonal of row r.  FWD sweeps
// the columns in ascending order and updates the rows below each column;
// otherwise the columns go in descending order and the rows above are
// updated.  Per element the operations are the column sweep of dtrsm_
// (src/flablas/f2c/dtrsm.c): x_j := x_j / d_j, then x_i -= c_ij x_j.
template <typename T, bool FWD>
__device__ __forceinline__ void wave_trsv16(T (&x)[2], const T (&cf)[16], T dg, int r, int g,
                                            int kb) {
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int j = FWD ? s : 15 - s;
    if (j < kb) {
      if (r == j) {
        x[0] = x[0] / dg;
        x[1] = x[1] / dg;
      }
      const T y0 = __shfl(x[0], j + 16 * g, 64), y1 = __shfl(x[1], j + 16 * g, 64);
      if (FWD ? r > j : r < j) {
        x[0] -= cf[j] * y0;
        x[1] -= cf[j] * y1;
      }
    }
  }
}

// Copies the m x kb panel at Ak (ld lda) into P (ld MAXN).  Thread t owns
// rows t and t + 256; all its loads are issued before the LDS stores.
template <typename T, int NB>
__device__ __forceinline__ void load_panel(const T* Ak, long lda, int kb, int m, T* P, int t) {
  T v[2][NB];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int r = t + 256 * q;
      v[q][c] = (r < m && c < kb) ? Ak[r + (long)c * lda] : (T)0;
    }
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int r = t + 256 * q;
      if (r < m && c < kb) P[r + c * MAXN] = v[q][c];
    }
}

// Rows [i0, m) of B (relative to the panel; columns cg .. cg + ncg - 1,
// ncg <= 32): b_ic -= l_ij y_jc for j < sb, one term at a time in column
// order, with l_ij = P[i, p0 + j] and y_jc = Ys[j + (c - cg)*NY].  Each
// thread keeps its row's sb multipliers in registers and works on 8
// right-hand sides at a time.
template <typename T, int NY>
__device__ __forceinline__ void rows_update(const T* P, int p0, int sb, int i0, int m, T* B,
                                            long ldb, int cg, int ncg, const T* Ys, int t) {
  for (int i = i0 + t; i < m; i += 256) {
    T l[NY];
#pragma unroll
    for (int j = 0; j < NY; ++j) l[j] = j < sb ? P[i + (p0 + j) * MAXN] : (T)0;
    for (int c8 = 0; c8 < ncg; c8 += 8) {
      T b[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) b[q] = c8 + q < ncg ? B[i + (long)(cg + c8 + q) * ldb] : (T)0;
#pragma unroll
      for (int j = 0; j < NY; ++j)
        if (j < sb) {
#pragma unroll
          for (int q = 0; q < 8; ++q) b[q] -= l[j] * Ys[j + (c8 + q) * NY];
        }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (c8 + q < ncg) B[i + (long)(cg + c8 + q) * ldb] = b[q];
    }
  }
}

// Forward substitution with the factored Cholesky panel in LDS (element
// (i, j) at P[i + j*MAXN], m rows, kb columns): the panel's rows of B (B
// points at the panel's first row) are solved against the diagonal block
// and the rows below are updated, 16 columns at a time.  Each wave solves 8
// right-hand sides per pass; the update subtracts one term at a time in
// column order (dtrsm_'s order for B := inv(L) B).
template <typename T>
__device__ __forceinline__ void chol_panel_forward(const T* P, int kb, int m, T* B, long ldb,
                                                   int nrhs, T* Ys, int t) {
  const int w = t >> 6, lane = t & 63, r = lane & 15, g = lane >> 4;
  for (int s0 = 0; s0 < kb; s0 += 16) {
    const int sb = min(16, kb - s0), row = s0 + r;
    for (int cg = 0; cg < nrhs; cg += 32) {
      const int ncg = min(32, nrhs - cg);
      if (8 * w < ncg) {
        T cf[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) cf[j] = (r < sb && j < sb) ? P[row + (s0 + j) * MAXN] : (T)0;
        const T dg = r < sb ? P[row + row * MAXN] : (T)1;
        T x[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int cc = 8 * w + g + 4 * q;
          x[q] = (r < sb && cc < ncg) ? B[row + (long)(cg + cc) * ldb] : (T)0;
        }
        wave_trsv16<T, true>(x, cf, dg, r, g, sb);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int cc = 8 * w + g + 4 * q;
          if (r < sb && cc < ncg) {
            B[row + (long)(cg + cc) * ldb] = x[q];
            Ys[r + cc * 16] = x[q];
          }
        }
      }
      __syncthreads();
      rows_update<T, 16>(P, s0, sb, s0 + sb, m, B, ldb, cg, ncg, Ys, t);
      __syncthreads();
    }
  }
}

// Column J of the register-resident Cholesky panel (compile-time J keeps
// v in registers): the threads of rows < kb publish their column-J values,
// then each thread scales its rows' column J and updates its rows' columns
// J+1..min(r, kb-1).  jfail: the first column with a non-positive (or NaN)
// diagonal; the later columns are skipped.
template <typename T, int NB, int J>
__device__ __forceinline__ void chol_reg_column(T (&v)[2][NB], T (*colb)[NB], int kb, int m, int t,
                                                int& jfail) {
  if (J < kb && jfail < 0) {
    if (t < kb) colb[J & 1][t] = v[0][J];
    __syncthreads();
    const T d = colb[J & 1][J];
    if (!(d > (T)0)) {
      jfail = J;
      return;
    }
    const T sd = sqrt(d);
    const T inv = (T)1 / sd;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int r = t + 256 * q;
      if (r == J) {
        v[q][J] = sd;
      } else if (r > J && r < m) {
        const T lr = v[q][J] * inv;
        v[q][J] = lr;
#pragma unroll
        for (int c = J + 1; c < NB; ++c)
          if (c < kb && c <= r) v[q][c] -= lr * (colb[J & 1][c] * inv);
      }
    }
  }
}
template <typename T, int NB, int... J>
__device__ __forceinline__ void chol_reg_columns(T (&v)[2][NB], T (*colb)[NB], int kb, int m,
                                                 int t, int& jfail,
                                                 std::integer_sequence<int, J...>) {
  (chol_reg_column<T, NB, J>(v, colb, kb, m, t, jfail), ...);
}

// ---- Cholesky of one matrix (lower), fused forward substitution -----------
template <typename T, int NB>
__device__ int chol_one(T* A, long lda, int n, T* B, long ldb, int nrhs, T* P, T* Ys, int t) {
  __shared__ T colb[2][NB];  // the published column, by column parity
  int fail = -1;
  for (int k = 0; k < n; k += NB) {
    const int kb = min(NB, n - k), m = n - k;
    // The panel A[k:n, k:k+kb] in registers: thread t owns rows t and
    // t + 256.  Per column j, the threads of rows < kb publish their current
    // column-j values (double-buffered by column parity, one barrier); every
    // thread then scales its rows' column j (l = a * (1/d)) and updates its
    // rows' columns j+1..min(r, kb-1) with l * (a_cj * (1/d)), the
    // multipliers of the top rows formed from the published values.  These
    // are dpotf2's operations in its order, per element.
    T v[2][NB];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const int r = t + 256 * q;
        v[q][c] = (r < m && c < kb) ? A[(k + r) + (long)(k + c) * lda] : (T)0;
      }
    int jfail = -1;
    chol_reg_columns<T, NB>(v, colb, kb, m, t, jfail, std::make_integer_sequence<int, NB>{});
    if (jfail >= 0) fail = k + jfail;
    // the factored panel to LDS (trailing update, forward solve) and, lower
    // part, back to A
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        const int r = t + 256 * q;
        if (r < m && c < kb) {
          P[r + c * MAXN] = v[q][c];
          if (r >= c) A[(k + r) + (long)(k + c) * lda] = v[q][c];
        }
      }
    __syncthreads();
    if (fail >= 0) return fail;
    // forward substitution with this panel: Y_k = L_kk^-1 B_k, B_rest -= L21 Y_k
    if (nrhs > 0) chol_panel_forward<T>(P, kb, m, B + k, ldb, nrhs, Ys, t);
    // trailing update
    if (m > kb)
      tile_update<T, true, NB, false>(m - kb, m - kb, kb, P + kb, P + kb, 1, MAXN,
                           A + (k + kb) + (long)(k + kb) * lda, lda, t);
    __syncthreads();
  }
  return fail;
}

// Backward substitution X := L^-T Y (L lower in A) for up to 8 right-hand
// sides, staged in LDS (X, ld MAXN).  From the bottom, 16 rows at a time:
// one wave solves the diagonal block in registers, then every thread
// updates the rows above with the block's solution.
template <typename T>
__device__ void chol_back(const T* A, long lda, int n, T* B, long ldb, int nrhs, T* X, int t) {
  const int w = t >> 6, lane = t & 63, r = lane & 15, g = lane >> 4;
  for (int e = t; e < n * nrhs; e += 256) X[e % n + (e / n) * MAXN] = B[e % n + (long)(e / n) * ldb];
  __syncthreads();
  for (int k = ((n - 1) / 16) * 16; k >= 0; k -= 16) {
    const int kb = min(16, n - k);
    // the rows-above update's multipliers, loaded before the block solve
    T l[2][16];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = t + 256 * q;
#pragma unroll
      for (int j = 0; j < 16; ++j) l[q][j] = (i < k && j < kb) ? A[(k + j) + (long)i * lda] : (T)0;
    }
    if (w == 0) {
      // row r of the block of L^T: L[k + j, k + r]
      T cf[16];
#pragma unroll
      for (int j = 0; j < 16; ++j)
        cf[j] = (r < kb && j < kb) ? A[(k + j) + (long)(k + r) * lda] : (T)0;
      const T dg = r < kb ? A[(k + r) + (long)(k + r) * lda] : (T)1;
      T x[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c = g + 4 * q;
        x[q] = (r < kb && c < nrhs) ? X[(k + r) + c * MAXN] : (T)0;
      }
      wave_trsv16<T, false>(x, cf, dg, r, g, kb);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c = g + 4 * q;
        if (r < kb && c < nrhs) X[(k + r) + c * MAXN] = x[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = t + 256 * q;
      if (i < k) {
#pragma unroll 1
        for (int c = 0; c < nrhs; ++c) {
          T v = X[i + c * MAXN];
#pragma unroll
          for (int j = 0; j < 16; ++j)
            if (j < kb) v -= l[q][j] * X[(k + j) + c * MAXN];
          X[i + c * MAXN] = v;
        }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < n * nrhs; e += 256) B[e % n + (long)(e / n) * ldb] = X[e % n + (e / n) * MAXN];
  __syncthreads();
}

// ---- LU with partial pivoting of one square matrix, fused forward solve ----
template <typename T, int NB>
__device__ int lu_one(T* A, long lda, int n, int* ipiv, T* B, long ldb, int nrhs, T* P,
                      AmaxSum* red, int* spiv, T* Ys, int t) {
  __shared__ int prow[2 * NB], psrc[2 * NB], pcnt;
  __shared__ int allpiv[MAXN];
  __shared__ PivRed fred[4];
  int first_null = -1;
  for (int k = 0; k < n; k += NB) {
    const int kb = min(NB, n - k), m = n - k;
    load_panel<T, NB>(A + k + (long)k * lda, lda, kb, m, P, t);
    __syncthreads();
    // Thread t owns rows t and t + 256 of the panel.  Per column: the
    // owners' candidates are combined in row order (shuffles within a
    // wave, the four waves' two partial results through LDS: one barrier),
    // the pivot row is swapped in (second barrier), and each thread scales
    // and updates its own rows with the multipliers formed on the spot
    // (the same rounded products l = a * (1/piv) the stored column holds).
    // The next column's search reads only the thread's own rows, so no
    // third barrier is needed.
    const int lane = t & 63, w = t >> 6;
    for (int j = 0; j < kb; ++j) {
      // candidates of the thread's rows: |a|, or -1 outside rows [j, m)
      const double a1 = (t >= j && t < m) ? (double)fabs_t(P[t + j * MAXN]) : -1.0;
      const double a2 =
          (t + 256 >= j && t + 256 < m) ? (double)fabs_t(P[t + 256 + j * MAXN]) : -1.0;
      const bool wnan = __ballot(a1 != a1 || a2 != a2) != 0;
      const double mw = wave_max_f64(fmax(a1, a2));
      const unsigned long long b1 = __ballot(a1 == mw && a1 >= 0.0);
      const unsigned long long b2 = __ballot(a2 == mw && a2 >= 0.0);
      if (lane == 0) {
        fred[w].best = mw;
        fred[w].i1 = b1 ? w * 64 + __ffsll(b1) - 1 : -1;
        fred[w].i2 = b2 ? 256 + w * 64 + __ffsll(b2) - 1 : -1;
        fred[w].nan = wnan;
      }
      __syncthreads();
      int pr = -1;
      if (!(fred[0].nan | fred[1].nan | fred[2].nan | fred[3].nan)) {
        // no NaN: the first row, in row order, holding the largest |a|
        const double M = fmax(fmax(fred[0].best, fred[1].best), fmax(fred[2].best, fred[3].best));
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (pr < 0 && fred[q].best == M && fred[q].i1 >= 0) pr = fred[q].i1;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (pr < 0 && fred[q].best == M && fred[q].i2 >= 0) pr = fred[q].i2;
      } else {
        // idamax_'s NaN rule: the exact summaries, combined in row order
        const AmaxSum none{0.0, -1, -1};
        AmaxSum s1 = none, s2 = none;
        if (t >= j && t < m) s1 = amax_of(a1, t);
        if (t + 256 >= j && t + 256 < m) s2 = amax_of(a2, t + 256);
        s1 = amax_wave(s1, lane);
        s2 = amax_wave(s2, lane);
        if (lane == 0) {
          red[w] = s1;
          red[4 + w] = s2;
        }
        __syncthreads();
        AmaxSum tot = red[0];
#pragma unroll
        for (int q = 1; q < 8; ++q) tot = amax_combine(tot, red[q]);
        pr = tot.bidx >= 0 ? tot.bidx : tot.nidx;
      }
      if (t == 0) {
        spiv[j] = pr - j;  // relative pivot
        allpiv[k + j] = pr - j;
      }
      if (pr != j)
        for (int c = t; c < kb; c += 256) {
          const T a = P[j + c * MAXN], b = P[pr + c * MAXN];
          P[j + c * MAXN] = b;
          P[pr + c * MAXN] = a;
        }
      __syncthreads();
      const T piv = P[j + j * MAXN];
      const bool scale = piv != (T)0;
      const T inv = scale ? (T)1 / piv : (T)1;
      if (!scale && first_null < 0) first_null = k + j;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int r = t + 256 * q;
        if (r > j && r < m) {
          T l = P[r + j * MAXN];
          if (scale) {
            l *= inv;
            P[r + j * MAXN] = l;
          }
          for (int c = j + 1; c < kb; ++c) P[r + c * MAXN] -= l * P[j + c * MAXN];
        }
      }
    }
    __syncthreads();
    // panel back to A; pivots out
    for (int c = 0; c < kb; ++c)
      for (int r = t; r < m; r += 256) A[(k + r) + (long)(k + c) * lda] = P[r + c * MAXN];
    for (int j = t; j < kb; j += 256) ipiv[k + j] = spiv[j];
    // The kb interchanges compose into one permutation of at most 2 kb rows:
    // slot s moves row psrc[s] to row prow[s] (relative to k).  Wave 0
    // builds it with one lane per slot, applying the swaps in order.  Every
    // column right of the panel and every right-hand side then reads all its
    // moved rows before it writes any (no read-modify-write chain per swap).
    // The columns left of the panel get the interchanges of all later
    // panels at once, after the last panel.
    if (w == 0) {
      int row = lane < kb ? lane : -1, src = row, nt = kb;
      for (int j = 0; j < kb; ++j) {
        const int i2 = j + spiv[j];
        if (i2 != j) {
          const unsigned long long hit = __ballot(row == i2);
          int s2;
          if (hit) {
            s2 = __ffsll((unsigned long long)hit) - 1;
          } else {
            s2 = nt;
            if (lane == nt) {
              row = i2;
              src = i2;
            }
            ++nt;
          }
          const int sj = __shfl(src, j, 64), ss = __shfl(src, s2, 64);
          if (lane == j) src = ss;
          if (lane == s2) src = sj;
        }
      }
      if (lane < nt) {
        prow[lane] = row;
        psrc[lane] = src;
      }
      if (lane == 0) pcnt = nt;
    }
    __syncthreads();
    {
      const int nt = pcnt;
      const int nright = n - k - kb;
      for (int c = t; c < nright + nrhs; c += 256) {
        T* col = c < nright ? A + (long)(k + kb + c) * lda : B + (long)(c - nright) * ldb;
        T v[2 * NB];
#pragma unroll
        for (int q = 0; q < 2 * NB; ++q)
          if (q < nt && psrc[q] != prow[q]) v[q] = col[k + psrc[q]];
#pragma unroll
        for (int q = 0; q < 2 * NB; ++q)
          if (q < nt && psrc[q] != prow[q]) col[k + prow[q]] = v[q];
      }
    }
    __syncthreads();
    const int n2 = n - k - kb;
    if (n2 + nrhs > 0) {
      // A12 := inv(L11) A12 (unit lower), one thread per column; the panel's
      // rows of the right-hand sides are solved alongside as further columns
      for (int c = t; c < n2 + nrhs; c += 256) {
        T* col = c < n2 ? A + k + (long)(k + kb + c) * lda : B + k + (long)(c - n2) * ldb;
        T x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) x[i] = (i < kb) ? col[i] : (T)0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          // x[j] is final here; keeps later columns' LDS reads from being
          // hoisted (hundreds of live values would spill)
          asm volatile("" : "+v"(x[j])::"memory");
#pragma unroll
          for (int i = j + 1; i < NB; ++i)
            if (i < kb && j < kb) x[i] -= P[i + j * MAXN] * x[j];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i < kb) col[i] = x[i];
      }
      __threadfence_block();
      __syncthreads();
    }
    if (n2 > 0) {
      const T* U12 = A + k + (long)(k + kb) * lda;  // (kk, col) at U12[kk + col*lda]
      tile_update<T, false, NB, true>(m - kb, n2, kb, P + kb, U12, lda, 1,
                            A + (k + kb) + (long)(k + kb) * lda, lda, t);
    }
    // forward substitution below the panel: b_i -= l_ij y_j, one term at a
    // time in column order (dtrsm_'s order for B := inv(L) B), with the
    // panel's solved rows of B staged in LDS
    for (int cg = 0; cg < nrhs; cg += 32) {
      const int ncg = min(32, nrhs - cg);
      for (int cc = t / NB; cc < ncg; cc += 256 / NB) {
        const int j = t % NB;
        if (j < kb) Ys[j + cc * NB] = B[(k + j) + (long)(cg + cc) * ldb];
      }
      __syncthreads();
      rows_update<T, NB>(P, 0, kb, kb, m, B + k, ldb, cg, ncg, Ys, t);
      __syncthreads();
    }
    __syncthreads();
  }
  // The interchanges of later panels on the columns of earlier ones, once
  // per matrix.  From the last panel back to the first, cont[q] is the row
  // whose content ends up in row t + 256 q after the swaps of the panels
  // after the current one; those swaps touch only rows below the current
  // panel, so its columns gather rows there (all reads before any write).
  {
    int cont[2] = {t, t + 256};
    for (int k0 = ((n - 1) / NB) * NB; k0 >= 0; k0 -= NB) {
      const int kb = min(NB, n - k0);
      T v[2][NB];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int r = t + 256 * q;
        const bool mv = r >= k0 + kb && r < n && cont[q] != r;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          v[q][c] = (mv && c < kb) ? A[cont[q] + (long)(k0 + c) * lda] : (T)0;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int r = t + 256 * q;
        const bool mv = r >= k0 + kb && r < n && cont[q] != r;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (mv && c < kb) A[r + (long)(k0 + c) * lda] = v[q][c];
      }
      // cont := the swaps of this panel, last to first, applied to cont
      for (int j = kb - 1; j >= 0; --j) {
        const int a = k0 + j, b = a + allpiv[a];
#pragma unroll
        for (int q = 0; q < 2; ++q) cont[q] = cont[q] == a ? b : (cont[q] == b ? a : cont[q]);
      }
    }
    __syncthreads();
  }
  return first_null;
}

// Backward substitution X := U^-1 Y (U upper in A) for up to 8 right-hand
// sides, staged in LDS (X, ld MAXN).  From the bottom, 16 rows at a time:
// one wave solves the diagonal block in registers, then every thread
// updates the rows above, one term at a time in descending column order.
// Per element this is dtrsm_'s column sweep for B := inv(U) B.
template <typename T>
__device__ void lu_back(const BatchItem* itp, int c0, int nrhs,
                        __attribute__((address_space(3))) T* X, int t) {
  // out of line (inlined, it pushes the LU kernel past its register
  // budget), so it reads its pointers from the item as global values and
  // takes X as an LDS pointer
  const T* A = global_field<T>(&itp->A);
  T* B = global_field<T>(&itp->B) + (long)c0 * itp->ldb;
  const long lda = itp->lda, ldb = itp->ldb;
  const int n = itp->n;
  const int w = t >> 6, lane = t & 63, r = lane & 15, g = lane >> 4;
  for (int e = t; e < n * nrhs; e += 256) X[e % n + (e / n) * MAXN] = B[e % n + (long)(e / n) * ldb];
  __syncthreads();
  for (int k = ((n - 1) / 16) * 16; k >= 0; k -= 16) {
    const int kb = min(16, n - k);
    // the rows-above update's coefficients, loaded before the block solve
    T u[2][16];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = t + 256 * q;
#pragma unroll
      for (int j = 0; j < 16; ++j) u[q][j] = (i < k && j < kb) ? A[i + (long)(k + j) * lda] : (T)0;
    }
    if (w == 0) {
      T cf[16];
#pragma unroll
      for (int j = 0; j < 16; ++j)
        cf[j] = (r < kb && j < kb) ? A[(k + r) + (long)(k + j) * lda] : (T)0;
      const T dg = r < kb ? A[(k + r) + (long)(k + r) * lda] : (T)1;
      T x[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c = g + 4 * q;
        x[q] = (r < kb && c < nrhs) ? X[(k + r) + c * MAXN] : (T)0;
      }
      wave_trsv16<T, false>(x, cf, dg, r, g, kb);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int c = g + 4 * q;
        if (r < kb && c < nrhs) X[(k + r) + c * MAXN] = x[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = t + 256 * q;
      if (i < k) {
#pragma unroll 1
        for (int c = 0; c < nrhs; ++c) {
          T v = X[i + c * MAXN];
#pragma unroll
          for (int j = 15; j >= 0; --j)
            if (j < kb) v -= u[q][j] * X[(k + j) + c * MAXN];
          X[i + c * MAXN] = v;
        }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < n * nrhs; e += 256) B[e % n + (long)(e / n) * ldb] = X[e % n + (e / n) * MAXN];
  __syncthreads();
}

// Persistent driver: OP 0 Cholesky (lower), 1 LU with partial pivoting.
// As many workgroups share a CU (up to 4) as the panel's LDS allows; the
// register budget is held to that occupancy.
template <typename T, int OP, int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    (160 * 1024) / ((MAXN + 32) * NB * sizeof(T) + 1024) >= 4
        ? 4
        : ((160 * 1024) / ((MAXN + 32) * NB * sizeof(T) + 1024) >= 2 ? 2 : 1)))) void batched_kernel(const BatchItem* __restrict__ items,
                                                      int count, int* counter, int* info,
                                                      int max_rhs) {
  __shared__ T P[MAXN * NB];
  __shared__ T Ys[NB * 32];  // solved rows of the right-hand sides, 32 columns
  __shared__ AmaxSum red[8];
  __shared__ int spiv[NB];
  __shared__ int cur;
  const int t = threadIdx.x;
  (void)max_rhs;
  for (;;) {
    if (t == 0) cur = atomicAdd(counter, 1);
    __syncthreads();
    const int idx = cur;
    __syncthreads();
    if (idx >= count) break;
    const BatchItem it = items[idx];
    // global pointers for the LU; the Cholesky kernels measured faster with
    // the generic ones
    T* A = OP == 1 ? global_field<T>(&items[idx].A) : static_cast<T*>(it.A);
    T* B = OP == 1 ? global_field<T>(&items[idx].B) : static_cast<T*>(it.B);
    const int nr = it.B ? it.nrhs : 0;
    int r;
    if (OP == 0) {
      r = chol_one<T, NB>(A, it.lda, it.n, B, it.ldb, nr, P, Ys, t);
      __syncthreads();
      if (r < 0 && nr > 0) {
        for (int c0 = 0; c0 < nr; c0 += 8)
          chol_back<T>(A, it.lda, it.n, B + (long)c0 * it.ldb, it.ldb, min(8, nr - c0), P, t);
      }
    } else {
      r = lu_one<T, NB>(A, it.lda, it.n, global_field<int>((void* const*)&items[idx].ipiv), B,
                        it.ldb, nr, P, red, spiv, Ys, t);
      __syncthreads();
      if (r < 0 && nr > 0) {
        for (int c0 = 0; c0 < nr; c0 += 8)
          lu_back<T>(&items[idx], c0, min(8, nr - c0), (__attribute__((address_space(3))) T*)P, t);
      }
    }
    if (t == 0) info[it.index] = r < 0 ? kSuccess : r;
    __syncthreads();
  }
}

__global__ void reset_counter_kernel(int* c) { *c = 0; }

// ---- host side: schedule, configurations, autotuner, graph replay ----------

struct BatchConfig {
  int nb = 32;        // panel width
  int wg_per_cu = 1;  // persistent workgroups per CU
};

enum { kDpotrf = 0, kSpotrf = 1, kDgetrf = 2, kKinds = 3 };

// Size signature of a batch: counts per 64-wide size bucket (n <= 512) and
// the presence of right-hand sides.
typedef std::vector<int> Signature;

struct TunedEntry {
  Signature sig;
  BatchConfig cfg;
};
std::vector<TunedEntry> g_tuned[kKinds];
bool g_use_graph = true;

struct GraphCache {
  bool valid = false;
  int kind = -1, count = 0;
  BatchConfig cfg;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
};
GraphCache g_graph;

// Device and pinned staging of the items (grow-only).
BatchItem* g_items_dev = nullptr;
BatchItem* g_items_pin = nullptr;
int* g_counter = nullptr;
size_t g_items_cap = 0;

Signature signature_of(int count, const int* n, const int* nrhs) {
  Signature s(10, 0);
  for (int i = 0; i < count; ++i) s[std::min(8, n[i] / 64)]++;
  bool rhs = false;
  for (int i = 0; nrhs && i < count; ++i) rhs |= nrhs[i] > 0;
  s[9] = rhs ? 1 : 0;
  return s;
}

// The fastest configurations measured on 4096 matrices of n ~ U{32..512}
// with 1-8 right-hand sides.
BatchConfig default_config(int kind) {
  BatchConfig c;
  c.nb = 16;
  c.wg_per_cu = (kind == kSpotrf) ? 4 : 2;
  return c;
}

// A tuned entry for this exact signature, else a forced configuration
// (empty signature, fla_hip_batched_set_config), else the default.
BatchConfig lookup_config(int kind, const Signature& sig) {
  for (const TunedEntry& e : g_tuned[kind])
    if (e.sig == sig) return e.cfg;
  for (const TunedEntry& e : g_tuned[kind])
    if (e.sig.empty()) return e.cfg;
  return default_config(kind);
}

template <typename T, int OP, int NB>
int launch_batched(hipStream_t s, int grid, const BatchItem* items, int count, int* counter,
                   int* info) {
  hipLaunchKernelGGL((batched_kernel<T, OP, NB>), dim3(grid), dim3(256), 0, s, items, count,
                     counter, info, 0);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

int launch_kind(hipStream_t s, int kind, const BatchConfig& cfg, int count, int* info) {
  const int grid = std::max(1, std::min(count, fla_ctx().num_cus * cfg.wg_per_cu));
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, g_counter);
  FLA_HIP_CHECK(hipGetLastError());
  if (kind == kDpotrf) {
    if (cfg.nb == 16) return launch_batched<double, 0, 16>(s, grid, g_items_dev, count, g_counter, info);
    return launch_batched<double, 0, 32>(s, grid, g_items_dev, count, g_counter, info);
  }
  if (kind == kSpotrf) {
    if (cfg.nb == 16) return launch_batched<float, 0, 16>(s, grid, g_items_dev, count, g_counter, info);
    return launch_batched<float, 0, 32>(s, grid, g_items_dev, count, g_counter, info);
  }
  return launch_batched<double, 1, 16>(s, grid, g_items_dev, count, g_counter, info);
}

// Panel widths: 16 or 32 for the Cholesky kinds, 16 for the LU (wider LU
// panels spill registers and measured slower).
bool valid_nb(int kind, int nb) {
  if (kind == kDgetrf) return nb == 16;
  return nb == 16 || nb == 32;
}

// Workgroups per CU that the panel's LDS allows (at most 4).
int max_wg_per_cu(int kind, int nb) {
  const int bytes = (MAXN + 32) * nb * (kind == kSpotrf ? 4 : 8) + 1024;
  return std::max(1, std::min(4, (160 * 1024) / bytes));
}

// Validates the arguments and stages the items, largest n first.
int stage_items(int count, const int* n, void* const* A, const int* lda, int* const* ipiv,
                void* const* B, const int* ldb, const int* nrhs, bool need_ipiv) {
  if (count < 0 || (count > 0 && (!n || !A || !lda))) return FLA_HIP_EARG;
  for (int i = 0; i < count; ++i) {
    if (n[i] < 0 || n[i] > MAXN || lda[i] < std::max(1, n[i]) || (n[i] > 0 && !A[i]))
      return FLA_HIP_EARG;
    if (need_ipiv && n[i] > 0 && (!ipiv || !ipiv[i])) return FLA_HIP_EARG;
    if (B && nrhs && nrhs[i] > 0) {
      if (!B[i] || !ldb || ldb[i] < std::max(1, n[i]) || nrhs[i] > 256) return FLA_HIP_EARG;
    }
  }
  if ((size_t)count > g_items_cap) {
    FLA_HIP_CHECK(hipDeviceSynchronize());
    if (g_items_dev) FLA_HIP_CHECK(hipFree(g_items_dev));
    if (g_items_pin) FLA_HIP_CHECK(hipHostFree(g_items_pin));
    g_items_dev = nullptr;
    g_items_pin = nullptr;
    g_graph.valid = false;
    const size_t cap = std::max<size_t>(1024, count);
    FLA_HIP_CHECK(hipMalloc(&g_items_dev, cap * sizeof(BatchItem)));
    FLA_HIP_CHECK(hipHostMalloc(&g_items_pin, cap * sizeof(BatchItem), hipHostMallocDefault));
    g_items_cap = cap;
  }
  if (!g_counter) FLA_HIP_CHECK(hipMalloc(&g_counter, 256));
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
  // the previous launch may still read the pinned items
  FLA_HIP_CHECK(hipStreamSynchronize(fla_ctx().s_update));
  for (int q = 0; q < count; ++q) {
    const int i = order[q];
    BatchItem it;
    std::memset(&it, 0, sizeof(it));
    it.A = A[i];
    it.lda = lda[i];
    it.n = n[i];
    it.index = i;
    it.ipiv = need_ipiv ? ipiv[i] : nullptr;
    const int r = (B && nrhs) ? nrhs[i] : 0;
    it.nrhs = r;
    it.B = r > 0 ? B[i] : nullptr;
    it.ldb = r > 0 ? ldb[i] : 1;
    g_items_pin[q] = it;
  }
  return FLA_HIP_OK;
}

// Copies the staged items and runs the persistent kernel, through a cached
// hipGraph (memcpy node, counter reset kernel, kernel) when enabled.
int run_staged(int kind, int count, const BatchConfig& cfg, int* info) {
  hipStream_t s = fla_ctx().s_update;
  if (count == 0) return FLA_HIP_OK;
  if (!g_use_graph) {
    FLA_HIP_CHECK(hipMemcpyAsync(g_items_dev, g_items_pin, sizeof(BatchItem) * count,
                                 hipMemcpyHostToDevice, s));
    return launch_kind(s, kind, cfg, count, info);
  }
  const bool hit = g_graph.valid && g_graph.kind == kind && g_graph.count == count &&
                   g_graph.cfg.nb == cfg.nb && g_graph.cfg.wg_per_cu == cfg.wg_per_cu;
  static int* g_graph_info = nullptr;
  if (!hit || g_graph_info != info) {
    if (g_graph.exec) FLA_HIP_CHECK(hipGraphExecDestroy(g_graph.exec));
    if (g_graph.graph) FLA_HIP_CHECK(hipGraphDestroy(g_graph.graph));
    g_graph.exec = nullptr;
    g_graph.graph = nullptr;
    g_graph.valid = false;
    hipStream_t cs;
    FLA_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    FLA_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int st = FLA_HIP_OK;
    if (hipMemcpyAsync(g_items_dev, g_items_pin, sizeof(BatchItem) * count,
                       hipMemcpyHostToDevice, cs) != hipSuccess)
      st = FLA_HIP_EARG;
    if (st == FLA_HIP_OK) st = launch_kind(cs, kind, cfg, count, info);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    FLA_HIP_CHECK(hipStreamDestroy(cs));
    if (st != FLA_HIP_OK) return st;
    FLA_HIP_CHECK(ec);
    g_graph.graph = g;
    FLA_HIP_CHECK(hipGraphInstantiate(&g_graph.exec, g, nullptr, nullptr, 0));
    g_graph.valid = true;
    g_graph.kind = kind;
    g_graph.count = count;
    g_graph.cfg = cfg;
    g_graph_info = info;
  }
  FLA_HIP_CHECK(hipGraphLaunch(g_graph.exec, s));
  return FLA_HIP_OK;
}

int batched_entry(int kind, int count, const int* n, void* const* A, const int* lda,
                  int* const* ipiv, void* const* B, const int* ldb, const int* nrhs, int* info) {
  FLA_TRY(fla_ensure_ready());
  if (count > 0 && !info) return FLA_HIP_EARG;
  FLA_TRY(stage_items(count, n, A, lda, ipiv, B, ldb, nrhs, kind == kDgetrf));
  const BatchConfig cfg = lookup_config(kind, signature_of(count, n, B ? nrhs : nullptr));
  return run_staged(kind, count, cfg, info);
}

// ---- autotuner ----------------------------------------------------------------

__global__ void tune_fill_kernel(int n, double* A, float* Af, int spd, unsigned long long seed) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long)n * n) return;
  const int i = (int)(idx % n), j = (int)(idx / n);
  unsigned long long z = seed + (unsigned long long)(spd ? (i > j ? i * 1315423911ULL + j : j * 1315423911ULL + i) : idx);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  double v = (double)(z >> 11) * 0x1.0p-53 - 0.5;
  if (spd && i == j) v = (double)n;
  if (A) A[idx] = v;
  if (Af) Af[idx] = (float)v;
}

}  // namespace

extern "C" int fla_hip_dposv_batched(int count, const int* n, double* const* A, const int* lda,
                                     double* const* B, const int* ldb, const int* nrhs,
                                     int* info) {
  FLA_USER_STREAM_SCOPE;
  return batched_entry(kDpotrf, count, n, reinterpret_cast<void* const*>(A), lda, nullptr,
                       reinterpret_cast<void* const*>(B), ldb, nrhs, info);
}

extern "C" int fla_hip_sposv_batched(int count, const int* n, float* const* A, const int* lda,
                                     float* const* B, const int* ldb, const int* nrhs, int* info) {
  FLA_USER_STREAM_SCOPE;
  return batched_entry(kSpotrf, count, n, reinterpret_cast<void* const*>(A), lda, nullptr,
                       reinterpret_cast<void* const*>(B), ldb, nrhs, info);
}

extern "C" int fla_hip_dgesv_batched(int count, const int* n, double* const* A, const int* lda,
                                     int* const* ipiv, double* const* B, const int* ldb,
                                     const int* nrhs, int* info) {
  FLA_USER_STREAM_SCOPE;
  return batched_entry(kDgetrf, count, n, reinterpret_cast<void* const*>(A), lda, ipiv,
                       reinterpret_cast<void* const*>(B), ldb, nrhs, info);
}

// Releases the cached graph and the item buffers (fla_hip_finalize; the
// device must be idle).  Everything is recreated on the next batched call.
int fla_batched_release() {
  int st = FLA_HIP_OK;
  if (g_graph.exec && hipGraphExecDestroy(g_graph.exec) != hipSuccess) st = FLA_HIP_EARG;
  if (g_graph.graph && hipGraphDestroy(g_graph.graph) != hipSuccess) st = FLA_HIP_EARG;
  g_graph.exec = nullptr;
  g_graph.graph = nullptr;
  g_graph.valid = false;
  if (g_items_dev && hipFree(g_items_dev) != hipSuccess) st = FLA_HIP_EARG;
  if (g_items_pin && hipHostFree(g_items_pin) != hipSuccess) st = FLA_HIP_EARG;
  if (g_counter && hipFree(g_counter) != hipSuccess) st = FLA_HIP_EARG;
  g_items_dev = nullptr;
  g_items_pin = nullptr;
  g_counter = nullptr;
  g_items_cap = 0;
  return st;
}

extern "C" int fla_hip_batched_set_graph(int on) {
  g_use_graph = on != 0;
  return FLA_HIP_OK;
}

extern "C" int fla_hip_batched_set_config(int kind, int nb, int wg_per_cu) {
  if (kind < 0 || kind >= kKinds || !valid_nb(kind, nb) || wg_per_cu < 1 || wg_per_cu > 4)
    return FLA_HIP_EARG;
  g_tuned[kind].clear();
  BatchConfig c;
  c.nb = nb;
  c.wg_per_cu = wg_per_cu;
  g_tuned[kind].push_back({Signature(), c});  // empty signature: see lookup below
  return FLA_HIP_OK;
}

// Times every candidate configuration on synthetic matrices of the given
// sizes (SPD for the Cholesky kinds, random for the LU) and keeps the
// fastest for this batch's size signature.  Returns the choice.
extern "C" int fla_hip_batched_tune(int kind, int count, const int* n, const int* nrhs,
                                    int* best_nb, int* best_wg, double* best_ms) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (kind < 0 || kind >= kKinds || count <= 0 || !n) return FLA_HIP_EARG;
  hipStream_t s = fla_ctx().s_update;
  const bool f32 = kind == kSpotrf;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  size_t total = 0, total_b = 0;
  for (int i = 0; i < count; ++i) {
    if (n[i] < 1 || n[i] > MAXN) return FLA_HIP_EARG;
    total += (size_t)n[i] * n[i];
    total_b += nrhs ? (size_t)n[i] * std::max(0, nrhs[i]) : 0;
  }
  char* buf = nullptr;
  FLA_HIP_CHECK(hipMalloc(&buf, (total + total_b) * esz + (size_t)count * (MAXN + 1) * sizeof(int) + 256));
  std::vector<void*> Ap(count), Bp(count);
  std::vector<int*> Pp(count);
  std::vector<int> lda(count), ldb(count);
  size_t off = 0, offb = total * esz;
  int* ipiv_base = reinterpret_cast<int*>(buf + (total + total_b) * esz);
  int* info = ipiv_base + (size_t)count * MAXN;
  for (int i = 0; i < count; ++i) {
    Ap[i] = buf + off;
    lda[i] = n[i];
    off += (size_t)n[i] * n[i] * esz;
    Bp[i] = buf + offb;
    ldb[i] = n[i];
    offb += nrhs ? (size_t)n[i] * std::max(0, nrhs[i]) * esz : 0;
    Pp[i] = ipiv_base + (size_t)i * MAXN;
  }
  std::vector<BatchConfig> cands;
  for (int nb : (kind == kDgetrf ? std::vector<int>{16} : std::vector<int>{16, 32}))
    for (int wg = 1; wg <= max_wg_per_cu(kind, nb); wg *= 2) {
      BatchConfig c;
      c.nb = nb;
      c.wg_per_cu = wg;
      cands.push_back(c);
    }
  hipEvent_t e0, e1;
  FLA_HIP_CHECK(hipEventCreate(&e0));
  FLA_HIP_CHECK(hipEventCreate(&e1));
  double best = 1e300;
  BatchConfig bestc = default_config(kind);
  const bool save_graph = g_use_graph;
  g_use_graph = false;
  int st = FLA_HIP_OK;
  for (const BatchConfig& c : cands) {
    float ms_sum = 0.f;
    for (int rep = 0; rep < 3 && st == FLA_HIP_OK; ++rep) {
      for (int i = 0; i < count; ++i) {
        const long nn = (long)n[i] * n[i];
        hipLaunchKernelGGL(tune_fill_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                           n[i], f32 ? nullptr : static_cast<double*>(Ap[i]),
                           f32 ? static_cast<float*>(Ap[i]) : nullptr, kind != kDgetrf ? 1 : 0,
                           (unsigned long long)i);
      }
      st = stage_items(count, n, Ap.data(), lda.data(), Pp.data(), nrhs ? Bp.data() : nullptr,
                       ldb.data(), nrhs, kind == kDgetrf);
      if (st != FLA_HIP_OK) break;
      FLA_HIP_CHECK(hipMemcpyAsync(g_items_dev, g_items_pin, sizeof(BatchItem) * count,
                                   hipMemcpyHostToDevice, s));
      FLA_HIP_CHECK(hipEventRecord(e0, s));
      st = launch_kind(s, kind, c, count, info);
      FLA_HIP_CHECK(hipEventRecord(e1, s));
      FLA_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      FLA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) ms_sum += ms;  // first run warms up
    }
    if (st != FLA_HIP_OK) break;
    if (ms_sum / 2 < best) {
      best = ms_sum / 2;
      bestc = c;
    }
  }
  g_use_graph = save_graph;
  FLA_HIP_CHECK(hipEventDestroy(e0));
  FLA_HIP_CHECK(hipEventDestroy(e1));
  FLA_HIP_CHECK(hipDeviceSynchronize());
  FLA_HIP_CHECK(hipFree(buf));
  if (st != FLA_HIP_OK) return st;
  const Signature sig = signature_of(count, n, nrhs);
  bool found = false;
  for (TunedEntry& e : g_tuned[kind])
    if (e.sig == sig) {
      e.cfg = bestc;
      found = true;
    }
  if (!found) g_tuned[kind].push_back({sig, bestc});
  if (best_nb) *best_nb = bestc.nb;
  if (best_wg) *best_wg = bestc.wg_per_cu;
  if (best_ms) *best_ms = best;
  return FLA_HIP_OK;
}
