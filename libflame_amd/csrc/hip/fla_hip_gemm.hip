// FP64 MFMA trailing-update kernel for gfx950 (CDNA4).
//
// Two kernels serve every level-3 update on the hot path:
//   * FLA_Chol_l_blk_var3's A22 := A22 - A21*A21'   (lower triangle only)
//     src/lapack/dec/chol/l/flamec/FLA_Chol_l_blk_var3.c:52-54 -> dsyrk_('L','N')
//   * FLA_LU_piv_blk_var5's A22 := A22 - A21*A12
//     src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_blk_var5.c:94-96 -> dgemm_('N','N')
//   * the FLA_Gemm / FLA_Syrk entry points of the thin C ABI.
//
// Tiling: a 512-thread workgroup (8 wave64s in a 4x2 grid) owns a 128x128
// tile of C; each wave owns 32x64 = 2x4 blocks of v_mfma_f64_16x16x4_f64.  K
// is streamed in chunks of 16 (8 when B is k-contiguous) through two LDS
// buffers (load chunk t+1 into registers while chunk t feeds the MFMAs, then
// one barrier per chunk).  The updates proper (A and B with the tile rows
// contiguous, full tiles) run dgemm_nt_kp_kernel, which keeps each chunk in
// LDS as k pairs so that one 16-byte read feeds two MFMA steps;
// dgemm_mfma_kernel (8-byte layout, any transposition, edge tiles) takes the
// rest.
// The MFMA is issued with the B fragment as its "A" operand so that each
// lane's four accumulators are four columns of one row index: for fixed
// accumulator slot, lanes 0..15 hold 16 consecutive rows of one column, and
// the epilogue stores are 128-byte contiguous runs of column-major C.
#include "fla_hip_internal.h"

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

// Loads the BM x BK chunk of op(X) starting at (r0, k0) into registers
// (NT threads, PER = BM*BK/NT values each).
//   T == false: op(X)(r,l) = X[r + l*ld]   (rows contiguous: lanes walk r)
//   T == true : op(X)(r,l) = X[l + r*ld]   (k contiguous: lanes walk l)
template <bool T, int BM, int BK, int NT>
__device__ __forceinline__ void load_chunk(const double* X, long ld, int rows, int kdim, int r0,
                                           int k0, double (&reg)[BM * BK / NT], int tid) {
  constexpr int PER = BM * BK / NT;
  const bool full = (r0 + BM <= rows) && (k0 + BK <= kdim);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + q * NT;
    const int r = T ? (e / BK) : (e % BM);
    const int l = T ? (e % BK) : (e / BM);
    const double* p = T ? X + (long)(k0 + l) + (long)(r0 + r) * ld
                        : X + (long)(r0 + r) + (long)(k0 + l) * ld;
    if (full)
      reg[q] = *p;
    else
      reg[q] = (r0 + r < rows && k0 + l < kdim) ? *p : 0.0;
  }
}

// Stores registers from load_chunk into LDS laid out as S[l][r] (stride BM+1).
template <bool T, int BM, int BK, int NT>
__device__ __forceinline__ void store_chunk(double* S, const double (&reg)[BM * BK / NT], int tid) {
  constexpr int PER = BM * BK / NT;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + q * NT;
    const int r = T ? (e / BK) : (e % BM);
    const int l = T ? (e % BK) : (e / BM);
    S[l * (BM + 1) + r] = reg[q];
  }
}

// Stores alpha * acc + beta * C for one BM x BN tile at (i0, j0) (mode 1:
// the lower triangle only).  v_mfma_f64_16x16x4f64 leaves D(x, y) in lane
// 16*(x%4) + y, register x/4 (measured on gfx950); with x = column, y = row
// (the B fragment is the MFMA's first operand), lane holds
// C(i0+wr*BM/WM+a*16+li, j0+wc*BN/WN+b*16+lk+4*r) in acc[a][b][r].
// Interior tiles (inside C and, in lower mode, strictly below the diagonal)
// take a branch-free path: per 16-row block a and group of EB column blocks
// every C value the lane needs is loaded first and only then combined and
// stored, so the loads are in flight together (one memory round trip per
// group instead of one per element; the group size keeps the kernel within
// 128 VGPRs).  Edge and diagonal tiles store through per-element guards.
// Addresses are the tile's base (uniform) plus a 32-bit lane offset.
template <int BM, int BN, int WM, int WN, int TM, int TN>
__device__ __forceinline__ void store_tile(const v4d (&acc)[TM][TN], int mode, int m, int n,
                                           double alpha, double beta, double* C, long ldc,
                                           int i0, int j0, int wr, int wc, int li, int lk) {
  constexpr int EB = TN >= 2 ? 2 : 1;
  double* Ct = C + (long)i0 + (long)j0 * ldc;
  const int ldci = (int)ldc;
  // (32-bit lane offsets: the tile spans fewer than 2^31 elements)
  const bool interior = (i0 + BM <= m) && (j0 + BN <= n) && (mode != 1 || i0 >= j0 + BN - 1) &&
                        ldc < (1L << 23);
  if (interior) {
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const int di = wr * (BM / WM) + a * 16 + li;
#pragma unroll
      for (int b0 = 0; b0 < TN; b0 += EB) {
        double cv[EB][4];
        if (beta != 0.0) {
#pragma unroll
          for (int b = 0; b < EB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              cv[b][r] = Ct[(unsigned)(di + (wc * (BN / WN) + (b0 + b) * 16 + lk + 4 * r) * ldci)];
        }
#pragma unroll
        for (int b = 0; b < EB; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double v = alpha * acc[a][b0 + b][r];
            Ct[(unsigned)(di + (wc * (BN / WN) + (b0 + b) * 16 + lk + 4 * r) * ldci)] =
                (beta == 0.0) ? v : v + beta * cv[b][r];
          }
      }
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < TM; ++a) {
    const int di = wr * (BM / WM) + a * 16 + li;
    const int i = i0 + di;
#pragma unroll
    for (int b = 0; b < TN; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dj = wc * (BN / WN) + b * 16 + lk + 4 * r;
        const int j = j0 + dj;
        if (i >= m || j >= n || (mode == 1 && i < j)) continue;
        double* c = Ct + (long)di + (long)dj * ldc;
        const double v = alpha * acc[a][b][r];
        *c = (beta == 0.0) ? v : v + beta * (*c);
      }
    }
  }
}

// C := alpha*op(A)*op(B) + beta*C on a BM x BN tile per workgroup of
// WM x WN waves; each wave owns (BM/WM) x (BN/WN) = TM x TN MFMA blocks.
template <bool TA, bool TB, int BM, int BN, int BK, int WM, int WN, bool DB>
__global__ __launch_bounds__(64 * WM * WN)
__attribute__((amdgpu_waves_per_eu(WM * WN >= 8 ? 4 : 2))) void dgemm_mfma_kernel(
    int mode, int m, int n, int k, double alpha, const double* A, long lda, const double* B,
    long ldb, double beta, double* C, long ldc, int Tm, int Tn, int rpc,
    const int* __restrict__ abort_flag) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;
  constexpr int LA = BM + 1, LB = BN + 1;
  constexpr int PA = BM * BK / NT, PB = BN * BK / NT;
  if (abort_flag && *abort_flag >= 0) return;
  __shared__ double As[2][BK * LA];
  __shared__ double Bs[2][BK * LB];

  int ti, tj;
  fla_tile_of(blockIdx.x, gridDim.x, mode, Tm, Tn, rpc, ti, tj);
  const int i0 = ti * BM, j0 = tj * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;

  v4d acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  const int nk = (k + BK - 1) / BK;
  double ra[PA], rb[PB];
  if (nk > 0) {
    load_chunk<TA, BM, BK, NT>(A, lda, m, k, i0, 0, ra, tid);
    load_chunk<!TB, BN, BK, NT>(B, ldb, n, k, j0, 0, rb, tid);
    store_chunk<TA, BM, BK, NT>(As[0], ra, tid);
    store_chunk<!TB, BN, BK, NT>(Bs[0], rb, tid);
  }
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      load_chunk<TA, BM, BK, NT>(A, lda, m, k, i0, (kt + 1) * BK, ra, tid);
      load_chunk<!TB, BN, BK, NT>(B, ldb, n, k, j0, (kt + 1) * BK, rb, tid);
    }
    const double* as = As[cur];
    const double* bs = Bs[cur];
    if constexpr (DB) {
      // Fragments of step kk+1 are read from LDS before the MFMAs of step kk
      // are issued (two register sets), so the LDS latency of the next step
      // overlaps this step's MFMAs.
      double fa[2][TM], fb[2][TN];
#pragma unroll
      for (int t = 0; t < TM; ++t) fa[0][t] = as[lk * LA + wr * (BM / WM) + t * 16 + li];
#pragma unroll
      for (int t = 0; t < TN; ++t) fb[0][t] = bs[lk * LB + wc * (BN / WN) + t * 16 + li];
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int cb = kk & 1;
        if (kk + 1 < BK / 4) {
#pragma unroll
          for (int t = 0; t < TM; ++t)
            fa[cb ^ 1][t] = as[((kk + 1) * 4 + lk) * LA + wr * (BM / WM) + t * 16 + li];
#pragma unroll
          for (int t = 0; t < TN; ++t)
            fb[cb ^ 1][t] = bs[((kk + 1) * 4 + lk) * LB + wc * (BN / WN) + t * 16 + li];
        }
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[cb][b], fa[cb][a], acc[a][b], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        double fa[TM], fb[TN];
#pragma unroll
        for (int t = 0; t < TM; ++t) fa[t] = as[(kk * 4 + lk) * LA + wr * (BM / WM) + t * 16 + li];
#pragma unroll
        for (int t = 0; t < TN; ++t) fb[t] = bs[(kk * 4 + lk) * LB + wc * (BN / WN) + t * 16 + li];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b], fa[a], acc[a][b], 0, 0, 0);
      }
    }
    if (more) {
      store_chunk<TA, BM, BK, NT>(As[cur ^ 1], ra, tid);
      store_chunk<!TB, BN, BK, NT>(Bs[cur ^ 1], rb, tid);
    }
    __syncthreads();
  }

  store_tile<BM, BN, WM, WN, TM, TN>(acc, mode, m, n, alpha, beta, C, ldc, i0, j0, wr, wc, li, lk);
}

// C := alpha*A*B' + beta*C on full 128 x 128 tiles (m, n multiples of 128,
// k a multiple of 16), A and B both stored with the tile's rows contiguous
// (the Cholesky SYRK update, the LU update with U12 transposed).  Same tiling
// as the 8-wave kernel above; the K chunk is kept in LDS as k PAIRS: entry
// (kp, r) holds X(r, 2kp) and X(r, 2kp+1) in 16 bytes, and the MFMA's four k
// slots over two consecutive steps s = 0, 1 of a step pair take k = 8p + 2*lk
// + s, so one 16-byte LDS read gives a lane its operands for both steps
// (half the LDS reads and writes of the 8-byte layout; the sum over a step
// pair covers k = 8p .. 8p+7 once).  The loads need no bounds guards: the
// chunk's base is uniform and each lane adds a 32-bit offset.
constexpr int KP_BM = 128, KP_BK = 16;
typedef double kp_d2 __attribute__((ext_vector_type(2)));

// BK: K chunk (a multiple of 16); WM x WN waves, each (128/WM) x (128/WN);
// PAD: 16-byte entries between the LDS rows of consecutive k pairs.
template <bool VEC, int BK, int WM, int WN, int PAD>
__global__ __launch_bounds__(64 * WM * WN)
__attribute__((amdgpu_waves_per_eu((WM * WN >= 8 && BK <= 16) ? 4 : 2))) void dgemm_nt_kp_kernel(
    int mode, int m, int n, int k, double alpha, const double* A, long lda, const double* B,
    long ldb, double beta, double* C, long ldc, int Tm, int Tn, int rpc,
    const int* __restrict__ abort_flag) {
  constexpr int NT = 64 * WM * WN, TM = KP_BM / WM / 16, TN = KP_BM / WN / 16, NKP = BK / 2;
  constexpr int KP_LD = KP_BM + PAD;  // LDS row of one k pair, in 16-byte entries
  // VEC: 2 x 2 blocks (two rows, one k pair) per thread; else single k pairs
  constexpr int NREG = VEC ? 2 * (64 * NKP / NT) : KP_BM * NKP / NT;
  static_assert(!VEC || (64 * NKP) % NT == 0, "whole 2 x 2 blocks per thread");
  static_assert(VEC || (KP_BM * NKP) % NT == 0, "whole k pairs per thread");
  if (abort_flag && *abort_flag >= 0) return;
  __shared__ kp_d2 As[2][NKP * KP_LD];
  __shared__ kp_d2 Bs[2][NKP * KP_LD];
  int ti, tj;
  fla_tile_of(blockIdx.x, gridDim.x, mode, Tm, Tn, rpc, ti, tj);
  const int i0 = ti * KP_BM, j0 = tj * KP_BM;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WN, wc = wave % WN;
  const int li = lane & 15, lk = lane >> 4;
  // This thread's entries: VEC, the 2 x 2 blocks of rows 2 (tid % 64) + {0, 1}
  // and k pairs tid / 64 + q NT / 64 (16-byte loads of two rows of one
  // column; A and B 16-byte aligned with even leading dimensions); else the
  // k pairs of row tid % 128, pairs tid / 128 + q NT / 128.  The per-q and
  // per-column parts of the addresses are uniform, so one 32-bit lane offset
  // per operand remains.
  constexpr int KSTEP = VEC ? NT / 64 : NT / KP_BM;  // k pairs between a thread's entries
  const unsigned r = VEC ? 2 * (tid & 63) : tid & (KP_BM - 1);
  const unsigned kp0 = VEC ? tid / 64 : tid / KP_BM;
  const unsigned offa = r + 2 * kp0 * (unsigned)lda, offb = r + 2 * kp0 * (unsigned)ldb;
  const int sbase = (int)(kp0 * KP_LD + r);
  v4d acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  const int nk = k / BK;
  const double* Ab = A + (long)i0;  // chunk base: + k0 * lda
  const double* Bb = B + (long)j0;
  kp_d2 ra[NREG], rb[NREG];
  auto load = [&](int kc) {
#pragma unroll
    for (int q = 0; q < (VEC ? NREG / 2 : NREG); ++q) {
      const double* a0 = Ab + ((long)kc * BK + 2 * KSTEP * q) * lda;
      const double* b0 = Bb + ((long)kc * BK + 2 * KSTEP * q) * ldb;
      if constexpr (VEC) {
        // rows (r, r+1) of columns 2kp and 2kp + 1
        ra[2 * q] = *reinterpret_cast<const kp_d2*>(a0 + offa);
        ra[2 * q + 1] = *reinterpret_cast<const kp_d2*>(a0 + lda + offa);
        rb[2 * q] = *reinterpret_cast<const kp_d2*>(b0 + offb);
        rb[2 * q + 1] = *reinterpret_cast<const kp_d2*>(b0 + ldb + offb);
      } else {
        ra[q].x = a0[offa];
        ra[q].y = (a0 + lda)[offa];
        rb[q].x = b0[offb];
        rb[q].y = (b0 + ldb)[offb];
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < (VEC ? NREG / 2 : NREG); ++q) {
      const int s0 = sbase + KSTEP * q * KP_LD;
      if constexpr (VEC) {
        As[buf][s0] = kp_d2{ra[2 * q].x, ra[2 * q + 1].x};
        As[buf][s0 + 1] = kp_d2{ra[2 * q].y, ra[2 * q + 1].y};
        Bs[buf][s0] = kp_d2{rb[2 * q].x, rb[2 * q + 1].x};
        Bs[buf][s0 + 1] = kp_d2{rb[2 * q].y, rb[2 * q + 1].y};
      } else {
        As[buf][s0] = ra[q];
        Bs[buf][s0] = rb[q];
      }
    }
  };
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) load(kt + 1);
    const kp_d2* as = As[cur];
    const kp_d2* bs = Bs[cur];
#pragma unroll
    for (int p = 0; p < NKP / 4; ++p) {
      kp_d2 fa[TM], fb[TN];
#pragma unroll
      for (int t = 0; t < TM; ++t) fa[t] = as[(4 * p + lk) * KP_LD + wr * (KP_BM / WM) + t * 16 + li];
#pragma unroll
      for (int t = 0; t < TN; ++t) fb[t] = bs[(4 * p + lk) * KP_LD + wc * (KP_BM / WN) + t * 16 + li];
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b].x, fa[a].x, acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[b].y, fa[a].y, acc[a][b], 0, 0, 0);
    }
    if (more) store(cur ^ 1);
    __syncthreads();
  }
  store_tile<KP_BM, KP_BM, WM, WN, TM, TN>(acc, mode, m, n, alpha, beta, C, ldc, i0, j0, wr, wc,
                                           li, lk);
}

template <int BK, int WM, int WN, int PAD = 4>
int launch_kp(hipStream_t s, bool vec, int mode, int m, int n, int k, double alpha,
              const double* A, long lda, const double* B, long ldb, double beta, double* C,
              long ldc, const int* abort_flag) {
  const int Tm = m / KP_BM, Tn = n / KP_BM;
  const long nblk = (mode == 0) ? (long)Tm * Tn : (long)Tn * Tm - (long)Tn * (Tn - 1) / 2;
  if (nblk > 0x7fffffffL) return FLA_HIP_EARG;
  constexpr bool vec_ok = (64 * (BK / 2)) % (64 * WM * WN) == 0;
  if (vec && vec_ok)
    hipLaunchKernelGGL((dgemm_nt_kp_kernel<vec_ok, BK, WM, WN, PAD>), dim3((unsigned)nblk),
                       dim3(64 * WM * WN), 0, s, mode, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc, Tm, Tn, g_tile_group, abort_flag);
  else
    hipLaunchKernelGGL((dgemm_nt_kp_kernel<false, BK, WM, WN, PAD>), dim3((unsigned)nblk),
                       dim3(64 * WM * WN), 0, s, mode, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc, Tm, Tn, g_tile_group, abort_flag);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// Tile configurations (selected by fla_hip_set_gemm_config for tuning):
// <BM, BN, BK, waves in M, waves in N, fragment double-buffering>.
int g_gemm_cfg = 0;

template <int BM, int BN, int BK, int WM, int WN, bool DB = false>
int launch_cfg(hipStream_t s, bool ta, bool tb, int mode, int m, int n, int k, double alpha,
               const double* A, long lda, const double* B, long ldb, double beta, double* C,
               long ldc, const int* abort_flag) {
  static_assert(BM == BN, "the lower-tile enumeration assumes square tiles");
  const int Tm = (m + BM - 1) / BM, Tn = (n + BN - 1) / BN;
  const int rpc_eff = g_tile_group;  // tile-column group width (fla_tile_of)
  const long nblk = (mode == 0) ? (long)Tm * Tn : (long)Tn * Tm - (long)Tn * (Tn - 1) / 2;
  if (nblk <= 0) return FLA_HIP_OK;
  if (nblk > 0x7fffffffL) return FLA_HIP_EARG;
  dim3 grid((unsigned)nblk), block(64 * WM * WN);
#define FLA_GEMM_LAUNCH(TA_, TB_)                                                              \
  hipLaunchKernelGGL((dgemm_mfma_kernel<TA_, TB_, BM, BN, BK, WM, WN, DB>), grid, block, 0, s, \
                     mode, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, Tm, Tn, rpc_eff,       \
                     abort_flag)
  if (!ta && !tb)
    FLA_GEMM_LAUNCH(false, false);
  else if (!ta && tb)
    FLA_GEMM_LAUNCH(false, true);
  else if (ta && !tb)
    FLA_GEMM_LAUNCH(true, false);
  else
    FLA_GEMM_LAUNCH(true, true);
#undef FLA_GEMM_LAUNCH
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

}  // namespace

extern "C" int fla_hip_set_gemm_group(int G) {
  if (G < 1 || G > 64) return FLA_HIP_EARG;
  g_tile_group = G;
  return FLA_HIP_OK;
}

extern "C" int fla_hip_set_gemm_config(int cfg) {
  if (cfg < 0 || cfg > 19) return FLA_HIP_EARG;
  g_gemm_cfg = cfg;
  return FLA_HIP_OK;
}

int launch_dgemm_mfma(hipStream_t s, bool ta, bool tb, int mode, int m, int n, int k,
                      double alpha, const double* A, long lda, const double* B, long ldb,
                      double beta, double* C, long ldc, const int* abort_flag) {
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  if (mode == 1 && m < n) return FLA_HIP_EARG;  // lower mode expects a tall C
  // Default (0) and 12..15: the k-pair kernel wherever it applies (A and B
  // with the tile rows contiguous, full tiles, K a multiple of the chunk),
  // measured on MI355X at 54.9 TFLOP/s on the n = 32768 Cholesky update
  // against 51.0 for the 8-byte-layout kernel (configuration 2).
  if (g_gemm_cfg == 0 || g_gemm_cfg >= 12) {
    const int bk = (g_gemm_cfg == 15) ? 32 : KP_BK;
    if (!ta && tb && m % KP_BM == 0 && n % KP_BM == 0 && k % bk == 0 && k > 0 &&
        lda < (1L << 26) && ldb < (1L << 26)) {
      // 16-byte loads (configuration 12) measured no faster than 8-byte ones
      const bool vec = g_gemm_cfg == 12 && lda % 2 == 0 && ldb % 2 == 0 &&
                       reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(B) % 16 == 0;
      switch (g_gemm_cfg) {
        case 13:
          return launch_kp<16, 2, 2>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                     abort_flag);
        case 14:
          return launch_kp<16, 4, 4>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                     abort_flag);
        case 15:
          return launch_kp<32, 4, 2>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                     abort_flag);
        case 16:
          return launch_kp<16, 4, 2, 0>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                        ldc, abort_flag);
        case 17:
          return launch_kp<16, 4, 2, 1>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                        ldc, abort_flag);
        case 18:
          return launch_kp<16, 4, 2, 2>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                        ldc, abort_flag);
        case 19:
          return launch_kp<16, 4, 2, 8>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                        ldc, abort_flag);
        default:
          return launch_kp<16, 4, 2>(s, vec, mode, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                     abort_flag);
      }
    }
  }
  switch (g_gemm_cfg) {
    case 1:
      return launch_cfg<128, 128, 32, 2, 2, true>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                beta, C, ldc, abort_flag);
    case 2:
      return launch_cfg<128, 128, 16, 4, 2, false>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                 beta, C, ldc, abort_flag);
    case 3:
      return launch_cfg<128, 128, 32, 4, 2, true>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                beta, C, ldc, abort_flag);
    case 4:
      return launch_cfg<128, 128, 16, 2, 4, false>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                 beta, C, ldc, abort_flag);
    case 5:
      return launch_cfg<128, 128, 16, 4, 4, true>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                beta, C, ldc, abort_flag);
    case 6:
      return launch_cfg<128, 128, 8, 4, 2, true>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                               beta, C, ldc, abort_flag);
    case 7:
      return launch_cfg<128, 128, 16, 2, 2, true>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                beta, C, ldc, abort_flag);
    case 8:
      return launch_cfg<128, 128, 32, 4, 2, false>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                 beta, C, ldc, abort_flag);
    case 9:
      return launch_cfg<128, 128, 16, 2, 2, false>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                 beta, C, ldc, abort_flag);
    case 10:
      return launch_cfg<128, 128, 32, 2, 2, false>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                 beta, C, ldc, abort_flag);
    case 11:
      return launch_cfg<128, 128, 16, 4, 4, false>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                                 beta, C, ldc, abort_flag);
    default:
      // Shapes the k-pair kernel does not take.  Measured on MI355X
      // (n = 16384..32768, k = 128..256): 8 waves per 128x128 tile; K chunks
      // of 16 when B is stored with j contiguous (~51 TFLOP/s on the SYRK
      // update), of 8 when B is k-contiguous, where the smaller LDS
      // footprint doubles occupancy.
      if (!tb)
        return launch_cfg<128, 128, 8, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb,
                                             beta, C, ldc, abort_flag);
      return launch_cfg<128, 128, 16, 4, 2>(s, ta, tb, mode, m, n, k, alpha, A, lda, B, ldb, beta,
                                            C, ldc, abort_flag);
  }
}

static bool is_trans(int t) { return t == kTrans || t == kConjTrans; }

extern "C" int fla_hip_dgemm(int transa, int transb, int m, int n, int k, double alpha,
                             const double* A, int lda, const double* B, int ldb, double beta,
                             double* C, int ldc) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || k < 0) return FLA_HIP_EARG;
  if (m == 0 || n == 0) return FLA_HIP_OK;
  const bool ta = is_trans(transa), tb = is_trans(transb);
  if (ldc < m || lda < (ta ? k : m) || ldb < (tb ? n : k)) return FLA_HIP_EARG;
  if (k == 0 || alpha == 0.0) {
    if (beta == 1.0) return FLA_HIP_OK;
    // C := beta*C without reading A or B (src/flablas/f2c/dgemm.c:194-214):
    // run with k = 0 (the K loop is skipped, the epilogue applies beta and
    // writes 0 where beta == 0).
    k = 0;
    alpha = 0.0;
  }
  return launch_dgemm_mfma(fla_ctx().s_update, ta, tb, 0, m, n, k, alpha, A, lda, B, ldb,
                           beta, C, ldc, nullptr);
}

extern "C" int fla_hip_dsyrk(int uplo, int trans, int n, int k, double alpha,
                             const double* A, int lda, double beta, double* C, int ldc) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (n < 0 || k < 0) return FLA_HIP_EARG;
  if (n == 0) return FLA_HIP_OK;
  if (uplo != kLower && uplo != kUpper) return FLA_HIP_EARG;
  const bool t = is_trans(trans);
  if (ldc < n || lda < (t ? k : n)) return FLA_HIP_EARG;
  if (k == 0 || alpha == 0.0) {
    if (beta == 1.0) return FLA_HIP_OK;
    k = 0;  // C := beta*C on the triangle without reading A (dsyrk.c's quick return)
    alpha = 0.0;
  }
  hipStream_t s = fla_ctx().s_update;
  if (uplo == kLower) {
    // C(lower) += alpha * op(A) op(A)'   with op(A) n x k.
    return launch_dgemm_mfma(s, t, !t, 1, n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                             nullptr);
  }
  // Upper triangle of C is the lower triangle of C' (same storage viewed with
  // swapped strides); C(i,j) for i <= j equals C'(j,i): compute lower of
  // C' = alpha * op(A) op(A)' + beta*C'  ==  identical values, transposed
  // storage.  Exchange triangles, update the lower one, exchange back.
  FLA_TRY(launch_transpose_swap(s, n, C, ldc));
  FLA_TRY(launch_dgemm_mfma(s, t, !t, 1, n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                            nullptr));
  return launch_transpose_swap(s, n, C, ldc);
}
