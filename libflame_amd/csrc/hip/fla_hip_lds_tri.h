// In-LDS inversion of a 128 x 128 lower-triangular block, shared by the
// diagonal-block inversion of the triangular solves and the fused
// factor-and-invert kernel of the Cholesky panel.
//
// S holds L column-major with leading dimension 129 (S[c*129 + i] = L(i,c)),
// identity beyond the block's order.  Recursive 2x2 blocking,
//   inv([A 0; B C]) = [inv(A) 0; -inv(C) B inv(A)  inv(C)],
// with the four 32x32 diagonal blocks solved by one wave each (lane = column
// of the inverse, forward substitution, L read by broadcast) and the
// off-diagonal blocks at sizes 64 and 128 formed as register-held products.
// On return S holds inv(L) in the same layout.  256 threads; the caller
// synchronizes before and after.
#pragma once

#include <hip/hip_runtime.h>

namespace fla_lds {

constexpr int TB = 128;
constexpr int TLD = TB + 1;

typedef double d4 __attribute__((ext_vector_type(4)));

// D := X * Y for H x H blocks of S (X at (xr, xc), Y at (yr, yc)), on
// v_mfma_f64_16x16x4f64.  Tile q of the (H/16)^2 output tiles of block-pair
// p is handled by wave (p * tiles + q) % 4; acc[u] holds the u-th tile this
// wave owns: lane (li, lk), register r = D(16*ti + li, 16*tj + lk + 4r).
template <int H, int NPAIR>
__device__ __forceinline__ void lds_mm(const double* S, int w, int lane, int xr0, int xc0,
                                       int yr0, int yc0, int pair_stride, d4 (&acc)[(H / 16) * (H / 16) * NPAIR / 4 > 0 ? (H / 16) * (H / 16) * NPAIR / 4 : 1]) {
  constexpr int NT = (H / 16) * (H / 16);
  constexpr int PER = NT * NPAIR / 4 > 0 ? NT * NPAIR / 4 : 1;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int g = w + 4 * u;  // global tile index
    const int p = g / NT, q = g % NT;
    const int ti = q % (H / 16), tj = q / (H / 16);
    const int xr = xr0 + p * pair_stride, xc = xc0 + p * pair_stride;
    const int yr = yr0 + p * pair_stride, yc = yc0 + p * pair_stride;
    d4 c = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < H / 4; ++ks) {
      const int k = 4 * ks + lk;
      const double xf = S[(xc + k) * TLD + xr + 16 * ti + li];   // X(row, k)
      const double yf = S[(yc + 16 * tj + li) * TLD + yr + k];   // Y(k, col)
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(yf, xf, c, 0, 0, 0);
    }
    acc[u] = c;
  }
}

template <int H, int NPAIR>
__device__ __forceinline__ void lds_store(double* S, int w, int lane, int r0, int c0,
                                          int pair_stride, double sign,
                                          const d4 (&acc)[(H / 16) * (H / 16) * NPAIR / 4 > 0 ? (H / 16) * (H / 16) * NPAIR / 4 : 1]) {
  constexpr int NT = (H / 16) * (H / 16);
  constexpr int PER = NT * NPAIR / 4 > 0 ? NT * NPAIR / 4 : 1;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int g = w + 4 * u;
    const int p = g / NT, q = g % NT;
    const int ti = q % (H / 16), tj = q / (H / 16);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      S[(c0 + p * pair_stride + 16 * tj + lk + 4 * r) * TLD + r0 + p * pair_stride + 16 * ti + li] =
          sign * acc[u][r];
  }
}

// One level of the recursive inversion: for each of the NPAIR 2H-blocks at
// offset o = p*2H, the lower-left H x H block B becomes -inv(C) * B * inv(A)
// (A, C already inverted; zero strictly upper parts, so the dense products
// equal the triangular ones).
template <int H>
__device__ __forceinline__ void trtri_combine(double* S, int t) {
  constexpr int NPAIR = TB / (2 * H);
  constexpr int PER = (H / 16) * (H / 16) * NPAIR / 4;
  const int w = t >> 6, lane = t & 63;
  d4 acc[PER];
  // Tmp = B * inv(A)
  lds_mm<H, NPAIR>(S, w, lane, H, 0, 0, 0, 2 * H, acc);
  __syncthreads();
  lds_store<H, NPAIR>(S, w, lane, H, 0, 2 * H, 1.0, acc);
  __syncthreads();
  // B := -inv(C) * Tmp
  lds_mm<H, NPAIR>(S, w, lane, H, H, H, 0, 2 * H, acc);
  __syncthreads();
  lds_store<H, NPAIR>(S, w, lane, H, 0, 2 * H, -1.0, acc);
  __syncthreads();
}

// inv(L) in place in S (see above).  Requires every thread of the block.
__device__ __forceinline__ void trtri_lower_lds(double* S, int t) {
  __shared__ double rdiag[TB];
  if (t < TB) rdiag[t] = 1.0 / S[t * TLD + t];
  __syncthreads();
  {
    const int w = t >> 6, lane = t & 63;
    if (lane < 32) {
      const int o = 32 * w, c = lane;
      double y[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        double sum = (i == c) ? 1.0 : 0.0;
#pragma unroll
        for (int kk = 0; kk < i; ++kk) sum -= S[(o + kk) * TLD + o + i] * y[kk];
        y[i] = sum * rdiag[o + i];
        // Pins y[i] (and the LDS reads) to this iteration: otherwise the
        // compiler hoists every broadcast read of later rows and spills.
        asm volatile("" : "+v"(y[i])::"memory");
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int i = 0; i < 32; ++i)
        if (i >= c) S[(o + c) * TLD + o + i] = y[i];
    }
  }
  __syncthreads();
  trtri_combine<32>(S, t);
  trtri_combine<64>(S, t);
}

}  // namespace fla_lds
