// Reduced-precision trailing updates of the mixed-precision factorizations:
//   C (FP32, column-major) += alpha * P * Q^T
// with P, Q rounded to BF16 (v_mfma_f32_16x16x32_bf16) or FP32
// (v_mfma_f32_16x16x4f32), FP32 accumulation in both cases.  Cholesky: P and
// Q are rows of the factored panel L21; LU: P = L21 and Q = U12^T.
//
// The panel is first packed into "fragment order": for each 16-row tile rt
// and k-block kb (32 k for BF16, 16 k for FP32) one 1 KB block holds, for
// lane l of a wave, exactly the 16 bytes that lane feeds to the MFMA:
//   BF16: row rt*16 + l%16, k = kb*32 + 8*(l/16) + e      (e = 0..7)
//   FP32: row rt*16 + l%16, k = kb*16 + 4*e + l/16        (e = 0..3, MFMA step e)
// so every operand load of the update kernel is one coalesced 1 KB wave read
// with no LDS staging.  Rows are zero-padded to a multiple of 128 (plus one
// spare tile row) and k to a whole block.
//
// Per-panel scaling: the packed values are multiplied by 2^-e, with e the
// binary exponent of the operand's largest magnitude (so they lie in
// [-1, 1)); the update multiplies its FP32 result by 2^(e_P + e_Q).  Both factors are powers of
// two, so the scaling is exact; it keeps the packed operands in the normal
// range of the narrow format.
#include "fla_hip_internal.h"

#include <hip/hip_bf16.h>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// |v| of the panel (m x k, FP64) -> atomic max on the float bit pattern
// (non-negative floats order like their bit patterns).
__global__ void panel_absmax_kernel(int m, int k, const double* __restrict__ P, long ldp,
                                    unsigned* __restrict__ out) {
  __shared__ float red[256];
  float mx = 0.0f;
  const long total = (long)m * k;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % m), j = (int)(e / m);
    const float v = (float)fabs(P[(long)i + (long)j * ldp]);
    mx = v > mx ? v : mx;
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x])
      red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

// 2^-e for the largest magnitude amax = f * 2^e, f in [0.5, 1); 1 for amax 0
// or a non-finite amax (the factorization reports those separately).
__device__ __forceinline__ float pow2_scale(const unsigned* amax_bits) {
  if (!amax_bits) return 1.0f;
  const float a = __uint_as_float(*amax_bits);
  if (!(a > 0.0f) || !isfinite(a)) return 1.0f;
  int e;
  frexpf(a, &e);
  return ldexpf(1.0f, -e);
}

// Element (row, kk) of the operand is P[row + kk*ldp], or P[kk + row*ldp]
// when `trans` (the rows of U12^T are the columns of U12).
template <bool BF16>
__global__ void lp_pack_kernel(int m, int k, const double* __restrict__ P, long ldp, int trans,
                               int rtiles, int nkb, const unsigned* __restrict__ amax_bits,
                               void* out) {
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const long total = (long)rtiles * nkb * 64;
  if (idx >= total) return;
  const int lane = (int)(idx & 63);
  const long rest = idx >> 6;
  const int kb = (int)(rest % nkb), rt = (int)(rest / nkb);
  const int row = rt * 16 + (lane & 15);
  const float sc = pow2_scale(amax_bits);
  if (BF16) {
    bf8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int kk = kb * 32 + 8 * (lane >> 4) + e;
      const long at = trans ? (long)kk + (long)row * ldp : (long)row + (long)kk * ldp;
      const float x = (row < m && kk < k) ? (float)P[at] * sc : 0.0f;
      v[e] = (__bf16)x;
    }
    reinterpret_cast<bf8*>(out)[idx] = v;
  } else {
    f4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int kk = kb * 16 + 4 * e + (lane >> 4);
      const long at = trans ? (long)kk + (long)row * ldp : (long)row + (long)kk * ldp;
      v[e] = (row < m && kk < k) ? (float)P[at] * sc : 0.0f;
    }
    reinterpret_cast<f4*>(out)[idx] = v;
  }
}

// One 128 x 128 C tile per workgroup of 4 waves (2 x 2), 64 x 64 per wave =
// 4 x 4 MFMA blocks.  Q fragments are the MFMA A operand and P fragments the
// B operand, so D(x, y) = sum_k Q(x,k) P(y,k) = C(y, x); lane l, register r
// holds C(i0 + l%16, j0 + 4*(l/16) + r) of its 16 x 16 block: 16 consecutive
// rows of a column per store.
template <bool BF16, bool LOWER>
__global__ __launch_bounds__(256) void lp_update_kernel(int m, int n, int nkb,
                                                        const void* __restrict__ Pv, int prt0,
                                                        const void* __restrict__ Qv, int qrt0,
                                                        float alpha,
                                                        const unsigned* __restrict__ amax_p,
                                                        const unsigned* __restrict__ amax_q,
                                                        float* __restrict__ C, long ldc, int Tm,
                                                        int Tn, int group,
                                                        const int* __restrict__ abort) {
  if (abort && *abort != kSuccess) return;
  int ti, tj;
  fla_tile_of(blockIdx.x, gridDim.x, LOWER ? 1 : 0, Tm, Tn, group, ti, tj);
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int wr = w >> 1, wc = w & 1;
  const int i0 = ti * 128 + wr * 64, j0 = tj * 128 + wc * 64;
  if (LOWER && i0 + 63 < j0) return;  // this wave's 64 x 64 block is above the diagonal
  const int prt = prt0 + (i0 >> 4), qrt = qrt0 + (j0 >> 4);
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  if (BF16) {
    const bf8* P = reinterpret_cast<const bf8*>(Pv);
    const bf8* Q = reinterpret_cast<const bf8*>(Qv);
    bf8 pa[4], qb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) pa[a] = P[((long)(prt + a) * nkb) * 64 + lane];
#pragma unroll
    for (int b = 0; b < 4; ++b) qb[b] = Q[((long)(qrt + b) * nkb) * 64 + lane];
    for (int kb = 0; kb < nkb; ++kb) {
      bf8 pn[4], qn[4];
      const int kn = (kb + 1 < nkb) ? kb + 1 : kb;
#pragma unroll
      for (int a = 0; a < 4; ++a) pn[a] = P[((long)(prt + a) * nkb + kn) * 64 + lane];
#pragma unroll
      for (int b = 0; b < 4; ++b) qn[b] = Q[((long)(qrt + b) * nkb + kn) * 64 + lane];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qb[b], pa[a], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a) pa[a] = pn[a];
#pragma unroll
      for (int b = 0; b < 4; ++b) qb[b] = qn[b];
    }
  } else {
    const f4* P = reinterpret_cast<const f4*>(Pv);
    const f4* Q = reinterpret_cast<const f4*>(Qv);
    f4 pa[4], qb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) pa[a] = P[((long)(prt + a) * nkb) * 64 + lane];
#pragma unroll
    for (int b = 0; b < 4; ++b) qb[b] = Q[((long)(qrt + b) * nkb) * 64 + lane];
    for (int kb = 0; kb < nkb; ++kb) {
      f4 pn[4], qn[4];
      const int kn = (kb + 1 < nkb) ? kb + 1 : kb;
#pragma unroll
      for (int a = 0; a < 4; ++a) pn[a] = P[((long)(prt + a) * nkb + kn) * 64 + lane];
#pragma unroll
      for (int b = 0; b < 4; ++b) qn[b] = Q[((long)(qrt + b) * nkb + kn) * 64 + lane];
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b)
            acc[a][b] =
                __builtin_amdgcn_mfma_f32_16x16x4f32(qb[b][e], pa[a][e], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a) pa[a] = pn[a];
#pragma unroll
      for (int b = 0; b < 4; ++b) qb[b] = qn[b];
    }
  }
  const float f = alpha / (pow2_scale(amax_p) * pow2_scale(amax_q));
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + a * 16 + (lane & 15);
    if (i >= m) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = j0 + b * 16 + 4 * (lane >> 4) + r;
        if (j >= n || (LOWER && i < j)) continue;
        float* c = C + (long)i + (long)j * ldc;
        *c += f * acc[a][b][r];
      }
    }
  }
}

}  // namespace

long lp_panel_elems(int m, int k, bool bf16) {
  const int kbw = bf16 ? 32 : 16;
  const long rows = ((long)(m + 127) / 128 + 1) * 128;
  const long nkb = (k + kbw - 1) / kbw;
  return rows * nkb * kbw;
}

int launch_lp_absmax(hipStream_t s, int m, int k, const double* P, long ldp, unsigned* out) {
  if (m <= 0 || k <= 0) return FLA_HIP_OK;
  const long total = (long)m * k;
  long blocks = (total + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(panel_absmax_kernel, dim3((unsigned)blocks), dim3(256), 0, s, m, k, P, ldp,
                     out);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

int launch_lp_pack(hipStream_t s, bool bf16, int m, int k, const double* P, long ldp,
                   const unsigned* amax_bits, void* out, bool trans) {
  const int kbw = bf16 ? 32 : 16;
  const int rtiles = (int)(((long)(m + 127) / 128 + 1) * 8);
  const int nkb = (k + kbw - 1) / kbw;
  const long total = (long)rtiles * nkb * 64;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  if (bf16)
    hipLaunchKernelGGL(lp_pack_kernel<true>, dim3(blocks), dim3(256), 0, s, m, k, P, ldp,
                       trans ? 1 : 0, rtiles, nkb, amax_bits, out);
  else
    hipLaunchKernelGGL(lp_pack_kernel<false>, dim3(blocks), dim3(256), 0, s, m, k, P, ldp,
                       trans ? 1 : 0, rtiles, nkb, amax_bits, out);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// C (m x n) += alpha * P[prow0 + i, :] * Q[qrow0 + j, :]^T, lower part only
// when `lower` (C's top-left element on the diagonal).  prow0 and qrow0 are
// multiples of 16; the packed panels must hold prow0 + 128*ceil(m/128) and
// qrow0 + 128*ceil(n/128) rows (lp_panel_elems pads for that).
int launch_lp_update(hipStream_t s, bool bf16, bool lower, int m, int n, int k, const void* P,
                     int prow0, const void* Q, int qrow0, float alpha, const unsigned* amax_p,
                     const unsigned* amax_q, float* C, long ldc, const int* abort_flag) {
  if (m <= 0 || n <= 0 || k <= 0) return FLA_HIP_OK;
  if ((prow0 & 15) || (qrow0 & 15)) return FLA_HIP_EARG;
  const int kbw = bf16 ? 32 : 16;
  const int nkb = (k + kbw - 1) / kbw;
  const int Tm = (m + 127) / 128, Tn = (n + 127) / 128;
  const long nblk = lower ? (long)Tn * Tm - (long)Tn * (Tn - 1) / 2 : (long)Tm * Tn;
  if (nblk <= 0) return FLA_HIP_OK;
  if (lower && Tm < Tn) return FLA_HIP_EARG;
  const dim3 grid((unsigned)nblk), block(256);
  if (bf16 && lower)
    hipLaunchKernelGGL((lp_update_kernel<true, true>), grid, block, 0, s, m, n, nkb, P, prow0 / 16,
                       Q, qrow0 / 16, alpha, amax_p, amax_q, C, ldc, Tm, Tn, g_tile_group, abort_flag);
  else if (bf16)
    hipLaunchKernelGGL((lp_update_kernel<true, false>), grid, block, 0, s, m, n, nkb, P,
                       prow0 / 16, Q, qrow0 / 16, alpha, amax_p, amax_q, C, ldc, Tm, Tn, g_tile_group, abort_flag);
  else if (lower)
    hipLaunchKernelGGL((lp_update_kernel<false, true>), grid, block, 0, s, m, n, nkb, P,
                       prow0 / 16, Q, qrow0 / 16, alpha, amax_p, amax_q, C, ldc, Tm, Tn, g_tile_group, abort_flag);
  else
    hipLaunchKernelGGL((lp_update_kernel<false, false>), grid, block, 0, s, m, n, nkb, P,
                       prow0 / 16, Q, qrow0 / 16, alpha, amax_p, amax_q, C, ldc, Tm, Tn, g_tile_group, abort_flag);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}
