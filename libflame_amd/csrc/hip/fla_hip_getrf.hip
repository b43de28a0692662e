// Blocked LU with partial pivoting on one MI355X.
//
// Reference algorithm: FLA_LU_piv_blk_var5
//   (src/lapack/dec/lu/piv/vars/flamec/FLA_LU_piv_blk_var5.c:13-114):
//     [A11;A21], p1 := LU_piv([A11;A21]);  apply p1 to [A10;A20] and [A12;A22];
//     A12 := inv(trilu(A11)) * A12;         A22 := A22 - A21 * A12
// with the leaf's pivot rule (bl1_damax -> idamax_, first maximal |x|;
// src/flablas/f2c/idamax.c:40-46), FLAME-native relative pivots
// (src/lapack/util/app/piv/ln/flamec/FLA_Apply_pivots_ln_opt_var1.c:323), and
// the null-pivot rule of FLA_LU_piv_opd_var4 (FLA_LU_piv_opt_var4.c:277-283,
// 325-332): the first null pivot's index is returned, its interchange and
// scaling are skipped, and the sweep continues.
//
// Panel: one cooperative kernel.  The (m-k) x nb panel is spread over P
// workgroups (256 rows each) and held in VGPRs for the whole panel; every
// column costs one grid-wide barrier, across which each workgroup publishes
// its local pivot candidate (and that row's values) and the owner of the
// diagonal row publishes that row.  All inter-workgroup traffic goes through
// uncached device memory with agent-scope atomic loads and stores (so no L2
// write-back or invalidate is needed across the XCDs), and every wait is
// bounded.
#include <hip/hip_runtime.h>

#include "fla_hip_internal.h"

namespace {

constexpr int PW = 128;           // maximum panel width
constexpr int ROWS_PER_WG = 256;  // rows of the panel held by one workgroup
constexpr int NT = 512;           // 8 waves: wave w owns columns [32(w&3), +32), row half w>>2
constexpr int RQ = 2;             // rows per thread
constexpr unsigned long long kSpinTicks = 200000000ULL;  // 2 s of the 100 MHz wall clock

// Per-workgroup candidate header (all doubles so one slot is uniform).
//   [0] best |v| over non-NaN rows >= j      [1] its row (or -1)  [2] its orig
//   [3] last NaN row (or -1)                  [4] its orig
//   [5] best |v| over non-NaN rows > last NaN [6] its row (or -1)  [7] its orig
constexpr int HDR = 8;
constexpr int MAX_P = 256;  // panel workgroups (one per CU at most)

struct PanelComm {
  double* hdr;      // [2][P][HDR]
  double* rows;     // [2][P][3][PW]   best row, NaN row, after-NaN row
  double* jrow;     // [2][PW + 1]     diagonal row values (+ orig)
  int* seq;         // [2][P]          column (+1) whose slot is published
  int* err;         // wait-limit flag
};

__device__ __forceinline__ void st_d(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_d(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// First-max combine of (value, row): larger value wins, ties -> smaller row.
__device__ __forceinline__ void better(double& bv, int& br, int& bo, double v, int r, int o) {
  if (r < 0) return;
  if (br < 0 || v > bv || (v == bv && r < br)) {
    bv = v;
    br = r;
    bo = o;
  }
}

__device__ __forceinline__ double pick(const double (&row)[32], int cj) {
  double x = 0.0;
#pragma unroll
  for (int c = 0; c < 32; ++c)
    if (c == cj) x = row[c];
  return x;
}

__global__ __launch_bounds__(NT) void lu_panel_kernel(int M, int W, double* __restrict__ A,
                                                      long lda, int k, int* __restrict__ piv,
                                                      int* __restrict__ info,
                                                      int* __restrict__ moved, PanelComm cm) {
  // A wait that expired in an earlier panel of this factorization aborts all later ones.
  if (__hip_atomic_load(cm.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  const unsigned P = gridDim.x;
  const int wg = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int cg = wave & 3;   // column group
  const int rh = wave >> 2;  // row half
  const int base = wg * ROWS_PER_WG;

  __shared__ double prow[PW];   // pivot row values
  __shared__ double jrow[PW];   // old diagonal row values
  __shared__ double cand[3][PW];
  __shared__ int jorig;
  __shared__ double lv[ROWS_PER_WG];
  __shared__ double wred[2][8];
  __shared__ double hsum[8];
  __shared__ int sel[4];        // winner row, winner orig, orig of the old diagonal row

  // Thread owns rows base + rh*128 + 64*q + lane (q = 0..1), columns 32*cg + c.
  double v[RQ][32];
  int orig[RQ];
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int r = base + rh * 128 + 64 * q + lane;
    orig[q] = r;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      const int col = 32 * cg + c;
      v[q][c] = (r < M && col < W) ? A[(long)r + (long)col * lda] : 0.0;
    }
  }

  for (int j = 0; j < W; ++j) {
    const int par = j & 1;
    const int cgj = j >> 5, cj = j & 31;
    double* hdr = cm.hdr + ((long)par * P + wg) * HDR;
    double* rows = cm.rows + ((long)par * P + wg) * 3 * PW;
    double* jslot = cm.jrow + (long)par * (PW + 1);

    // ---- local candidates: the two waves holding column j ----
    if (cg == cgj) {
      double bv = 0.0, av = 0.0;
      int br = -1, bo = -1, ar = -1, ao = -1, nr = -1, no = -1;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = base + rh * 128 + 64 * q + lane;
        const double x = pick(v[q], cj);
        if (r < M && r >= j) {
          if (x != x) {
            if (r > nr) {
              nr = r;
              no = orig[q];
            }
          } else {
            better(bv, br, bo, fabs(x), r, orig[q]);
          }
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int orr = __shfl_xor(br, off), oo = __shfl_xor(bo, off);
        better(bv, br, bo, ov, orr, oo);
        const int onr = __shfl_xor(nr, off), ono = __shfl_xor(no, off);
        if (onr > nr) {
          nr = onr;
          no = ono;
        }
      }
      if (lane == 0) {
        wred[rh][0] = bv;
        wred[rh][1] = br;
        wred[rh][2] = bo;
        wred[rh][3] = nr;
        wred[rh][4] = no;
      }
    }
    __syncthreads();
    if (t == 0) {
      double bv = wred[0][0];
      int br = (int)wred[0][1], bo = (int)wred[0][2];
      better(bv, br, bo, wred[1][0], (int)wred[1][1], (int)wred[1][2]);
      int nr = (int)wred[0][3], no = (int)wred[0][4];
      if ((int)wred[1][3] > nr) {
        nr = (int)wred[1][3];
        no = (int)wred[1][4];
      }
      hsum[0] = bv;
      hsum[1] = br;
      hsum[2] = bo;
      hsum[3] = nr;
      hsum[4] = no;
    }
    __syncthreads();
    // after-NaN candidate: first max over non-NaN rows > last NaN of this slab
    const int nr_wg = (int)hsum[3];
    if (nr_wg >= 0) {
      if (cg == cgj) {
        double av = 0.0;
        int ar = -1, ao = -1;
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          const int r = base + rh * 128 + 64 * q + lane;
          const double x = pick(v[q], cj);
          if (r < M && r > nr_wg && x == x) better(av, ar, ao, fabs(x), r, orig[q]);
        }
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(av, off);
          const int orr = __shfl_xor(ar, off), oo = __shfl_xor(ao, off);
          better(av, ar, ao, ov, orr, oo);
        }
        if (lane == 0) {
          wred[rh][5] = av;
          wred[rh][6] = ar;
          wred[rh][7] = ao;
        }
      }
      __syncthreads();
      if (t == 0) {
        double av = wred[0][5];
        int ar = (int)wred[0][6], ao = (int)wred[0][7];
        better(av, ar, ao, wred[1][5], (int)wred[1][6], (int)wred[1][7]);
        hsum[5] = av;
        hsum[6] = ar;
        hsum[7] = ao;
      }
    } else if (t == 0) {
      hsum[5] = 0.0;
      hsum[6] = -1.0;
      hsum[7] = -1.0;
    }
    __syncthreads();
    const int br = (int)hsum[1], nr = (int)hsum[3], ar = (int)hsum[6];
    // Stage candidate rows (and the diagonal row j) in LDS: one branch per
    // row around its 32 stores; waves holding none skip (a uniform test).
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const int r = base + rh * 128 + 64 * q + lane;
      if (!__any(r == br || r == nr || r == ar || r == j)) continue;
      if (r == br) {
#pragma unroll
        for (int c = 0; c < 32; ++c) cand[0][32 * cg + c] = v[q][c];
      }
      if (r == nr) {
#pragma unroll
        for (int c = 0; c < 32; ++c) cand[1][32 * cg + c] = v[q][c];
      }
      if (r == ar) {
#pragma unroll
        for (int c = 0; c < 32; ++c) cand[2][32 * cg + c] = v[q][c];
      }
      if (r == j) {
#pragma unroll
        for (int c = 0; c < 32; ++c) jrow[32 * cg + c] = v[q][c];
        if (cg == 0) jorig = orig[q];
      }
    }
    __syncthreads();
    // ---- wave 0: publish this workgroup's slot for column j, wait for every
    // slot of column j, then pick the global pivot ----
    if (wave == 0) {
      const bool owns_j = (j >= base && j < base + ROWS_PER_WG);
      if (lane < HDR) st_d(&hdr[lane], hsum[lane]);
      for (int e = lane; e < W; e += 64) {
        if (br >= 0) st_d(&rows[0 * PW + e], cand[0][e]);
        if (nr >= 0) st_d(&rows[1 * PW + e], cand[1][e]);
        if (ar >= 0) st_d(&rows[2 * PW + e], cand[2][e]);
        if (owns_j) st_d(&jslot[e], jrow[e]);
      }
      if (owns_j && lane == 0) st_d(&jslot[PW], (double)jorig);
      // The exchange buffers are uncached: waiting for the stores to be
      // acknowledged orders them before the sequence number, with no L2
      // write-back.
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0)
        __hip_atomic_store(&cm.seq[par * P + wg], j + 1, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      bool ok = true;
      const unsigned long long t0 = wall_clock64();
      while (true) {
        bool done = true;
        for (unsigned w = lane; w < P; w += 64)
          if (__hip_atomic_load(&cm.seq[par * P + w], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT) != j + 1)
            done = false;
        if (__all(done)) break;
        if (__hip_atomic_load(cm.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            wall_clock64() - t0 > kSpinTicks) {
          if (lane == 0) __hip_atomic_store(cm.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");  // uncached: no invalidate
      if (!ok) {
        if (lane == 0) sel[0] = -2;
      } else {
        double gnr = -1.0, gno = -1.0;
        // the first slot's no-NaN candidate comes in the same round trip
        double f0 = 0.0, f1 = -1.0, f2 = -1.0;
        for (unsigned w = lane; w < P; w += 64) {
          const double* h = cm.hdr + ((long)par * P + w) * HDR;
          const double hn = ld_d(&h[3]), ho = ld_d(&h[4]);
          const double h0 = ld_d(&h[0]), h1 = ld_d(&h[1]), h2 = ld_d(&h[2]);
          if (w == lane) {
            f0 = h0;
            f1 = h1;
            f2 = h2;
          }
          if (hn > gnr) {
            gnr = hn;
            gno = ho;
          }
        }
        for (int off = 32; off > 0; off >>= 1) {
          const double o = __shfl_xor(gnr, off), oo = __shfl_xor(gno, off);
          if (o > gnr) {
            gnr = o;
            gno = oo;
          }
        }
        // Candidates: without a NaN, every slab's best; with the last NaN at
        // row gnr (idamax then restarts after it), the after-NaN candidate of
        // the slab holding gnr plus the best of every slab wholly below it.
        double bv = 0.0;
        int bw = -1, brow = -1, bo = -1, bslot = 0;
        for (unsigned w = lane; w < P; w += 64) {
          const double* h = cm.hdr + ((long)par * P + w) * HDR;
          double cv = 0.0;
          int cr = -1, co = -1, cs = 0;
          if (gnr < 0.0) {
            cv = (w == lane) ? f0 : ld_d(&h[0]);
            cr = (int)((w == lane) ? f1 : ld_d(&h[1]));
            co = (int)((w == lane) ? f2 : ld_d(&h[2]));
          } else {
            const int hn = (int)ld_d(&h[3]);
            if (hn == (int)gnr) {
              cv = ld_d(&h[5]);
              cr = (int)ld_d(&h[6]);
              co = (int)ld_d(&h[7]);
              cs = 2;
            } else if ((long)w * ROWS_PER_WG > (long)gnr) {
              cv = ld_d(&h[0]);
              cr = (int)ld_d(&h[1]);
              co = (int)ld_d(&h[2]);
            }
          }
          if (cr >= 0 && (brow < 0 || cv > bv || (cv == bv && cr < brow))) {
            bv = cv;
            brow = cr;
            bo = co;
            bw = (int)w;
            bslot = cs;
          }
        }
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(bv, off);
          const int orow = __shfl_xor(brow, off), oo = __shfl_xor(bo, off),
                    ow = __shfl_xor(bw, off), os = __shfl_xor(bslot, off);
          if (orow >= 0 && (brow < 0 || ov > bv || (ov == bv && orow < brow))) {
            bv = ov;
            brow = orow;
            bo = oo;
            bw = ow;
            bslot = os;
          }
        }
        if (brow < 0 && gnr >= 0.0) {  // nothing after the last NaN: it is the pivot
          brow = (int)gnr;
          bo = (int)gno;
          bw = brow / ROWS_PER_WG;
          bslot = 1;
        }
        // the winning row and the old diagonal row (from its owner's slot)
        const double* src = cm.rows + (((long)par * P + bw) * 3 + bslot) * PW;
        for (int e = lane; e < W; e += 64) {
          prow[e] = ld_d(&src[e]);
          jrow[e] = ld_d(&jslot[e]);
        }
        if (lane == 0) {
          sel[0] = brow;
          sel[1] = bo;
          sel[2] = (int)ld_d(&jslot[PW]);
        }
      }
    }
    __syncthreads();
    if (sel[0] == -2) return;
    const int r = sel[0];
    const int origj = sel[2];

    const double piv_val = prow[j];
    const bool null_pivot = (piv_val == 0.0);
    if (wg == 0 && t == 0) {
      piv[k + j] = r - j;
      if (null_pivot) atomicCAS(info, -1, k + j);
    }
    // ---- interchange rows j and r (skipped for a null pivot) ----
    if (!null_pivot && r != j) {
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int rr = base + rh * 128 + 64 * q + lane;
        if (!__any(rr == j || rr == r)) continue;  // this wave holds neither row
        if (rr == j) {
#pragma unroll
          for (int c = 0; c < 32; ++c) v[q][c] = prow[32 * cg + c];
          orig[q] = sel[1];
        } else if (rr == r) {
#pragma unroll
          for (int c = 0; c < 32; ++c) v[q][c] = jrow[32 * cg + c];
          orig[q] = origj;
        }
      }
    }
    // ---- scale column j below the diagonal (by the reciprocal, skipped when
    // the pivot is 1, as bl1_dinvscalv), then the rank-1 update ----
    const bool scale = !null_pivot && piv_val != 1.0;
    const double inv = scale ? 1.0 / piv_val : 1.0;
    if (cg == cgj) {
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int lr = rh * 128 + 64 * q + lane;
        const int rr = base + lr;
        double x = pick(v[q], cj);
        if (rr > j && rr < M) {
          if (scale) x = x * inv;
#pragma unroll
          for (int c = 0; c < 32; ++c)
            if (c == cj) v[q][c] = x;
        } else {
          x = 0.0;
        }
        lv[lr] = x;
      }
    }
    __syncthreads();
    if (32 * cg + 31 > j) {
      // Branch-free (results selected), so the prow loads are batched.
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int lr = rh * 128 + 64 * q + lane;
        const int rr = base + lr;
        const bool act = rr > j && rr < M;
        const double l = lv[lr];
#pragma unroll
        for (int c = 0; c < 32; ++c) {
          const int col = 32 * cg + c;
          const double u = v[q][c] - l * prow[col];
          v[q][c] = (act && col > j && col < W) ? u : v[q][c];
        }
      }
    }
    __syncthreads();
  }

  // ---- write back; record rows whose content moved ----
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int r = base + rh * 128 + 64 * q + lane;
    if (r < M) {
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        const int col = 32 * cg + c;
        if (col < W) A[(long)r + (long)col * lda] = v[q][c];
      }
      if (cg == 0 && orig[q] != r) {
        const int slot = atomicAdd(&moved[0], 1);
        moved[1 + 2 * slot] = r;
        moved[2 + 2 * slot] = orig[q];
      }
    }
  }
}

// Applies the panel's row moves to columns [0, ncols) of the (M-row) block
// at A: new row q := old row orig, through registers (all reads of a
// workgroup's columns precede its writes).
constexpr int GC = 16;  // columns per workgroup
template <typename T>
__global__ __launch_bounds__(256) void lu_row_gather_kernel(int ncols, T* __restrict__ A, long lda,
                                                            const int* __restrict__ moved) {
  const int cnt = moved[0];
  const int c0 = blockIdx.x * GC;
  T val[GC];
  const int t = threadIdx.x;
  // cnt <= 2 * PW = 256: one pair per thread.
  int dst = -1, src = -1;
  if (t < cnt) {
    dst = moved[1 + 2 * t];
    src = moved[2 + 2 * t];
  }
  if (src >= 0) {
#pragma unroll
    for (int c = 0; c < GC; ++c)
      val[c] = (c0 + c < ncols) ? A[(long)src + (long)(c0 + c) * lda] : T(0);
  }
  __syncthreads();
  if (dst >= 0) {
#pragma unroll
    for (int c = 0; c < GC; ++c)
      if (c0 + c < ncols) A[(long)dst + (long)(c0 + c) * lda] = val[c];
  }
}

// dst(j, l) := src(l, j) for an r x c block (32 x 32 tiles through LDS).
__global__ __launch_bounds__(256) void transpose_kernel(int r, int c, const double* __restrict__ src,
                                                        long lds, double* __restrict__ dst,
                                                        long ldd) {
  __shared__ double tile[32][33];
  const int l0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int q = ty; q < 32; q += 8) {
    const int l = l0 + tx, j = j0 + q;
    tile[q][tx] = (l < r && j < c) ? src[(long)l + (long)j * lds] : 0.0;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int j = j0 + tx, l = l0 + q;
    if (l < r && j < c) dst[(long)j + (long)l * ldd] = tile[tx][q];
  }
}

__global__ void lu_reset_kernel(int* seq, int nseq, int* moved) {
  for (int i = threadIdx.x; i < nseq; i += blockDim.x) seq[i] = 0;
  if (threadIdx.x == 0) moved[0] = 0;
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

}  // namespace

// Buffers for one panel launch (two sets, alternated between steps).
struct LuBufs {
  PanelComm cm;
  int* moved;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t lu_set_bytes(int Pmax) {
  return align256(sizeof(double) * 2 * Pmax * HDR) + align256(sizeof(double) * 2 * Pmax * 3 * PW) +
         align256(sizeof(double) * 2 * (PW + 1)) + align256(sizeof(int) * 2 * Pmax) +
         align256(sizeof(int)) + align256(sizeof(int) * (1 + 4 * PW));
}

static LuBufs lu_carve(char* p, int Pmax) {
  LuBufs b;
  b.cm.hdr = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 2 * Pmax * HDR);
  b.cm.rows = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 2 * Pmax * 3 * PW);
  b.cm.jrow = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * 2 * (PW + 1));
  b.cm.seq = reinterpret_cast<int*>(p);
  p += align256(sizeof(int) * 2 * Pmax);
  b.cm.err = reinterpret_cast<int*>(p);
  p += align256(sizeof(int));
  b.moved = reinterpret_cast<int*>(p);
  return b;
}

static int lu_panel(hipStream_t s, int M, int W, double* Ap, long lda, int k, int* piv,
                    int* dinfo, const LuBufs& b) {
  hipLaunchKernelGGL(lu_reset_kernel, dim3(1), dim3(256), 0, s, b.cm.seq, 2 * ((M + ROWS_PER_WG - 1) / ROWS_PER_WG), b.moved);
  const int P = (M + ROWS_PER_WG - 1) / ROWS_PER_WG;
  void* tok = fla_stats_begin(s);
  hipLaunchKernelGGL(lu_panel_kernel, dim3(P), dim3(NT), 0, s, M, W, Ap, lda, k, piv, dinfo,
                     b.moved, b.cm);
  double fl = 0.0;
  for (int j = 0; j < W; ++j) fl += (double)(M - j - 1) * (1.0 + 2.0 * (W - j - 1));
  fla_stats_end(s, tok, 2, fl);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

static int lu_gather(hipStream_t s, int ncols, double* A, long lda, const int* moved) {
  if (ncols <= 0) return FLA_HIP_OK;
  hipLaunchKernelGGL(lu_row_gather_kernel<double>, dim3((ncols + GC - 1) / GC), dim3(256), 0, s,
                     ncols, A, lda, moved);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

// Row interchanges, A12 solve and A22 update for columns [c0, c0+nc) after
// the panel at (k, kb).  U12 is copied transposed into Ut (ld ldu) so that
// the update's right operand is read with j contiguous, as in the Cholesky
// update (16-deep K chunks, coalesced loads).
static int lu_update_cols(hipStream_t s, int m, double* A, long lda, int k, int kb, int c0,
                          int nc, const int* moved, double* Ut, long ldu) {
  if (nc <= 0) return FLA_HIP_OK;
  double* Ak = A + (long)k;  // row k
  FLA_TRY(lu_gather(s, nc, Ak + (long)c0 * lda, lda, moved));
  double* A11 = A + (long)k + (long)k * lda;
  double* A12 = A + (long)k + (long)c0 * lda;
  FLA_TRY(launch_trsm_llnu(s, nc, kb, A11, lda, A12, lda));
  const int mr = m - k - kb;
  if (mr > 0) {
    hipLaunchKernelGGL(transpose_kernel, dim3((kb + 31) / 32, (nc + 31) / 32), dim3(256), 0, s, kb,
                       nc, A12, lda, Ut, ldu);
    FLA_HIP_CHECK(hipGetLastError());
    void* tok = fla_stats_begin(s);
    FLA_TRY(launch_dgemm_mfma(s, false, true, 0, mr, nc, kb, -1.0, A11 + kb, lda, Ut, ldu, 1.0,
                              A + (long)(k + kb) + (long)c0 * lda, lda, nullptr));
    fla_stats_end(s, tok, 1, 2.0 * mr * (double)nc * kb);
  }
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dgetrf(int m, int n, double* A, int lda, int* p, int* info) {
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  *info = kSuccess;
  const int mn = m < n ? m : n;
  if (mn == 0) return FLA_HIP_OK;
  FlaHipCtx& c = fla_ctx();
  const int nb = c.lu_nb <= PW ? c.lu_nb : PW;
  const int Pmax = (m + ROWS_PER_WG - 1) / ROWS_PER_WG;
  // Every panel workgroup must be resident at once (one per CU at most).
  if (Pmax > c.num_cus || Pmax > MAX_P) {
    std::fprintf(stderr, "libflame_amd: dgetrf supports m <= %d on this device\n",
                 c.num_cus * ROWS_PER_WG);
    return FLA_HIP_EARG;
  }
  // exchange buffers in the uncached region, laid out for MAX_P workgroups
  const size_t set = align256(lu_set_bytes(MAX_P));
  static_assert(2 * (8 * 2 * MAX_P * HDR + 8 * 2 * MAX_P * 3 * PW) < kUcLuBytes, "LU exchange");
  char* sp = c.uc;
  LuBufs bufs[2] = {lu_carve(sp, MAX_P), lu_carve(sp + set, MAX_P)};
  // transposed U12 of one step (n x nb, ld n)
  void* wut = nullptr;
  FLA_TRY(fla_workspace((size_t)(n > 1 ? n : 1) * PW * sizeof(double), &wut));
  double* Ut = static_cast<double*>(wut);
  const long ldu = n > 1 ? n : 1;
  int* dinfo = c.dscratch;
  int* derr = c.dscratch + 4;  // one wait-failure flag for the whole factorization
  bufs[0].cm.err = derr;
  bufs[1].cm.err = derr;
  hipStream_t su = c.s_update, sq = c.s_panel;
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, su, dinfo, kSuccess);
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, su, derr, 0);
  int evi = 0;
  auto next_event = [&]() {
    hipEvent_t e = c.ev[evi];
    evi = (evi + 1) % FlaHipCtx::kEvents;
    return e;
  };
  hipEvent_t start = next_event();
  FLA_HIP_CHECK(hipEventRecord(start, su));
  FLA_HIP_CHECK(hipStreamWaitEvent(sq, start, 0));

  int step = 0;
  int kb = mn < nb ? mn : nb;
  FLA_TRY(lu_panel(sq, m, kb, A, lda, 0, p, dinfo, bufs[0]));
  for (int k = 0; k < mn; ++step) {
    const LuBufs& b = bufs[step & 1];
    hipEvent_t pd = next_event();
    FLA_HIP_CHECK(hipEventRecord(pd, sq));
    FLA_HIP_CHECK(hipStreamWaitEvent(su, pd, 0));
    // Interchanges on the columns left of the panel.
    FLA_TRY(lu_gather(su, k, A + (long)k, lda, b.moved));
    const int k2 = k + kb;
    const int kb2 = (mn - k2) < nb ? (mn - k2) : nb;  // next panel width (0 at the end)
    // (a) the next panel's columns, so that panel can start
    FLA_TRY(lu_update_cols(su, m, A, lda, k, kb, k2, kb2 > 0 ? kb2 : 0, b.moved, Ut, ldu));
    hipEvent_t ready = next_event();
    FLA_HIP_CHECK(hipEventRecord(ready, su));
    // (b) the remaining columns
    FLA_TRY(lu_update_cols(su, m, A, lda, k, kb, k2 + kb2, n - k2 - kb2, b.moved, Ut + kb2, ldu));
    if (kb2 <= 0) break;
    FLA_HIP_CHECK(hipStreamWaitEvent(sq, ready, 0));
    FLA_TRY(lu_panel(sq, m - k2, kb2, A + (long)k2 + (long)k2 * lda, lda, k2, p, dinfo,
                     bufs[(step + 1) & 1]));
    k = k2;
    kb = kb2;
  }
  hipEvent_t fin = next_event();
  FLA_HIP_CHECK(hipEventRecord(fin, sq));
  FLA_HIP_CHECK(hipStreamWaitEvent(su, fin, 0));
  int h[2] = {kSuccess, 0};
  FLA_HIP_CHECK(hipMemcpyAsync(&h[0], dinfo, sizeof(int), hipMemcpyDeviceToHost, su));
  FLA_HIP_CHECK(hipMemcpyAsync(&h[1], derr, sizeof(int), hipMemcpyDeviceToHost, su));
  FLA_HIP_CHECK(hipStreamSynchronize(su));
  fla_stats_harvest();
  if (h[1] != 0) return FLA_HIP_ESPIN;
  *info = h[0];
  return FLA_HIP_OK;
}

// ---- building blocks of the reduced-precision LU (fla_hip_mixed.hip) ----------

static LuBufs g_mp_bufs[2];

int fla_lu_panel_setup(int m, int* derr) {
  FlaHipCtx& c = fla_ctx();
  const int Pmax = (m + ROWS_PER_WG - 1) / ROWS_PER_WG;
  if (Pmax > c.num_cus || Pmax > MAX_P) {
    std::fprintf(stderr, "libflame_amd: the LU panel supports m <= %d on this device\n",
                 c.num_cus * ROWS_PER_WG);
    return FLA_HIP_EARG;
  }
  const size_t set = align256(lu_set_bytes(MAX_P));
  g_mp_bufs[0] = lu_carve(c.uc, MAX_P);
  g_mp_bufs[1] = lu_carve(c.uc + set, MAX_P);
  g_mp_bufs[0].cm.err = derr;
  g_mp_bufs[1].cm.err = derr;
  return FLA_HIP_OK;
}

int fla_lu_panel_run(hipStream_t s, int set, int M, int W, double* Ap, long lda, int k, int* piv,
                     int* dinfo) {
  if (W > PW) return FLA_HIP_EARG;
  return lu_panel(s, M, W, Ap, lda, k, piv, dinfo, g_mp_bufs[set & 1]);
}

int fla_lu_gather_rows(hipStream_t s, int set, int ncols, double* A, long lda) {
  return lu_gather(s, ncols, A, lda, g_mp_bufs[set & 1].moved);
}

int fla_lu_gather_rows_f(hipStream_t s, int set, int ncols, float* A, long lda) {
  if (ncols <= 0) return FLA_HIP_OK;
  hipLaunchKernelGGL(lu_row_gather_kernel<float>, dim3((ncols + GC - 1) / GC), dim3(256), 0, s,
                     ncols, A, lda, g_mp_bufs[set & 1].moved);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}
