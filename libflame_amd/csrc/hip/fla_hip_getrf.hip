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
// scaling are skipped, the update uses the unchanged row j, and the sweep
// continues.
//
// Panel: one cooperative kernel.  The (m-k) x nb panel is spread over P
// workgroups (256 rows each) and held in VGPRs for the whole panel; every
// column costs one all-to-all exchange: each workgroup publishes its local
// pivot candidates (value, row, original row) and their rows, the owner of
// the diagonal row publishes that row, and every workgroup reads all
// candidates, picks the same pivot, and reads the two rows it needs.
// The exchange words live in uncached device memory (no L2 write-back or
// invalidate across the XCDs) and carry their own tags, so no store needs
// to be acknowledged before another one and no separate flag is polled:
// each 8-byte word holds a 32-bit payload and the 32-bit tag of the
// (launch, column) it belongs to, and a reader takes a word only once its
// tag is current.  Every wait is bounded.
#include <hip/hip_runtime.h>

#include "fla_hip_internal.h"

namespace {

constexpr int PW = 128;           // maximum panel width
constexpr int ROWS_PER_WG = 256;  // rows of the panel held by one workgroup
constexpr int NT = 512;  // 8 waves: wave w owns columns [32(w>>1), +32) of row half w&1
constexpr int RQ = 2;    // rows per thread
constexpr unsigned long long kSpinTicks = 200000000ULL;  // 2 s of the 100 MHz wall clock
constexpr int MAX_P = 256;  // panel workgroups (one per CU at most)
constexpr int HW = 8;       // header entries per workgroup slot (one 128-byte line)

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) u64x2 g_u64x2;
typedef __attribute__((address_space(1))) unsigned long long g_u64;

// Exchange area (tagged 16-byte entries, two parities: column j uses j & 1).
//   hdr [2][P][HW]: {best |v| over non-NaN rows >= j}, {its row, its orig},
//                   {last NaN row, its orig}, {best |v| over non-NaN rows
//                   after the last NaN}, {its row, its orig}   (row -1: none),
//                   {orig of row j} (workgroup 0), two pad entries
//   rows[2][P][3][PW]: the best, the last-NaN and the after-NaN row
//   jrow[2][PW]: the diagonal row's values
// Every store covers whole 128-byte lines (a wave's lanes write consecutive
// entries): a partial-line write to uncached memory is a read-modify-write.
struct PanelComm {
  u64x2* hdr;
  u64x2* rows;
  u64x2* jrow;
  int* err;  // wait-limit flag
};

// Stores need no ordering among themselves (each word carries its tag): a
// plain 16-byte store.  Loads are relaxed atomics of the two 8-byte words, so
// a polling loop re-reads memory without waiting after every load.
__device__ __forceinline__ void put(u64x2* p, unsigned lo, unsigned hi, unsigned tag) {
  u64x2 w;
  w.x = (unsigned long long)lo | ((unsigned long long)tag << 32);
  w.y = (unsigned long long)hi | ((unsigned long long)tag << 32);
  *(g_u64x2*)p = w;
}
__device__ __forceinline__ void put_d(u64x2* p, double d, unsigned tag) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  put(p, (unsigned)b, (unsigned)(b >> 32), tag);
}
__device__ __forceinline__ u64x2 get(const u64x2* p) {
  g_u64* q = (g_u64*)p;
  u64x2 w;
  w.x = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  w.y = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return w;
}
__device__ __forceinline__ bool tagged(u64x2 w, unsigned tag) {
  return (unsigned)(w.x >> 32) == tag && (unsigned)(w.y >> 32) == tag;
}
__device__ __forceinline__ double as_d(u64x2 w) {
  return __longlong_as_double((long long)((w.x & 0xffffffffull) | (w.y << 32)));
}
__device__ __forceinline__ int lo_i(u64x2 w) { return (int)(unsigned)w.x; }
__device__ __forceinline__ int hi_i(u64x2 w) { return (int)(unsigned)w.y; }

// Wave reductions with DPP: prefix within each row of 16 lanes (row_shr 1,
// 2, 4, 8), then across rows (row_bcast 15, 31); lane 63 ends with the
// result.  Lanes whose source is out of range keep their own value.
template <int CTRL, int ROWM>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, ROWM, 0xf, false);
}
template <int CTRL, int ROWM>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_i<CTRL, ROWM>((int)b), hi = dpp_i<CTRL, ROWM>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_max_d(double x) {
  x = fmax(x, dpp_d<0x111, 0xf>(x));
  x = fmax(x, dpp_d<0x112, 0xf>(x));
  x = fmax(x, dpp_d<0x114, 0xf>(x));
  x = fmax(x, dpp_d<0x118, 0xf>(x));
  x = fmax(x, dpp_d<0x142, 0xa>(x));
  x = fmax(x, dpp_d<0x143, 0xc>(x));
  return readlane_d(x, 63);
}
__device__ __forceinline__ int wave_max_i(int x) {
  x = max(x, dpp_i<0x111, 0xf>(x));
  x = max(x, dpp_i<0x112, 0xf>(x));
  x = max(x, dpp_i<0x114, 0xf>(x));
  x = max(x, dpp_i<0x118, 0xf>(x));
  x = max(x, dpp_i<0x142, 0xa>(x));
  x = max(x, dpp_i<0x143, 0xc>(x));
  return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ int wave_min_i(int x) {
  x = min(x, dpp_i<0x111, 0xf>(x));
  x = min(x, dpp_i<0x112, 0xf>(x));
  x = min(x, dpp_i<0x114, 0xf>(x));
  x = min(x, dpp_i<0x118, 0xf>(x));
  x = min(x, dpp_i<0x142, 0xa>(x));
  x = min(x, dpp_i<0x143, 0xc>(x));
  return __builtin_amdgcn_readlane(x, 63);
}

// First maximum over the wave of key[i] (the lane's rows lane + 64 i, i =
// 0..3, so smaller i and then smaller lane means the smaller row); keys below
// zero take no part.  Returns the local row, or -1, and the maximum.
__device__ __forceinline__ int first_max(const double (&key)[4], double& best) {
  const double m = wave_max_d(fmax(fmax(key[0], key[1]), fmax(key[2], key[3])));
  best = m;
  if (m < 0.0) return -1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const unsigned long long b = __ballot(key[i] == m);
    if (b) return 64 * i + __builtin_ctzll(b);
  }
  return -1;
}

__global__ __launch_bounds__(NT) void lu_panel_kernel(int M, int W, double* __restrict__ A,
                                                      long lda, int k, int* __restrict__ piv,
                                                      int* __restrict__ info,
                                                      int* __restrict__ moved, PanelComm cm,
                                                      unsigned tagbase) {
  // A wait that expired in an earlier panel of this factorization aborts all later ones.
  if (__hip_atomic_load(cm.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  const int P = gridDim.x;
  const int wg = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int cg = wave >> 1;  // column group (its two waves sit on different SIMDs)
  const int rh = wave & 1;   // row half
  const int base = wg * ROWS_PER_WG;

  __shared__ double colbuf[2][ROWS_PER_WG];  // column j of every row, before its interchange
  __shared__ int origs[ROWS_PER_WG];         // original row of each local row
  __shared__ double prow[PW];                // pivot row values
  __shared__ double jrow[PW];                // old diagonal row values
  __shared__ int sel[4];                     // pivot row, its orig, orig of the old row j
  // The panel's pivots and its first null pivot, written to global memory
  // after the column loop: a global store inside the loop would hold the
  // storing wave at the next column's wait for outstanding memory operations
  // until the store is acknowledged, and every workgroup waits on workgroup 0.
  __shared__ int pivs[PW];
  __shared__ int first_null;
  __shared__ double stage[NT / 64][32];      // a row's 32 columns, per wave, to store coalesced

  if (wg == 0 && t == 0) {
    // The move list is appended to only after every workgroup has read this
    // workgroup's first header, which is stored after this reset completes.
    __hip_atomic_store(&moved[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    first_null = -1;
  }

  // Thread owns local rows rh*128 + 64*q + lane (q = 0..1), columns 32*cg + c.
  double v[RQ][32];
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int r = base + rh * 128 + 64 * q + lane;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      const int col = 32 * cg + c;
      v[q][c] = (r < M && col < W) ? A[(long)r + (long)col * lda] : 0.0;
    }
    if (cg == 0) {
      colbuf[0][rh * 128 + 64 * q + lane] = v[q][0];
      origs[rh * 128 + 64 * q + lane] = r;
    }
  }
  // All panel loads complete here, so that inside the column loop no use of
  // the panel registers waits on memory counters (which also count the
  // exchange stores in flight).
  __builtin_amdgcn_s_waitcnt(0);

  for (int j = 0; j < W; ++j) {
    const int par = j & 1;
    const int cgj = j >> 5, cj = j & 31;
    const unsigned tag = tagbase | (unsigned)(j + 1);
    __syncthreads();  // column j and the row origins are in LDS

    // ---- local candidates, found by every wave from the LDS copy ----
    double key[4];
    unsigned long long nanb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int lr = lane + 64 * i;
      const int r = base + lr;
      const double x = colbuf[par][lr];
      const bool act = r >= j && r < M;
      key[i] = (act && x == x) ? fabs(x) : -1.0;
      nanb[i] = __ballot(act && x != x);
    }
    double bv, av = -1.0;
    const int lb = first_max(key, bv);
    int ln = -1;
#pragma unroll
    for (int i = 3; i >= 0; --i)
      if (ln < 0 && nanb[i]) ln = 64 * i + 63 - __builtin_clzll(nanb[i]);
    int la = -1;
    if (ln >= 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (lane + 64 * i <= ln) key[i] = -1.0;
      la = first_max(key, av);
    }

    // ---- publish: the header (wave 0), each candidate row (its owners) ----
    u64x2* myrows = cm.rows + ((long)par * P + wg) * 3 * PW;
    u64x2* js = cm.jrow + (long)par * PW;
    if (wave == 0 && lane < HW) {
      u64x2* h = cm.hdr + ((long)par * P + wg) * HW;
      const int br = lb >= 0 ? base + lb : -1, bo = lb >= 0 ? origs[lb] : -1;
      const int nr = ln >= 0 ? base + ln : -1, no = ln >= 0 ? origs[ln] : -1;
      const int ar = la >= 0 ? base + la : -1, ao = la >= 0 ? origs[la] : -1;
      const int jo = (wg == 0) ? origs[j] : -1;
      const double d = lane == 0 ? bv : av;
      const unsigned long long db = (unsigned long long)__double_as_longlong(d);
      unsigned lo = 0, hi = 0;
      if (lane == 0 || lane == 3) {
        lo = (unsigned)db;
        hi = (unsigned)(db >> 32);
      } else if (lane == 1) {
        lo = (unsigned)br;
        hi = (unsigned)bo;
      } else if (lane == 2) {
        lo = (unsigned)nr;
        hi = (unsigned)no;
      } else if (lane == 4) {
        lo = (unsigned)ar;
        hi = (unsigned)ao;
      } else if (lane == 5) {
        lo = (unsigned)jo;
      }
      put(h + lane, lo, hi, tag);
    }
    {
      // rows lb, ln, la of this workgroup, and row j in workgroup 0: the
      // owner lane stages its 32 columns, then 32 lanes store them
      const int jl = (wg == 0) ? j : -1;
      const int want[4] = {lb, ln, la, jl};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int lr = want[s];
        if (lr < 0 || (lr >> 7) != rh) continue;  // uniform
        u64x2* dst = (s < 3) ? myrows + s * PW : js;
        if (lane == (lr & 63)) {
          if (((lr >> 6) & 1) == 0) {
#pragma unroll
            for (int c = 0; c < 32; ++c) stage[wave][c] = v[0][c];
          } else {
#pragma unroll
            for (int c = 0; c < 32; ++c) stage[wave][c] = v[1][c];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < 32 && 32 * cg + lane < W) put_d(&dst[32 * cg + lane], stage[wave][lane], tag);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }

    // ---- wave 0: read every header, pick the pivot, read its row and row j ----
    if (wave == 0) {
      bool ok = true;
      const unsigned long long t0 = wall_clock64();
      auto expired = [&]() {
        if (__hip_atomic_load(cm.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            wall_clock64() - t0 > kSpinTicks) {
          if (lane == 0) __hip_atomic_store(cm.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          return true;
        }
        return false;
      };
      const u64x2* hb = cm.hdr + (long)par * P * HW;
      // pass 1: best of every slab, and the last NaN row overall
      double gv = 0.0;
      int grow = -1, go = -1, gw = -1, gslot = 0, gnr = -1, gno = -1, jo = 0;
      for (int w0 = 0; ok && w0 < P; w0 += 64) {
        const int w = w0 + lane;
        u64x2 e0, e1, e2, e3, e4, e5;
        while (true) {
          bool done = true;
          if (w < P) {
            const u64x2* h = hb + (long)w * HW;
            e0 = get(h);
            e1 = get(h + 1);
            e2 = get(h + 2);
            e3 = get(h + 3);
            e4 = get(h + 4);
            e5 = get(h + 5);
            done = tagged(e0, tag) && tagged(e1, tag) && tagged(e2, tag) && tagged(e3, tag) &&
                   tagged(e4, tag) && tagged(e5, tag);
          }
          if (__all(done)) break;
          if (expired()) {
            ok = false;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (ok && w == 0) jo = lo_i(e5);
        if (ok && w < P) {
          // a lane visits its slabs in row order: keep the first maximum
          const double cv = as_d(e0);
          const int cr = lo_i(e1), co = hi_i(e1), hn = lo_i(e2), hno = hi_i(e2);
          if (hn > gnr) {
            gnr = hn;
            gno = hno;
          }
          if (cr >= 0 && (grow < 0 || cv > gv)) {
            gv = cv;
            grow = cr;
            go = co;
            gw = w;
          }
        }
      }
      if (ok) {
        const unsigned long long nan_lanes = __ballot(gnr >= 0);
        if (nan_lanes == 0) {
          // first maximum over the lanes: the wave's maximum, then the
          // smallest row among the lanes that hold it
          const double m = wave_max_d(grow >= 0 ? gv : -1.0);
          const unsigned long long b = __ballot(grow >= 0 && gv == m);
          int L = b ? __builtin_ctzll(b) : 0;
          if (__builtin_popcountll(b) > 1) {
            const bool in = (b >> lane) & 1ull;
            const int mr = wave_min_i(in ? grow : 0x7fffffff);
            L = __builtin_ctzll(__ballot(in && grow == mr));
          }
          grow = __builtin_amdgcn_readlane(grow, L);
          go = __builtin_amdgcn_readlane(go, L);
          gw = __builtin_amdgcn_readlane(gw, L);
        } else {
          // the last NaN row overall, and its orig from the lane holding it
          const int m = wave_max_i(gnr);
          gno = __builtin_amdgcn_readlane(gno, __builtin_ctzll(__ballot(gnr == m)));
          gnr = m;
          // With the last NaN at row gnr (idamax restarts after it): the
          // after-NaN candidate of the slab holding gnr plus the best of
          // every slab wholly below it.  All headers are current by now.
          gv = 0.0;
          grow = go = gw = -1;
          for (int w = lane; w < P; w += 64) {
            const u64x2* h = hb + (long)w * HW;
            double cv = 0.0;
            int cr = -1, co = -1, cs = 0;
            if (lo_i(get(h + 2)) == gnr) {
              const u64x2 e3 = get(h + 3), e4 = get(h + 4);
              cv = as_d(e3);
              cr = lo_i(e4);
              co = hi_i(e4);
              cs = 2;
            } else if ((long)w * ROWS_PER_WG > (long)gnr) {
              const u64x2 e0 = get(h), e1 = get(h + 1);
              cv = as_d(e0);
              cr = lo_i(e1);
              co = hi_i(e1);
            }
            if (cr >= 0 && (grow < 0 || cv > gv || (cv == gv && cr < grow))) {
              gv = cv;
              grow = cr;
              go = co;
              gw = w;
              gslot = cs;
            }
          }
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(gv, off);
            const int orow = __shfl_xor(grow, off), oo = __shfl_xor(go, off),
                      ow = __shfl_xor(gw, off), os = __shfl_xor(gslot, off);
            if (orow >= 0 && (grow < 0 || ov > gv || (ov == gv && orow < grow))) {
              gv = ov;
              grow = orow;
              go = oo;
              gw = ow;
              gslot = os;
            }
          }
        }
        if (grow < 0 && gnr >= 0) {  // nothing after the last NaN: it is the pivot
          grow = gnr;
          go = gno;
          gw = gnr / ROWS_PER_WG;
          gslot = 1;
        }
        if (gw < 0) gw = 0;  // unreachable: row j itself takes part (W <= M)
        // the winning row and the old diagonal row
        const u64x2* src = cm.rows + (((long)par * P + gw) * 3 + gslot) * PW;
        while (true) {
          bool done = true;
          for (int e = lane; e < W; e += 64) {
            const u64x2 a = get(&src[e]), b = get(&js[e]);
            if (tagged(a, tag) && tagged(b, tag)) {
              prow[e] = as_d(a);
              jrow[e] = as_d(b);
            } else {
              done = false;
            }
          }
          if (__all(done)) break;
          if (expired()) {
            ok = false;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (lane == 0) {
          sel[0] = grow;
          sel[1] = go;
          sel[2] = jo;
        }
      }
      if (!ok && lane == 0) sel[0] = -2;
    }
    __syncthreads();
    if (sel[0] == -2) return;
    const int r = sel[0];
    const double piv_val = prow[j];
    const bool null_pivot = (piv_val == 0.0);
    const bool swap = !null_pivot && r != j;
    if (wg == 0 && t == 0) {
      pivs[j] = r - j;
      if (null_pivot && first_null < 0) first_null = j;
    }
    // ---- interchange rows j and r (skipped for a null pivot) ----
    if (swap) {
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int rr = base + rh * 128 + 64 * q + lane;
        if (!__any(rr == j || rr == r)) continue;  // this wave holds neither row
        if (rr == j) {
#pragma unroll
          for (int c = 0; c < 32; ++c) v[q][c] = prow[32 * cg + c];
        } else if (rr == r) {
#pragma unroll
          for (int c = 0; c < 32; ++c) v[q][c] = jrow[32 * cg + c];
        }
      }
      if (t == 0) {
        if (j >= base && j < base + ROWS_PER_WG) origs[j - base] = sel[1];
        if (r >= base && r < base + ROWS_PER_WG) origs[r - base] = sel[2];
      }
    }
    // ---- multipliers (column j scaled by the reciprocal, skipped for a null
    // pivot or a pivot of 1, as bl1_dinvscalv), then the rank-1 update with
    // the diagonal row (the old row j when there was no interchange) ----
    const bool scale = !null_pivot && piv_val != 1.0;
    const double inv = scale ? 1.0 / piv_val : 1.0;
    const double* u = swap ? prow : jrow;
    const int jn = j + 1;
    if (cg >= cgj) {
      double l[RQ], xn[RQ] = {0.0, 0.0};
      bool act[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int lr = rh * 128 + 64 * q + lane;
        const int rr = base + lr;
        act[q] = rr > j && rr < M;
        double x = colbuf[par][lr];
        if (swap && rr == r) x = jrow[j];
        l[q] = scale ? x * inv : x;
      }
      // 8-column chunks [c0, c0 + 8): wholly right of j without selects; the
      // chunk holding j (only in the waves owning it) with them.  The chunk's
      // pivot-row values are read once for both rows of the thread.
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const int c0 = 32 * cg + 8 * cc;
        if (c0 + 7 < j) continue;  // uniform
        double uu[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) uu[c] = u[c0 + c];
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          double* a = &v[q][8 * cc];
          if (c0 > j) {
            if (act[q]) {
#pragma unroll
              for (int c = 0; c < 8; ++c) a[c] = fma(-l[q], uu[c], a[c]);
            }
          } else if (act[q]) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
              const double y = fma(-l[q], uu[c], a[c]);
              a[c] = (c0 + c > j) ? y : ((c0 + c == j) ? l[q] : a[c]);
            }
          }
          if (jn < W && c0 <= jn && jn < c0 + 8) {  // uniform: the next column's chunk
#pragma unroll
            for (int c = 0; c < 8; ++c) xn[q] = (c0 + c == jn) ? a[c] : xn[q];
          }
        }
      }
      if (jn < W && cg == (jn >> 5)) {
#pragma unroll
        for (int q = 0; q < RQ; ++q) colbuf[par ^ 1][rh * 128 + 64 * q + lane] = xn[q];
      }
    }
  }
  __syncthreads();

  if (wg == 0) {
    for (int jj = t; jj < W; jj += NT) piv[k + jj] = pivs[jj];
    if (t == 0 && first_null >= 0) atomicCAS(info, -1, k + first_null);
  }
  // ---- write back; record rows whose content moved ----
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int lr = rh * 128 + 64 * q + lane;
    const int r = base + lr;
    if (r < M) {
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        const int col = 32 * cg + c;
        if (col < W) A[(long)r + (long)col * lda] = v[q][c];
      }
      if (cg == 0 && origs[lr] != r) {
        const int slot = atomicAdd(&moved[0], 1);
        moved[1 + 2 * slot] = r;
        moved[2 + 2 * slot] = origs[lr];
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

__global__ void set_int_kernel(int* p, int v) { *p = v; }

}  // namespace

// Exchange area of the panel kernel (one, shared by every launch: the tags
// keep launches apart) and four move lists: a panel must not reset a list
// that row gathers still queued on the update stream read (two per 256-column
// block, alternated between blocks; the 128-column path alternates two).
struct LuBufs {
  PanelComm cm;
  int* moved;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr size_t kHdrBytes = sizeof(u64x2) * 2 * MAX_P * HW;
constexpr size_t kRowsBytes = sizeof(u64x2) * 2 * MAX_P * 3 * PW;
constexpr size_t kJrowBytes = sizeof(u64x2) * 2 * PW;
constexpr size_t kMovedBytes = sizeof(int) * (1 + 4 * PW);
static_assert(kHdrBytes + kRowsBytes + kJrowBytes + 256 * 8 + 4 * kMovedBytes < kUcLuBytes,
              "LU exchange area");

static void lu_carve(char* p, LuBufs (&b)[4]) {
  PanelComm cm;
  cm.hdr = reinterpret_cast<u64x2*>(p);
  p += align256(kHdrBytes);
  cm.rows = reinterpret_cast<u64x2*>(p);
  p += align256(kRowsBytes);
  cm.jrow = reinterpret_cast<u64x2*>(p);
  p += align256(kJrowBytes);
  cm.err = nullptr;  // set by the caller
  for (int s = 0; s < 4; ++s) {
    b[s].cm = cm;
    b[s].moved = reinterpret_cast<int*>(p);
    p += align256(kMovedBytes);
  }
}

static __global__ void zero_kernel(unsigned long long* p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Launch counter: its low 24 bits tag every exchange word of a launch.  When
// it wraps, the exchange area is cleared, so no word written 2^24 launches
// earlier can carry a current tag.
static unsigned g_lu_epoch = 0;

static int lu_panel(hipStream_t s, int M, int W, double* Ap, long lda, int k, int* piv,
                    int* dinfo, const LuBufs& b) {
  const int P = (M + ROWS_PER_WG - 1) / ROWS_PER_WG;
  g_lu_epoch = (g_lu_epoch + 1) & 0xffffffu;
  if (g_lu_epoch == 0) {
    const long words = (long)(reinterpret_cast<char*>(b.cm.jrow) + kJrowBytes -
                              reinterpret_cast<char*>(b.cm.hdr)) / 8;
    hipLaunchKernelGGL(zero_kernel, dim3(256), dim3(256), 0, s,
                       reinterpret_cast<unsigned long long*>(b.cm.hdr), words);
    FLA_HIP_CHECK(hipGetLastError());
    g_lu_epoch = 1;
  }
  void* tok = fla_stats_begin(s);
  hipLaunchKernelGGL(lu_panel_kernel, dim3(P), dim3(NT), 0, s, M, W, Ap, lda, k, piv, dinfo,
                     b.moved, b.cm, g_lu_epoch << 8);
  double fl = 0.0;
  for (int j = 0; j < W; ++j) fl += (double)(M - j - 1) * (1.0 + 2.0 * (W - j - 1));
  fla_stats_end(s, tok, 2, fl);
  FLA_HIP_CHECK(hipGetLastError());
  return FLA_HIP_OK;
}

static int lu_gather(hipStream_t s, int ncols, double* A, long lda, const int* moved) {
  if (ncols <= 0) return FLA_HIP_OK;
  // measurement kind 6: time of the interchange kernels (the moved bytes
  // depend on the pivots and are counted by the caller from them)
  void* tok = fla_stats_begin(s);
  hipLaunchKernelGGL(lu_row_gather_kernel<double>, dim3((ncols + GC - 1) / GC), dim3(256), 0, s,
                     ncols, A, lda, moved);
  fla_stats_end(s, tok, 6, 0.0);
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

// Blocks of one panel of width nb <= 128 (look-ahead over panels): the
// next panel's columns first, then the remaining columns.
static int lu_blocked1(FlaHipCtx& c, int m, int n, double* A, long lda, int* p, int* dinfo,
                       LuBufs (&bufs)[4], double* Ut, long ldu) {
  const int mn = m < n ? m : n;
  const int nb = c.lu_nb <= PW ? c.lu_nb : PW;
  hipStream_t su = c.s_update, sq = c.s_panel;
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
    const int k2 = k + kb;
    const int kb2 = (mn - k2) < nb ? (mn - k2) : nb;  // next panel width (0 at the end)
    // (a) the next panel's columns, so that panel can start
    FLA_TRY(lu_update_cols(su, m, A, lda, k, kb, k2, kb2 > 0 ? kb2 : 0, b.moved, Ut, ldu));
    hipEvent_t ready = next_event();
    FLA_HIP_CHECK(hipEventRecord(ready, su));
    // Interchanges on the columns left of the panel (off the next panel's path).
    FLA_TRY(lu_gather(su, k, A + (long)k, lda, b.moved));
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
  return FLA_HIP_OK;
}

// The same for a 256-column block factored as two panels, a at (k, k) of
// width kba and b at (k + kba, k + kba) of width kbb: both panels'
// interchanges, U12 := inv(L11) A12 with L11 = [La 0; Lba Lb] (two 128
// solves around Bb -= Lba Xa), and one update with K = kba + kbb.
static int lu_update_cols2(hipStream_t s, int m, double* A, long lda, int k, int kba, int kbb,
                           int c0, int nc, const int* moved_a, const int* moved_b, double* Ut,
                           long ldu) {
  if (nc <= 0) return FLA_HIP_OK;
  if (kbb <= 0) return lu_update_cols(s, m, A, lda, k, kba, c0, nc, moved_a, Ut, ldu);
  const int kb = kba + kbb, kb1 = k + kba;
  FLA_TRY(lu_gather(s, nc, A + (long)k + (long)c0 * lda, lda, moved_a));
  FLA_TRY(lu_gather(s, nc, A + (long)kb1 + (long)c0 * lda, lda, moved_b));
  double* Xa = A + (long)k + (long)c0 * lda;
  double* Xb = A + (long)kb1 + (long)c0 * lda;
  FLA_TRY(launch_trsm_llnu(s, nc, kba, A + (long)k + (long)k * lda, lda, Xa, lda));
  FLA_TRY(launch_dgemm_mfma(s, false, false, 0, kbb, nc, kba, -1.0, A + (long)kb1 + (long)k * lda,
                            lda, Xa, lda, 1.0, Xb, lda, nullptr));
  FLA_TRY(launch_trsm_llnu(s, nc, kbb, A + (long)kb1 + (long)kb1 * lda, lda, Xb, lda));
  const int mr = m - k - kb;
  if (mr > 0) {
    hipLaunchKernelGGL(transpose_kernel, dim3((kb + 31) / 32, (nc + 31) / 32), dim3(256), 0, s, kb,
                       nc, Xa, lda, Ut, ldu);
    FLA_HIP_CHECK(hipGetLastError());
    void* tok = fla_stats_begin(s);
    FLA_TRY(launch_dgemm_mfma(s, false, true, 0, mr, nc, kb, -1.0, A + (long)(k + kb) + (long)k * lda,
                              lda, Ut, ldu, 1.0, A + (long)(k + kb) + (long)c0 * lda, lda, nullptr));
    fla_stats_end(s, tok, 1, 2.0 * mr * (double)nc * kb);
  }
  return FLA_HIP_OK;
}

// 256-column blocks ("pairs") of two 128-column panels A and B.  The update
// stream carries the chain: panel A; B's columns updated with A; panel B; B's
// interchanges on A's columns; the next pair's 256 columns updated with the
// pair (K = 256), after which the next panel A starts.  The bulk stream
// (s_update_lp: its CU mask leaves 64 CUs to the panel kernel) updates every
// column beyond the next pair (K = 256) and applies the pair's
// interchanges left of it, concurrently with the next pair's two panels; the
// chain waits for it only before the following pair's columns.  Ut holds the
// transposed U12 of the bulk update in rows [0, n) and of the chain's updates
// in rows [n, n + 256) (ld ldu >= n + 256).
static int lu_blocked2(FlaHipCtx& c, int m, int n, double* A, long lda, int* p, int* dinfo,
                       LuBufs (&bufs)[4], double* Ut, long ldu) {
  const int mn = m < n ? m : n;
  hipStream_t su = c.s_update, sb = c.s_update_lp, sq = c.s_panel;
  double* Utn = Ut + n;
  int evi = 0;
  auto next_event = [&]() {
    hipEvent_t e = c.ev[evi];
    evi = (evi + 1) % FlaHipCtx::kEvents;
    return e;
  };
  hipEvent_t start = next_event();
  FLA_HIP_CHECK(hipEventRecord(start, su));
  FLA_HIP_CHECK(hipStreamWaitEvent(sq, start, 0));
  FLA_HIP_CHECK(hipStreamWaitEvent(sb, start, 0));
  int blk = 0;
  int kba = mn < PW ? mn : PW;
  hipEvent_t bulk_done = nullptr;
  FLA_TRY(lu_panel(sq, m, kba, A, lda, 0, p, dinfo, bufs[0]));
  for (int k = 0;; ++blk) {
    const LuBufs& ba = bufs[2 * (blk & 1)];
    const LuBufs& bb = bufs[2 * (blk & 1) + 1];
    hipEvent_t pa = next_event();
    FLA_HIP_CHECK(hipEventRecord(pa, sq));
    FLA_HIP_CHECK(hipStreamWaitEvent(su, pa, 0));
    const int kb1 = k + kba;
    const int kbb = (mn - kb1) < PW ? (mn - kb1) : PW;
    if (kbb > 0) {
      FLA_TRY(lu_update_cols(su, m, A, lda, k, kba, kb1, kbb, ba.moved, Utn, ldu));
      hipEvent_t rb = next_event();
      FLA_HIP_CHECK(hipEventRecord(rb, su));
      FLA_HIP_CHECK(hipStreamWaitEvent(sq, rb, 0));
      FLA_TRY(lu_panel(sq, m - kb1, kbb, A + (long)kb1 + (long)kb1 * lda, lda, kb1, p, dinfo, bb));
      hipEvent_t pb = next_event();
      FLA_HIP_CHECK(hipEventRecord(pb, sq));
      FLA_HIP_CHECK(hipStreamWaitEvent(su, pb, 0));
      FLA_TRY(lu_gather(su, kba, A + (long)kb1 + (long)k * lda, lda, bb.moved));
    }
    const int k2 = kb1 + (kbb > 0 ? kbb : 0);  // the next pair
    const int nw = (mn - k2) < 2 * PW ? (mn - k2) : 2 * PW;
    if (bulk_done) FLA_HIP_CHECK(hipStreamWaitEvent(su, bulk_done, 0));
    FLA_TRY(lu_update_cols2(su, m, A, lda, k, kba, kbb, k2, nw > 0 ? nw : 0, ba.moved, bb.moved,
                            Utn, ldu));
    hipEvent_t ready = next_event();
    FLA_HIP_CHECK(hipEventRecord(ready, su));
    const int c0 = k2 + (nw > 0 ? nw : 0);
    FLA_HIP_CHECK(hipStreamWaitEvent(sb, ready, 0));
    FLA_TRY(lu_update_cols2(sb, m, A, lda, k, kba, kbb, c0, n - c0, ba.moved, bb.moved, Ut, ldu));
    FLA_TRY(lu_gather(sb, k, A + (long)k, lda, ba.moved));
    if (kbb > 0) FLA_TRY(lu_gather(sb, k, A + (long)kb1, lda, bb.moved));
    bulk_done = next_event();
    FLA_HIP_CHECK(hipEventRecord(bulk_done, sb));
    if (nw <= 0) break;
    const int kbn = nw < PW ? nw : PW;
    FLA_HIP_CHECK(hipStreamWaitEvent(sq, ready, 0));
    FLA_TRY(lu_panel(sq, m - k2, kbn, A + (long)k2 + (long)k2 * lda, lda, k2, p, dinfo,
                     bufs[2 * ((blk + 1) & 1)]));
    k = k2;
    kba = kbn;
  }
  hipEvent_t fin = next_event();
  FLA_HIP_CHECK(hipEventRecord(fin, sq));
  FLA_HIP_CHECK(hipStreamWaitEvent(su, fin, 0));
  FLA_HIP_CHECK(hipStreamWaitEvent(su, bulk_done, 0));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dgetrf(int m, int n, double* A, int lda, int* p, int* info) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLA_HIP_EARG;
  *info = kSuccess;
  const int mn = m < n ? m : n;
  if (mn == 0) return FLA_HIP_OK;
  FlaHipCtx& c = fla_ctx();
  const int Pmax = (m + ROWS_PER_WG - 1) / ROWS_PER_WG;
  // Every panel workgroup must be resident at once (one per CU at most).
  if (Pmax > c.num_cus || Pmax > MAX_P) {
    std::fprintf(stderr, "libflame_amd: dgetrf supports m <= %d on this device\n",
                 c.num_cus * ROWS_PER_WG);
    return FLA_HIP_EARG;
  }
  // exchange area in the uncached region, laid out for MAX_P workgroups
  LuBufs bufs[4];
  lu_carve(c.uc, bufs);
  // transposed U12 of a step: (n + 2 PW) x 2 PW, ld n + 2 PW
  void* wut = nullptr;
  const long ldu = (long)n + 2 * PW;
  FLA_TRY(fla_workspace((size_t)ldu * 2 * PW * sizeof(double), &wut));
  double* Ut = static_cast<double*>(wut);
  int* dinfo = c.dscratch;
  int* derr = c.dscratch + 4;  // one wait-failure flag for the whole factorization
  for (auto& b : bufs) b.cm.err = derr;
  hipStream_t su = c.s_update, sq = c.s_panel;
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, su, dinfo, kSuccess);
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, su, derr, 0);
  if (c.lu_nb > PW) {
    FLA_TRY(lu_blocked2(c, m, n, A, lda, p, dinfo, bufs, Ut, ldu));
  } else {
    FLA_TRY(lu_blocked1(c, m, n, A, lda, p, dinfo, bufs, Ut, ldu));
  }
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

static LuBufs g_mp_bufs[4];

int fla_lu_panel_setup(int m, int* derr) {
  FlaHipCtx& c = fla_ctx();
  const int Pmax = (m + ROWS_PER_WG - 1) / ROWS_PER_WG;
  if (Pmax > c.num_cus || Pmax > MAX_P) {
    std::fprintf(stderr, "libflame_amd: the LU panel supports m <= %d on this device\n",
                 c.num_cus * ROWS_PER_WG);
    return FLA_HIP_EARG;
  }
  lu_carve(c.uc, g_mp_bufs);
  for (auto& b : g_mp_bufs) b.cm.err = derr;
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
