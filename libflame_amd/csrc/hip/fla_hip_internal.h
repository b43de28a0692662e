// Internal declarations shared by the HIP translation units of libflame_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cmath>
#include <cstdio>

#include "FLA_hip.h"

// FLAME constants (same numeric values as the reference's FLA_macro_defs.h).
enum : int {
  kSuccess = -1,
  kLeft = 210,
  kRight = 211,
  kLower = 300,
  kUpper = 301,
  kFull = 311,
  kNoTrans = 400,
  kTrans = 401,
  kConjTrans = 402,
  kUnit = 500,
  kNonUnit = 501,
  kNativePivots = 700,
  kLapackPivots = 701,
};

struct FlaHipCtx {
  bool ready = false;
  int device = 0;
  int num_cus = 0;
  hipStream_t s_update = nullptr;  // trailing updates (and the API's default stream)
  hipStream_t s_panel = nullptr;   // panel factorizations (look-ahead)
  // Trailing updates of the reduced-precision factor: their panel chain, not
  // the update, is the critical path, so this stream leaves more CUs free.
  // 64 CUs left to the panels: the reduced-precision factor's updates and the
  // LU's bulk updates beside its panel chain
  hipStream_t s_update_lp = nullptr;
  static constexpr int kEvents = 8;
  hipEvent_t ev[kEvents] = {};
  // The caller's stream (fla_hip_set_stream; null: none) and the events that
  // order the library's three streams after it on entry and it after them on
  // return.
  hipStream_t user_stream = nullptr;
  hipEvent_t ev_user[4] = {};
  // Grow-only device workspaces: slot 0 general scratch (dtrsm, dlaswp),
  // slot 1 the mixed-precision solver's buffers, slot 2 its solve scratch,
  // slot 3 generators and small reductions, slot 4 the hierarchical-storage
  // copy tables.
  static constexpr int kSlots = 5;
  void* work[kSlots] = {};
  size_t work_bytes[kSlots] = {};
  // Small persistent device scratch: [0] info word, [1..] counters.
  int* dscratch = nullptr;
  // Device memory the GPU does not cache (hipDeviceMallocUncached) for data
  // that workgroups exchange inside one kernel: stores and loads of it need
  // no L2 write-back or invalidate across the XCDs.  Allocated once and kept
  // until the process ends (this kind of memory must not be freed after a
  // kernel read it).  [0, kUcLuBytes): the LU panel's exchange area and
  // move lists; then the triangular solve's copy of the finished right-hand-side
  // blocks (16 columns of n <= kUcSolveRows rows).
  char* uc = nullptr;
  // 128 x 128 FP64 block: inverse of the current Cholesky diagonal block.
  double* dlinv = nullptr;
  int chol_nb = 0;  // 0: chosen from n (fla_chol_nb)
  int mp_nb = 0;    // reduced-precision factor block; 0: fla_chol_nb
  int lu_nb = 256;  // <= 128: panel width; 256: blocks of two 128-column panels
};

constexpr size_t kUcLuBytes = 4u << 20;
constexpr long kUcSolveRows = 131072;
constexpr size_t kUcBytes = kUcLuBytes + (size_t)kUcSolveRows * 16 * sizeof(double);

FlaHipCtx& fla_ctx();
// Outer Cholesky block for order n: the configured one, else n/32 rounded
// down to a multiple of 128 and clamped to [256, 1024] (measured on MI355X:
// larger blocks raise the trailing update's arithmetic intensity).
int fla_chol_nb(int n);
int fla_ensure_ready();
// Orders a public entry point's device work after the work already queued on
// the caller's stream (on construction) and that stream's later work after
// the entry point's (on destruction), when a caller's stream is set.  Uses
// events created at initialisation only: no allocation.
struct FlaUserStreamScope {
  FlaUserStreamScope();
  ~FlaUserStreamScope();
  FlaUserStreamScope(const FlaUserStreamScope&) = delete;
  FlaUserStreamScope& operator=(const FlaUserStreamScope&) = delete;
};
#define FLA_USER_STREAM_SCOPE FlaUserStreamScope fla_user_stream_scope_

// Returns a device workspace of at least `bytes` bytes (grow-only).
int fla_workspace(size_t bytes, void** ptr);
int fla_workspace_slot(int slot, size_t bytes, void** ptr);
// Resolves the device address of the no-allocation scratch array for `device`
// (called by fla_hip_init with that device current).
int fla_noalloc_setup(int device);

#define FLA_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "libflame_amd: %s failed at %s:%d: %s\n", #expr,    \
                   __FILE__, __LINE__, hipGetErrorString(e_));                 \
      return (int)e_;                                                          \
    }                                                                          \
  } while (0)

#define FLA_TRY(expr)              \
  do {                             \
    int s_ = (expr);               \
    if (s_ != FLA_HIP_OK) return s_; \
  } while (0)

// Releases the batched path's cached graph and item buffers, and the
// measurement event pool (called by fla_hip_finalize with the device idle).
int fla_batched_release();
void fla_stats_release();

// ---- measurement hooks (no-ops unless fla_hip_stats_enable(1)) -------------
bool fla_stats_on();
void* fla_stats_begin(hipStream_t s);
void fla_stats_end(hipStream_t s, void* token, int kind, double flops);
void fla_stats_harvest();
// Algorithmic flops of a lower update of an m x n block (top n x n part on
// the diagonal) with inner dimension k: 2k per stored element.
inline double lower_update_flops(int m, int n, int k) {
  const double nn = n, mm = m;
  return 2.0 * k * ((mm - nn) * nn + nn * (nn + 1.0) / 2.0);
}

// Maps a linear workgroup id to a C tile.  Workgroups are dealt to the eight
// XCDs round-robin; the id is first permuted so that each XCD receives a
// contiguous run of tiles.  Tiles are ordered by groups of G tile columns
// (row by row inside a group), so the workgroups an XCD runs at the same time
// share both operand panels in its L2: a G x (32/G) patch of C reads G
// column panels and 32/G row panels instead of 1 and 32.
// mode 0: all Tm x Tn tiles; mode 1: the lower tiles (ti >= tj) of a
// Tm x Tn grid whose diagonal starts at the top-left (Tm >= Tn).
__host__ __device__ inline long fla_group_offset(int s, int G, int Tm, int mode) {
  // tiles before group s (all of them G columns wide)
  if (mode == 0) return (long)s * G * Tm;
  return (long)s * G * Tm - (long)G * G * s * (s - 1) / 2 - (long)s * G * (G - 1) / 2;
}

__host__ __device__ inline void fla_tile_of_linear(long t, int mode, int Tm, int Tn, int G,
                                                   int& ti, int& tj) {
  // group s: largest s with offset(s) <= t
  int s;
  if (mode == 0) {
    s = (int)(t / ((long)G * Tm));
  } else {
    const double a = 0.5 * G * G, bq = (double)G * Tm + 0.5 * G;
    const double disc = bq * bq - 4.0 * a * (double)t;
    s = (int)((bq - sqrt(disc > 0.0 ? disc : 0.0)) / (2.0 * a));
  }
  const int ng = (Tn + G - 1) / G;
  if (s < 0) s = 0;
  if (s > ng - 1) s = ng - 1;
  while (s > 0 && fla_group_offset(s, G, Tm, mode) > t) --s;
  while (s + 1 < ng && fla_group_offset(s + 1, G, Tm, mode) <= t) ++s;
  const long u = t - fla_group_offset(s, G, Tm, mode);
  const int c0 = s * G;
  const int w = (Tn - c0) < G ? (Tn - c0) : G;
  if (mode == 0) {
    ti = (int)(u / w);
    tj = c0 + (int)(u % w);
    return;
  }
  const long tri = (long)w * (w + 1) / 2;
  if (u < tri) {
    int r = (int)((sqrt(8.0 * (double)u + 1.0) - 1.0) * 0.5);
    while ((long)r * (r + 1) / 2 > u) --r;
    while ((long)(r + 1) * (r + 2) / 2 <= u) ++r;
    ti = c0 + r;
    tj = c0 + (int)(u - (long)r * (r + 1) / 2);
  } else {
    const long v = u - tri;
    ti = c0 + w + (int)(v / w);
    tj = c0 + (int)(v % w);
  }
}

__device__ __forceinline__ void fla_tile_of(int bid, int nblk, int mode, int Tm, int Tn, int G,
                                            int& ti, int& tj) {
  const int x = bid & 7;
  const int q = nblk >> 3, rr = nblk & 7;
  const long t = (long)x * q + (x < rr ? x : rr) + (bid >> 3);
  fla_tile_of_linear(t, mode, Tm, Tn, G, ti, tj);
}

// Tile-column group width of the update kernels' tile order (fla_tile_of).
extern int g_tile_group;

// ---- kernel launchers implemented in the other translation units ----------
// MFMA FP64 update: C := alpha*op(A)*op(B) + beta*C (mode 0 full, 1 lower).
int launch_dgemm_mfma(hipStream_t s, bool ta, bool tb, int mode, int m, int n, int k,
                      double alpha, const double* A, long lda, const double* B, long ldb,
                      double beta, double* C, long ldc, const int* abort_flag);
// Single-workgroup Cholesky of an nb x nb lower diagonal block (nb <= 128).
int launch_potrf_diag(hipStream_t s, int nb, double* A, long lda, int col_offset,
                      int* info);
// A21 := A21 * inv(L11)' for an m x nb strip, L11 lower non-unit (nb <= 128).
int launch_trsm_rlt(hipStream_t s, int m, int nb, const double* L11, long ldl, double* A21,
                    long lda, const int* abort_flag);
// A12 := inv(L11) * A12, L11 unit lower nb x nb (nb <= 128): one lane per column.
int launch_trsm_llnu(hipStream_t s, int ncols, int nb, const double* L11, long ldl, double* A12,
                     long lda);
// X := inv(op(T)) X for na x nrhs X (left side), W = inverted 128x128 diagonal
// blocks of T in T's orientation; one persistent launch per 16 columns.
// done/err are two device ints the call resets; *err != 0 after a bounded
// wait expired.  With ill != nullptr, block k with ill[k] != 0 is solved by
// substitution through op(T_kk) itself (unit: unit diagonal) instead of by
// its inverse.
constexpr int kNarrowRhs = 32;
int launch_trsm_left_narrow(hipStream_t s, bool op_lower, bool trans, int na, int nrhs,
                            const double* T, long ldt, const double* W, double* X, long ldx,
                            int* done, int* err, const int* ill = nullptr, bool unit = false);
// Inverts the 128x128 diagonal blocks of a triangle (lower or upper storage)
// into W, block k at W + k*128*128 (ld 128, A's orientation, other triangle 0).
// Substitution (no block inverses, the reference's dtrsm_ order): A12 :=
// inv(trilu(L11)) A12 one lane per column; A21 := A21 inv(triu(U11)) one lane
// per row (nb <= 128).
int launch_trsm_llnu_subst(hipStream_t s, int ncols, int nb, const double* L11, long ldl,
                           double* A12, long lda);
int launch_trsm_runn_rows(hipStream_t s, int nrows, int nb, const double* U11, long ldu,
                          double* A21, long lda);
// With ill != nullptr, ill[k] (one int per block) is set to 1 when block k's
// condition number exceeds the bound up to which the few-right-hand-side
// solve multiplies by the inverse, else 0.
int launch_trtri_blocks(hipStream_t s, int na, const double* A, long lda, bool upper, bool unit,
                        double* W, int* ill = nullptr);
// FP64 factor of columns [k, k+kb) of a lower matrix (rows k..n-1): diagonal
// factors, panel solves and in-panel updates, all on stream s.
int chol_panel_f64(hipStream_t s, int n, double* A, long lda, int k, int kb, int* info);
// The same with the panel solve done as a product with the explicit inverse
// of each 128 x 128 diagonal factor (fewer launches, preconditioner quality).
int chol_panel_f64_inv(hipStream_t s, int n, double* A, long lda, int k, int kb, int* info);
// LU panel building blocks (fla_hip_getrf.hip) for the reduced-precision LU:
// two alternating buffer sets for panel launches on matrices of up to m
// rows; derr is the device int set when a bounded wait expires.
int fla_lu_panel_setup(int m, int* derr);
// FP64 LU with partial pivoting of the M x W panel at Ap (W <= 128), whose
// first column is column k of the matrix: pivots into piv[k..k+W), the first
// null pivot's column into *dinfo; the row moves are kept in buffer set `set`.
int fla_lu_panel_run(hipStream_t s, int set, int M, int W, double* Ap, long lda, int k, int* piv,
                     int* dinfo);
// Applies buffer set `set`'s row moves to ncols columns of the block whose
// first row is the panel's first row (FP64 / FP32).
int fla_lu_gather_rows(hipStream_t s, int set, int ncols, double* A, long lda);
int fla_lu_gather_rows_f(hipStream_t s, int set, int ncols, float* A, long lda);
// Row interchanges of fla_hip_dlaswp on stream s, rows bounded by `rows`.
int launch_laswp(hipStream_t s, int rows, int ncols, double* A, long lda, int k1, int k2,
                 const int* p, int incp);
// In-place exchange of the strictly lower and strictly upper triangles.
int launch_transpose_swap(hipStream_t s, int n, double* A, long lda);
