// Runtime of libflame_amd: device selection, the two HIP streams that replace
// SuperMatrix's worker threads, a grow-only workspace, and host<->device copies.
//
// The reference binds one rocBLAS handle per SuperMatrix thread and device
// (FLASH_Queue_bind_hip, src/base/flamec/supermatrix/hip/main/FLASH_Queue_hip.c:297-318);
// here one process drives one GPU through three streams: s_update for trailing
// updates, s_panel for the look-ahead panel chain, and s_update_lp (64 CUs
// left to the panels) for the reduced-precision factor's updates and the
// LU's bulk updates.
#include <cstdlib>
#include <cstring>

#include "fla_hip_internal.h"

static FlaHipCtx g_ctx;

int g_tile_group = 4;

FlaHipCtx& fla_ctx() { return g_ctx; }

// The update stream may leave the last `reserved` CUs to the panel stream:
// the trailing update launches thousands of LDS-heavy workgroups, and a
// single-workgroup panel kernel that needs a whole CU's LDS would otherwise
// wait for a CU to drain completely.
static int g_reserved_cus = 8;
// The second update stream leaves 64 CUs to the panels.  It carries the
// reduced-precision factor's updates (measured on MI355X at n = 16384) and
// the LU's bulk updates, whose panel kernel needs up to 64 whole CUs with all
// its workgroups resident (with 8 left, LU panels waited for CUs: 119 ms
// instead of 87 ms per factorization at n = 16384).  One stream serves both
// so that the library's three streams, with the null stream, fit the four
// hardware queues HIP gives a process: two busy streams sharing a queue
// serialize (an LU with its bulk stream on the panel's queue took 182 ms
// instead of 134 ms).
static int g_lp_reserved_cus = 64;

static int create_masked_stream(hipStream_t* s, int reserved) {
  if (reserved <= 0 || reserved >= g_ctx.num_cus) {
    FLA_HIP_CHECK(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
    return FLA_HIP_OK;
  }
  const int words = (g_ctx.num_cus + 31) / 32;
  uint32_t mask[16] = {};
  if (words > 16) {
    FLA_HIP_CHECK(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
    return FLA_HIP_OK;
  }
  const int use = g_ctx.num_cus - reserved;
  for (int c = 0; c < use; ++c) mask[c / 32] |= 1u << (c % 32);
  if (hipExtStreamCreateWithCUMask(s, (uint32_t)words, mask) != hipSuccess) {
    (void)hipGetLastError();
    FLA_HIP_CHECK(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
  }
  return FLA_HIP_OK;
}

static int create_update_stream(int reserved) {
  return create_masked_stream(&g_ctx.s_update, reserved);
}

int fla_ensure_ready() {
  if (g_ctx.ready) return FLA_HIP_OK;
  return fla_hip_init();
}

// Process exit with the library's streams, events and graph still alive
// leaves their destruction to the HIP runtime's own teardown, after tools
// such as the profiler have already shut down.  The library therefore
// releases them itself from an exit handler registered after the runtime
// initialised (handlers run in reverse order of registration, so this one
// runs while the runtime is still whole).
static void fla_release_at_exit() { (void)fla_hip_finalize(); }

extern "C" int fla_hip_init(void) {
  if (g_ctx.ready) return FLA_HIP_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "libflame_amd: no HIP device available\n");
    return FLA_HIP_ENODEV;
  }
  FLA_HIP_CHECK(hipSetDevice(g_ctx.device));
  hipDeviceProp_t prop;
  FLA_HIP_CHECK(hipGetDeviceProperties(&prop, g_ctx.device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    std::fprintf(stderr, "libflame_amd: device 0 is %s, this build targets gfx950\n",
                 prop.gcnArchName);
    return FLA_HIP_ENODEV;
  }
  g_ctx.num_cus = prop.multiProcessorCount;
  // Default-priority streams only.
  FLA_TRY(create_update_stream(g_reserved_cus));
  FLA_TRY(create_masked_stream(&g_ctx.s_update_lp, g_lp_reserved_cus));
  FLA_HIP_CHECK(hipStreamCreateWithFlags(&g_ctx.s_panel, hipStreamNonBlocking));
  for (int i = 0; i < FlaHipCtx::kEvents; ++i)
    FLA_HIP_CHECK(hipEventCreateWithFlags(&g_ctx.ev[i], hipEventDisableTiming));
  for (int i = 0; i < 4; ++i)
    FLA_HIP_CHECK(hipEventCreateWithFlags(&g_ctx.ev_user[i], hipEventDisableTiming));
  FLA_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&g_ctx.dscratch), 4096));
  FLA_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&g_ctx.dlinv), 128 * 128 * sizeof(double)));
  // uncached exchange memory: allocated once per process, never freed
  static char* uc = nullptr;
  if (!uc) {
    FLA_HIP_CHECK(hipExtMallocWithFlags(reinterpret_cast<void**>(&uc), kUcBytes,
                                        hipDeviceMallocUncached));
    FLA_HIP_CHECK(hipMemset(uc, 0, kUcBytes));  // no exchange word carries a valid tag
  }
  g_ctx.uc = uc;
  FLA_TRY(fla_noalloc_setup(g_ctx.device));
  g_ctx.ready = true;
  static bool exit_hook = false;
  if (!exit_hook) {
    exit_hook = true;
    std::atexit(fla_release_at_exit);
  }
  return FLA_HIP_OK;
}

extern "C" int fla_hip_finalize(void) {
  if (!g_ctx.ready) return FLA_HIP_OK;
  FLA_HIP_CHECK(hipDeviceSynchronize());
  fla_stats_release();
  FLA_TRY(fla_batched_release());
  for (int i = 0; i < FlaHipCtx::kSlots; ++i) {
    if (g_ctx.work[i]) FLA_HIP_CHECK(hipFree(g_ctx.work[i]));
    g_ctx.work[i] = nullptr;
    g_ctx.work_bytes[i] = 0;
  }
  if (g_ctx.dscratch) FLA_HIP_CHECK(hipFree(g_ctx.dscratch));
  g_ctx.dscratch = nullptr;
  if (g_ctx.dlinv) FLA_HIP_CHECK(hipFree(g_ctx.dlinv));
  g_ctx.dlinv = nullptr;
  for (int i = 0; i < FlaHipCtx::kEvents; ++i) FLA_HIP_CHECK(hipEventDestroy(g_ctx.ev[i]));
  for (int i = 0; i < 4; ++i) FLA_HIP_CHECK(hipEventDestroy(g_ctx.ev_user[i]));
  FLA_HIP_CHECK(hipStreamDestroy(g_ctx.s_update));
  FLA_HIP_CHECK(hipStreamDestroy(g_ctx.s_update_lp));
  FLA_HIP_CHECK(hipStreamDestroy(g_ctx.s_panel));
  g_ctx.user_stream = nullptr;
  g_ctx.ready = false;
  return FLA_HIP_OK;
}

// ---- the caller's stream -------------------------------------------------
extern "C" int fla_hip_set_stream(void* stream) {
  FLA_TRY(fla_ensure_ready());
  g_ctx.user_stream = static_cast<hipStream_t>(stream);
  return FLA_HIP_OK;
}

extern "C" void* fla_hip_get_stream(void) { return g_ctx.user_stream; }

FlaUserStreamScope::FlaUserStreamScope() {
  if (!g_ctx.ready || !g_ctx.user_stream) return;
  // errors here resurface at the entry point's own calls on the same streams
  if (hipEventRecord(g_ctx.ev_user[0], g_ctx.user_stream) != hipSuccess) return;
  (void)hipStreamWaitEvent(g_ctx.s_update, g_ctx.ev_user[0], 0);
  (void)hipStreamWaitEvent(g_ctx.s_panel, g_ctx.ev_user[0], 0);
  (void)hipStreamWaitEvent(g_ctx.s_update_lp, g_ctx.ev_user[0], 0);
}

FlaUserStreamScope::~FlaUserStreamScope() {
  if (!g_ctx.ready || !g_ctx.user_stream) return;
  const hipStream_t mine[3] = {g_ctx.s_update, g_ctx.s_panel, g_ctx.s_update_lp};
  for (int i = 0; i < 3; ++i) {
    if (hipEventRecord(g_ctx.ev_user[1 + i], mine[i]) == hipSuccess)
      (void)hipStreamWaitEvent(g_ctx.user_stream, g_ctx.ev_user[1 + i], 0);
  }
}

extern "C" int fla_hip_synchronize(void) {
  FLA_TRY(fla_ensure_ready());
  FLA_HIP_CHECK(hipStreamSynchronize(g_ctx.s_panel));
  FLA_HIP_CHECK(hipStreamSynchronize(g_ctx.s_update_lp));
  FLA_HIP_CHECK(hipStreamSynchronize(g_ctx.s_update));
  return FLA_HIP_OK;
}

int fla_workspace_slot(int slot, size_t bytes, void** ptr) {
  FLA_TRY(fla_ensure_ready());
  if (slot < 0 || slot >= FlaHipCtx::kSlots) return FLA_HIP_EARG;
  if (bytes > g_ctx.work_bytes[slot]) {
    // Every queued kernel may still use the old buffer.
    FLA_HIP_CHECK(hipDeviceSynchronize());
    if (g_ctx.work[slot]) FLA_HIP_CHECK(hipFree(g_ctx.work[slot]));
    g_ctx.work[slot] = nullptr;
    g_ctx.work_bytes[slot] = 0;
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    FLA_HIP_CHECK(hipMalloc(&g_ctx.work[slot], want));
    g_ctx.work_bytes[slot] = want;
  }
  *ptr = g_ctx.work[slot];
  return FLA_HIP_OK;
}

int fla_workspace(size_t bytes, void** ptr) { return fla_workspace_slot(0, bytes, ptr); }

extern "C" int fla_hip_malloc(void** ptr, size_t bytes) {
  FLA_TRY(fla_ensure_ready());
  *ptr = nullptr;
  if (bytes == 0) return FLA_HIP_OK;
  FLA_HIP_CHECK(hipMalloc(ptr, bytes));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_free(void* ptr) {
  if (!ptr) return FLA_HIP_OK;
  FLA_TRY(fla_ensure_ready());
  FLA_HIP_CHECK(hipDeviceSynchronize());
  FLA_HIP_CHECK(hipFree(ptr));
  return FLA_HIP_OK;
}

static int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  FLA_TRY(fla_ensure_ready());
  if (bytes == 0) return FLA_HIP_OK;
  FLA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, g_ctx.s_update));
  FLA_HIP_CHECK(hipStreamSynchronize(g_ctx.s_update));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  FLA_USER_STREAM_SCOPE;
  return copy_sync(dst, src, bytes, hipMemcpyHostToDevice);
}
extern "C" int fla_hip_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  FLA_USER_STREAM_SCOPE;
  return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost);
}
extern "C" int fla_hip_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  FLA_USER_STREAM_SCOPE;
  return copy_sync(dst, src, bytes, hipMemcpyDeviceToDevice);
}

static int copy2d(int m, int n, const double* s, int lds, double* d, int ldd,
                  hipMemcpyKind kind) {
  FLA_TRY(fla_ensure_ready());
  if (m <= 0 || n <= 0) return FLA_HIP_OK;
  if (lds < m || ldd < m) return FLA_HIP_EARG;
  FLA_HIP_CHECK(hipMemcpy2DAsync(d, (size_t)ldd * sizeof(double), s,
                                 (size_t)lds * sizeof(double), (size_t)m * sizeof(double),
                                 (size_t)n, kind, g_ctx.s_update));
  FLA_HIP_CHECK(hipStreamSynchronize(g_ctx.s_update));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_dmatrix_h2d(int m, int n, const double* h, int ldh, double* d,
                                   int ldd) {
  FLA_USER_STREAM_SCOPE;
  return copy2d(m, n, h, ldh, d, ldd, hipMemcpyHostToDevice);
}
extern "C" int fla_hip_dmatrix_d2h(int m, int n, const double* d, int ldd, double* h,
                                   int ldh) {
  FLA_USER_STREAM_SCOPE;
  return copy2d(m, n, d, ldd, h, ldh, hipMemcpyDeviceToHost);
}
extern "C" int fla_hip_dmatrix_d2d(int m, int n, const double* s, int lds, double* d,
                                   int ldd) {
  FLA_USER_STREAM_SCOPE;
  return copy2d(m, n, s, lds, d, ldd, hipMemcpyDeviceToDevice);
}

extern "C" int fla_hip_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch,
                                size_t width, size_t height, int kind) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (width == 0 || height == 0) return FLA_HIP_OK;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  FLA_HIP_CHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k, g_ctx.s_update));
  FLA_HIP_CHECK(hipStreamSynchronize(g_ctx.s_update));
  return FLA_HIP_OK;
}

extern "C" int fla_hip_is_device_ptr(const void* ptr) {
  if (!ptr) return 0;
  if (fla_ensure_ready() != FLA_HIP_OK) return 0;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
    (void)hipGetLastError();  // plain host memory: clear the sticky error
    return 0;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

extern "C" const char* fla_hip_error_string(int status) {
  switch (status) {
    case FLA_HIP_OK: return "success";
    case FLA_HIP_EARG: return "invalid argument";
    case FLA_HIP_ENODEV: return "no gfx950 device";
    case FLA_HIP_ESPIN: return "inter-workgroup wait expired";
    case FLA_HIP_ERANGE: return "input does not fit the reduced precision";
    case FLA_HIP_ENOINIT: return "library not initialised (call FLA_Init or fla_hip_init)";
    case FLA_HIP_ESCRATCH: return "problem exceeds the no-allocation scratch limit";
    default: return hipGetErrorString((hipError_t)status);
  }
}

// 0 leaves a block size as it is; -1 restores the default (chol_nb: chosen
// from n; lu_nb: 256).
extern "C" int fla_hip_set_blocksizes(int chol_nb, int lu_nb) {
  if (chol_nb < -1 || lu_nb < -1) return FLA_HIP_EARG;
  if (chol_nb > 0 && (chol_nb % 128 != 0 || chol_nb > 1024)) return FLA_HIP_EARG;
  if (lu_nb > 0 && (lu_nb % 32 != 0 || (lu_nb > 128 && lu_nb != 256))) return FLA_HIP_EARG;
  if (chol_nb != 0) g_ctx.chol_nb = chol_nb > 0 ? chol_nb : 0;
  if (lu_nb != 0) g_ctx.lu_nb = lu_nb > 0 ? lu_nb : 256;
  return FLA_HIP_OK;
}

int fla_chol_nb(int n) {
  if (g_ctx.chol_nb > 0) return g_ctx.chol_nb;
  int nb = (n / 32) / 128 * 128;
  if (nb < 256) nb = 256;
  if (nb > 1024) nb = 1024;
  return nb;
}

extern "C" int fla_hip_chol_blocksize(int n) { return fla_chol_nb(n); }

extern "C" int fla_hip_get_blocksizes(int* chol_nb, int* lu_nb) {
  if (chol_nb) *chol_nb = g_ctx.chol_nb;
  if (lu_nb) *lu_nb = g_ctx.lu_nb;
  return FLA_HIP_OK;
}

// ---- kernel timing for measurement (HIP events on the launching stream) ----
#include <vector>

namespace {
struct PendingTiming {
  hipEvent_t start, stop;
  int kind;
  double flops;
};
struct StatsState {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<PendingTiming> pending;
  static constexpr int kKinds = 8;
  double flops[kKinds] = {};
  double ms[kKinds] = {};
  long launches[kKinds] = {};
};
StatsState g_stats;

hipEvent_t take_event() {
  if (!g_stats.pool.empty()) {
    hipEvent_t e = g_stats.pool.back();
    g_stats.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool fla_stats_on() { return g_stats.on; }

void* fla_stats_begin(hipStream_t s) {
  if (!g_stats.on) return nullptr;
  hipEvent_t e = take_event();
  if (e) (void)hipEventRecord(e, s);
  return e;
}

void fla_stats_end(hipStream_t s, void* token, int kind, double flops) {
  if (!g_stats.on || !token) return;
  hipEvent_t e = take_event();
  if (!e) return;
  (void)hipEventRecord(e, s);
  g_stats.pending.push_back({static_cast<hipEvent_t>(token), e, kind, flops});
}

void fla_stats_harvest() {
  for (auto& p : g_stats.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess && p.kind >= 0 &&
        p.kind < StatsState::kKinds) {
      g_stats.ms[p.kind] += ms;
      g_stats.flops[p.kind] += p.flops;
      g_stats.launches[p.kind] += 1;
    }
    g_stats.pool.push_back(p.start);
    g_stats.pool.push_back(p.stop);
  }
  g_stats.pending.clear();
}

void fla_stats_release() {
  for (auto& p : g_stats.pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  g_stats.pending.clear();
  for (hipEvent_t e : g_stats.pool) (void)hipEventDestroy(e);
  g_stats.pool.clear();
}

extern "C" int fla_hip_stats_enable(int on) {
  g_stats.on = on != 0;
  return FLA_HIP_OK;
}

extern "C" int fla_hip_stats_reset(void) {
  fla_stats_harvest();
  for (int i = 0; i < StatsState::kKinds; ++i) {
    g_stats.flops[i] = 0;
    g_stats.ms[i] = 0;
    g_stats.launches[i] = 0;
  }
  return FLA_HIP_OK;
}

extern "C" int fla_hip_stats_get(int kind, double* flops, double* ms, long* launches) {
  if (kind < 0 || kind >= StatsState::kKinds) return FLA_HIP_EARG;
  fla_stats_harvest();
  if (flops) *flops = g_stats.flops[kind];
  if (ms) *ms = g_stats.ms[kind];
  if (launches) *launches = g_stats.launches[kind];
  return FLA_HIP_OK;
}

extern "C" int fla_hip_set_reserved_cus(int reserved) {
  if (reserved < 0) return FLA_HIP_EARG;
  g_reserved_cus = reserved;
  if (!g_ctx.ready) return FLA_HIP_OK;
  FLA_HIP_CHECK(hipDeviceSynchronize());
  FLA_HIP_CHECK(hipStreamDestroy(g_ctx.s_update));
  g_ctx.s_update = nullptr;
  return create_update_stream(reserved);
}

extern "C" int fla_hip_mp_set_reserved_cus(int reserved) {
  if (reserved < 0) return FLA_HIP_EARG;
  g_lp_reserved_cus = reserved;
  if (!g_ctx.ready) return FLA_HIP_OK;
  FLA_HIP_CHECK(hipDeviceSynchronize());
  FLA_HIP_CHECK(hipStreamDestroy(g_ctx.s_update_lp));
  g_ctx.s_update_lp = nullptr;
  return create_masked_stream(&g_ctx.s_update_lp, reserved);
}

extern "C" int fla_hip_mp_set_blocksize(int nb) {
  if (nb < 0 || nb % 128 != 0) return FLA_HIP_EARG;
  FLA_TRY(fla_ensure_ready());
  g_ctx.mp_nb = nb;
  return FLA_HIP_OK;
}
