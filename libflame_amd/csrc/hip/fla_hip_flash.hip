// Kernels of the hierarchical (FLASH) storage: moving data between the
// leaves of a storage-by-blocks matrix and a flat column-major matrix.
//
// The reference walks the hierarchy on the host and copies block by block
// (FLASH_Copy_flat_to_hier / FLASH_Copy_hier_to_flat,
// src/base/flamec/hierarchy/main/FLASH_Copy_other.c:81-138).  Here the host
// flattens the walk into a table of pieces (device pointer, leading
// dimension, rows, cols, row and column in the flat matrix) and ONE launch
// moves every piece: workgroup (t, c) copies chunk c of piece t, lanes
// walking rows so that both sides are read and written in 8-byte (or 4-byte)
// coalesced runs.  HBM-bound: 2 * elem bytes per element.
#include "fla_hip_internal.h"

namespace {

constexpr int kCopyThreads = 256;
constexpr int kChunk = 16 * kCopyThreads;  // elements per workgroup

template <typename T>
__global__ __launch_bounds__(kCopyThreads) void tile_copy_kernel(const long long* __restrict__ desc,
                                                                 T* flat, long ldf,
                                                                 int to_tiles) {
  const long long* d = desc + 6 * (long)blockIdx.x;
  T* tile = reinterpret_cast<T*>((size_t)d[0]);
  const long ldt = (long)d[1];
  const long rows = (long)d[2], cols = (long)d[3];
  T* f = flat + (long)d[4] + (long)d[5] * ldf;
  const long total = rows * cols;
  const long e0 = (long)blockIdx.y * kChunk;
  for (long e = e0 + threadIdx.x; e < total && e < e0 + kChunk; e += kCopyThreads) {
    const long i = e % rows, j = e / rows;
    if (to_tiles)
      tile[i + j * ldt] = f[i + j * ldf];
    else
      f[i + j * ldf] = tile[i + j * ldt];
  }
}

}  // namespace

extern "C" int fla_hip_tile_copy(int ntiles, const long long* desc, void* flat, long ldf,
                                 int elem_bytes, int to_tiles) {
  FLA_USER_STREAM_SCOPE;
  FLA_TRY(fla_ensure_ready());
  if (ntiles < 0 || (elem_bytes != 4 && elem_bytes != 8) || ldf < 1) return FLA_HIP_EARG;
  if (ntiles == 0) return FLA_HIP_OK;
  long long maxel = 0;
  for (int t = 0; t < ntiles; ++t) {
    const long long* d = desc + 6 * (long)t;
    if (d[1] < 1 || d[2] < 0 || d[3] < 0 || d[4] < 0 || d[5] < 0 || (d[2] > 0 && d[1] < d[2]) ||
        d[4] + d[2] > ldf)
      return FLA_HIP_EARG;
    if (d[2] * d[3] > maxel) maxel = d[2] * d[3];
  }
  if (maxel == 0) return FLA_HIP_OK;
  const size_t bytes = (size_t)ntiles * 6 * sizeof(long long);
  void* w = nullptr;
  FLA_TRY(fla_workspace_slot(4, bytes, &w));
  hipStream_t s = fla_ctx().s_update;
  FLA_HIP_CHECK(hipMemcpyAsync(w, desc, bytes, hipMemcpyHostToDevice, s));
  const dim3 grid((unsigned)ntiles, (unsigned)((maxel + kChunk - 1) / kChunk));
  if (elem_bytes == 8)
    hipLaunchKernelGGL(tile_copy_kernel<double>, grid, dim3(kCopyThreads), 0, s,
                       static_cast<const long long*>(w), static_cast<double*>(flat), ldf,
                       to_tiles);
  else
    hipLaunchKernelGGL(tile_copy_kernel<int>, grid, dim3(kCopyThreads), 0, s,
                       static_cast<const long long*>(w), static_cast<int*>(flat), ldf, to_tiles);
  FLA_HIP_CHECK(hipGetLastError());
  // the descriptor table is host memory the caller may release on return
  FLA_HIP_CHECK(hipStreamSynchronize(s));
  return FLA_HIP_OK;
}
