// util.hip - device utilities of the library that are not on the reionization
// path itself: the streaming copy the benchmark takes as the device's practical
// copy rate (mrx_stream_copy, include/mrx_reion.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/mrx_reion.h"
#include "ctx.hpp"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kCopyThreads = 256;
constexpr int kCopyUnroll = 4;  // 16-byte loads in flight per thread before its first store

// Grid-stride copy of n 16-byte elements: each thread issues kCopyUnroll loads,
// then stores them with non-temporal stores (the written lines bypass the
// caches the reads go through).  Every index is checked against n.
__global__ __launch_bounds__(kCopyThreads) void k_stream_copy(f32x4* __restrict__ dst, const f32x4* __restrict__ src,
                                                              size_t n)
{
  const size_t step = (size_t)gridDim.x * kCopyThreads * kCopyUnroll;
  for (size_t base = (size_t)blockIdx.x * kCopyThreads * kCopyUnroll + threadIdx.x; base < n; base += step) {
    f32x4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const size_t i = base + (size_t)u * kCopyThreads;
      if (i < n) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const size_t i = base + (size_t)u * kCopyThreads;
      if (i < n) __builtin_nontemporal_store(v[u], dst + i);
    }
  }
}

}  // namespace

extern "C" int mrx_stream_copy(mrx_ctx* ctx, void* dst, const void* src, size_t bytes, int reps, double* ms_per_copy)
{
  if (!ctx) return MRX_EINVAL;
  if (!dst || !src || bytes == 0 || (bytes % 16) != 0 || reps < 1 || ((size_t)dst % 16) || ((size_t)src % 16))
    return mrx::ctx_fail(ctx, MRX_EINVAL, "stream_copy: 16-byte aligned device buffers, bytes a multiple of 16");
  if (hipSetDevice(mrx::ctx_device(ctx)) != hipSuccess) return mrx::ctx_fail(ctx, MRX_EDEVICE, "hipSetDevice");
  hipStream_t st = mrx::ctx_stream(ctx);
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, mrx::ctx_device(ctx)) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_stream_copy, kCopyThreads, 0) != hipSuccess)
    return mrx::ctx_fail(ctx, MRX_EDEVICE, "stream_copy: device attributes");
  const size_t n = bytes / 16;
  const size_t need = (n + (size_t)kCopyThreads * kCopyUnroll - 1) / ((size_t)kCopyThreads * kCopyUnroll);
  const unsigned grid = (unsigned)std::min<size_t>(need, (size_t)std::max(1, cus * per_cu));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return mrx::ctx_fail(ctx, MRX_EDEVICE, "stream_copy: events");
  }
  int rc = MRX_OK;
  k_stream_copy<<<grid, kCopyThreads, 0, st>>>((f32x4*)dst, (const f32x4*)src, n);  // warm-up
  (void)hipEventRecord(e0, st);
  for (int r = 0; r < reps; ++r) k_stream_copy<<<grid, kCopyThreads, 0, st>>>((f32x4*)dst, (const f32x4*)src, n);
  (void)hipEventRecord(e1, st);
  if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
    rc = mrx::ctx_fail(ctx, MRX_EDEVICE, "stream_copy: launch failed");
  } else {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = (double)ms / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}
