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

// One 16-byte element per thread over a grid that covers the buffer, stored
// with a non-temporal store (the written lines bypass the caches the reads go
// through).  Measured on MI355X over 2 GiB this form reaches 6.1 TB/s read +
// write, where grid-stride loops with four or eight loads in flight per thread
// over one to sixteen workgroups per CU reach 5.1-5.5 TB/s.
__global__ __launch_bounds__(kCopyThreads) void k_stream_copy(f32x4* __restrict__ dst, const f32x4* __restrict__ src,
                                                              size_t n)
{
  const size_t i = (size_t)blockIdx.x * kCopyThreads + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(src[i], dst + i);
}

}  // namespace

extern "C" int mrx_stream_copy(mrx_ctx* ctx, void* dst, const void* src, size_t bytes, int reps, double* ms_per_copy)
{
  if (!ctx) return MRX_EINVAL;
  if (!dst || !src || bytes == 0 || (bytes % 16) != 0 || reps < 1 || ((size_t)dst % 16) || ((size_t)src % 16))
    return mrx::ctx_fail(ctx, MRX_EINVAL, "stream_copy: 16-byte aligned device buffers, bytes a multiple of 16");
  if (hipSetDevice(mrx::ctx_device(ctx)) != hipSuccess) return mrx::ctx_fail(ctx, MRX_EDEVICE, "hipSetDevice");
  hipStream_t st = mrx::ctx_stream(ctx);
  const size_t n = bytes / 16;
  const size_t need = (n + kCopyThreads - 1) / kCopyThreads;
  if (need > 0x7FFFFFFFu) return mrx::ctx_fail(ctx, MRX_EINVAL, "stream_copy: at most 2^31 workgroups (32 TiB)");
  const unsigned grid = (unsigned)need;
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
