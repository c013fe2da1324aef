// ctx.hpp - the part of a context (mrx_ctx, defined in mrx_reion.hip) that
// the library's other translation units use.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

struct mrx_ctx;

namespace mrx {

hipStream_t ctx_stream(mrx_ctx* c);
int ctx_device(mrx_ctx* c);
// Records the message of a failure and returns `code`.
int ctx_fail(mrx_ctx* c, int code, const char* fmt, ...);
// Device buffer number `slot` (< kCtxSlots) of at least `bytes`, kept by the
// context until it is freed or a larger one is asked for; null on failure
// (the message is recorded).
constexpr int kCtxSlots = 24;
void* ctx_buffer(mrx_ctx* c, int slot, size_t bytes);
// Waits for the context's stream unless the caller supplied it.
int ctx_finish(mrx_ctx* c);

}  // namespace mrx
