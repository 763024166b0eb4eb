// Context, stream and memory entry points of the C ABI.
#include "common.hpp"

#include <algorithm>

#include <cstring>

namespace bzg {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

bzg_status workspace(bzg_context* ctx, int slot, size_t bytes, void** out) {
  Workspace& w = ctx->ws[slot];
  if (w.bytes < bytes) {
    // small buffers grow geometrically so that slowly growing requests (batch metadata)
    // do not reallocate on every call; large ones exactly
    if (w.bytes && bytes < (64u << 20)) bytes = std::max(bytes, 2 * w.bytes);
    if (w.ptr) {
      BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      BZG_HIP_CHECK(hipFree(w.ptr));
      w.ptr = nullptr;
      w.bytes = 0;
    }
    hipError_t e = hipMalloc(&w.ptr, bytes);
    if (e != hipSuccess) {
      w.ptr = nullptr;
      set_error(std::string("workspace hipMalloc: ") + hipGetErrorString(e));
      return BZG_ERR_OUT_OF_MEMORY;
    }
    w.bytes = bytes;
    ++w.generation;
  }
  *out = w.ptr;
  return BZG_OK;
}
}  // namespace bzg

extern "C" {

bzg_status bzg_create(bzg_context** out) {
  BZG_REQUIRE(out != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_create: null output pointer");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    bzg::set_error("bzg_create: no HIP device visible");
    return BZG_ERR_NO_DEVICE;
  }
  auto* ctx = new bzg_context();
  BZG_HIP_CHECK(hipGetDevice(&ctx->device));
  hipDeviceProp_t prop;
  BZG_HIP_CHECK(hipGetDeviceProperties(&prop, ctx->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    bzg::set_error(std::string("bzg_create: device is ") + prop.gcnArchName +
                   ", this library is built for gfx950 only");
    delete ctx;
    return BZG_ERR_NO_DEVICE;
  }
  ctx->num_cus = prop.multiProcessorCount;
  {
    const bzg_status st = bzg::noalloc_scratch_init(ctx->device);
    if (st) {
      delete ctx;
      return st;
    }
  }
  BZG_HIP_CHECK(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
  ctx->stream = ctx->own_stream;
  *out = ctx;
  return BZG_OK;
}

bzg_status bzg_destroy(bzg_context* ctx) {
  if (!ctx) return BZG_OK;
  hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->prof)
    for (auto& pr : kv.second) {
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
  for (auto& w : ctx->ws)
    if (w.ptr) hipFree(w.ptr);
  if (ctx->host_stats) hipHostFree(ctx->host_stats);
  if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return BZG_OK;
}

bzg_status bzg_set_stream(bzg_context* ctx, void* s) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_set_stream: null context");
  ctx->stream = static_cast<hipStream_t>(s);
  return BZG_OK;
}

void* bzg_get_stream(bzg_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }
void* bzg_own_stream(bzg_context* ctx) { return ctx ? (void*)ctx->own_stream : nullptr; }

bzg_status bzg_synchronize(bzg_context* ctx) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_synchronize: null context");
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return BZG_OK;
}

const char* bzg_status_string(bzg_status s) {
  if (s > 0) return "singular matrix (exactly zero pivot)";
  switch (s) {
    case BZG_OK: return "ok";
    case BZG_ERR_INVALID_ARGUMENT: return "invalid argument";
    case BZG_ERR_SIZE_MISMATCH: return "sizes do not match";
    case BZG_ERR_HIP: return "HIP runtime error";
    case BZG_ERR_OUT_OF_MEMORY: return "out of device memory";
    case BZG_ERR_NOT_SQUARE: return "non-square system matrix";
    case BZG_ERR_UNSUPPORTED: return "unsupported shape or option";
    case BZG_ERR_NO_DEVICE: return "no gfx950 device";
    case BZG_ERR_BUSY: return "busy";
    case BZG_ERR_CANCELLED: return "cancelled";
    default: return "unknown status";
  }
}

const char* bzg_last_error(void) { return bzg::g_last_error.c_str(); }

bzg_status bzg_profile_enable(bzg_context* ctx, int32_t on) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_profile_enable: null context");
  ctx->profile = on != 0;
  return BZG_OK;
}

bzg_status bzg_profile_read(bzg_context* ctx, const char* kernel, double* total_ms, int64_t* launches) {
  BZG_REQUIRE(ctx && kernel && total_ms && launches, BZG_ERR_INVALID_ARGUMENT,
              "bzg_profile_read: null argument");
  *total_ms = 0.0;
  *launches = 0;
  auto it = ctx->prof.find(kernel);
  if (it == ctx->prof.end()) return BZG_OK;
  for (auto& pr : it->second) {
    BZG_HIP_CHECK(hipEventSynchronize(pr.second));
    float ms = 0.f;
    BZG_HIP_CHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
    *total_ms += ms;
    *launches += 1;
    hipEventDestroy(pr.first);
    hipEventDestroy(pr.second);
  }
  it->second.clear();
  return BZG_OK;
}

bzg_status bzg_malloc(void** dptr, size_t bytes) {
  BZG_REQUIRE(dptr != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_malloc: null output pointer");
  *dptr = nullptr;
  if (bytes == 0) return BZG_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    *dptr = nullptr;
    bzg::set_error(std::string("bzg_malloc: ") + hipGetErrorString(e));
    return BZG_ERR_OUT_OF_MEMORY;
  }
  return BZG_OK;
}

bzg_status bzg_free(void* dptr) {
  if (dptr) BZG_HIP_CHECK(hipFree(dptr));
  return BZG_OK;
}

bzg_status bzg_memcpy(bzg_context* ctx, void* dst, const void* src, size_t bytes, int kind) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_memcpy: null context");
  if (bytes == 0) return BZG_OK;
  BZG_REQUIRE(dst && src, BZG_ERR_INVALID_ARGUMENT, "bzg_memcpy: null pointer");
  hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost
                                : hipMemcpyDeviceToDevice;
  BZG_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return BZG_OK;
}

bzg_status bzg_memset_zero(bzg_context* ctx, void* dptr, size_t bytes) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "bzg_memset_zero: null context");
  if (bytes == 0) return BZG_OK;
  BZG_HIP_CHECK(hipMemsetAsync(dptr, 0, bytes, ctx->stream));
  return BZG_OK;
}

}  // extern "C"
