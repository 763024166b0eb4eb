// Shared definitions of the gfx950 kernels and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>
#include "blaze_gpu.h"

namespace bzg {

// Per-thread text of the last failure, returned by bzg_last_error().
void set_error(const std::string& msg);

#define BZG_HIP_CHECK(expr)                                                          \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::bzg::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
      return BZG_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

#define BZG_REQUIRE(cond, code, msg)                                                 \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      ::bzg::set_error(msg);                                                         \
      return (code);                                                                 \
    }                                                                                \
  } while (0)

// Launch check: the launch itself (bad configuration) is reported immediately.
#define BZG_LAUNCH_CHECK() BZG_HIP_CHECK(hipGetLastError())

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
  unsigned generation = 0;  // bumped on every (re)allocation: contents cached in it are gone
};

}  // namespace bzg

struct bzg_context {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  // Cached scratch for kernels that need it (split GEMM slices, solver staging).
  bzg::Workspace ws[20];
  // Optional HIP-event timing of named kernels (bzg_profile_enable / bzg_profile_read).
  // plan of the last FP64-accurate GEMM on this context (bzg_fp64acc_last_plan)
  int fp64acc_moduli = 0, fp64acc_bits = 0;
  double fp64acc_est = 0.0;
  double* host_stats = nullptr;       // pinned readback of the split GEMM's operand statistics
  unsigned fp64acc_cc_gen = 0;       // workspace generation holding the split GEMM's CRT constants
  int fp64acc_cc_key = -1;            // ... and for which (slice type, moduli count); -1: none
  // the residue GEMM's 128 KB dynamic-LDS limit has been raised on this context's device, per
  // mode (0 BF16 slices, 1 INT8 slices, 2 INT32 digits)
  bool residue_lds_attr[3] = {false, false, false};
  bool profile = false;
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> prof;
};

namespace bzg {
// Brackets the kernels launched in its scope with HIP events on the context stream when
// profiling is enabled (the events measure device time of exactly those launches).
struct ProfScope {
  bzg_context* ctx;
  const char* name;
  hipEvent_t b = nullptr, e = nullptr;
  ProfScope(bzg_context* c, const char* n) : ctx(c), name(n) {
    if (ctx && ctx->profile && hipEventCreate(&b) == hipSuccess && hipEventCreate(&e) == hipSuccess)
      hipEventRecord(b, ctx->stream);
    else
      b = e = nullptr;
  }
  ~ProfScope() {
    if (b && e) {
      hipEventRecord(e, ctx->stream);
      ctx->prof[name].push_back({b, e});
    }
  }
};

// Resolves the device address of the no-allocation solver's scratch array on `device` (current).
bzg_status noalloc_scratch_init(int device);

// On-chip mixed-precision batched solver (lu_mixed.hip): stable partition of a batch into
// order classes, and the launches of the class kernels.
bzg_status launch_partition(bzg_context* ctx, int64_t count, const int32_t* n, const int32_t* nrhs,
                            const int32_t* order, int32_t* info, int32_t* iter, int32_t* lists, int cap,
                            int* counts);
bzg_status launch_dsgesv_onchip(bzg_context* ctx, const int32_t* lists, int cap, const int* counts,
                                const int32_t* n, const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                                const double* A, const double* B, double* X, int32_t* iter, int32_t* info,
                                int* counters, int32_t* fb_list, int* fb_count, double tol, int max_iter,
                                const int* wg_per_cu, int64_t count);

// True when the CSR plan holds a sliced-ELL copy of element type dtype (0 double, 1 float,
// 2 int32; spmv.hip).
bool plan_has_sell(const bzg_csr_plan* plan, int dtype);

// Returns a cached device buffer of at least `bytes` (slot-indexed); grows on demand.
bzg_status workspace(bzg_context* ctx, int slot, size_t bytes, void** out);

// INT32 GEMM C = alpha op(A) op(B) + beta C (mod 2^32) from balanced INT8 digits on the INT8
// MFMA (ozaki.hip); the caller has checked the arguments and m, n, k >= 1.  One pass covers
// k <= kIgemmDigitsMaxK; longer K runs in chunks of at most kIgemmDigitsChunk.
constexpr int64_t kIgemmDigitsMaxK = 32767;
constexpr int64_t kIgemmDigitsChunk = 32640;  // 255 x 128
bzg_status igemm_digits_impl(bzg_context* ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k, int32_t alpha,
                             const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb, int32_t beta,
                             int32_t* C, int64_t ldc);

typedef double d4 __attribute__((ext_vector_type(4)));
// four doubles that need only 8-byte alignment (one 32-byte global access, split as needed)
typedef double dvec4u __attribute__((ext_vector_type(4), aligned(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Wavefront-wide helpers (wave64).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace bzg
