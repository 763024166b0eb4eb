// CSR SpMV for gfx950:  y = A * x, A a row-major CompressedMatrix packed as CSR.
//
// Replaces the element-wise evaluation of SMatDVecMultExpr (blaze/math/expressions/
// SMatDVecMultExpr.h:204-207), i.e. one sparse row times x per element of y
// (SVecDVecInnerExpr.h:127-135: sp = v0*x[c0]; sp += v_k*x[c_k] in storage order).
//
// Two kernels, chosen per row by a plan built from the row-length histogram:
//  * vector-CSR: G lanes (1..32) per row, several rows per wavefront.  The G lanes load G
//    consecutive nonzeros at a time (coalesced); their products are then added in storage
//    order through lane shuffles, so every row's sum is the reference's sequential sum
//    (products rounded, then added left to right: bit-identical to the reference built
//    without floating-point contraction).
//  * wavefront-per-row: rows longer than the threshold get a whole wavefront; lanes stride
//    by 64 and the 64 partials are combined by a fixed butterfly (deterministic; within the
//    usual gamma_n bound of the sequential sum).
// Nonzero values and column indices are streamed with non-temporal loads so that x stays in
// L2 / MALL.
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

// Every product is rounded before it is added: never contract v*x + s into an FMA here.
#pragma clang fp contract(off)

struct bzg_csr_plan {
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int lanes = 1;            // lanes per row in the vector kernel
  int long_threshold = 0;   // rows with more nonzeros go to the wavefront kernel
  int32_t* short_rows = nullptr;  // device row list; nullptr = all rows, contiguous
  int64_t n_short = 0;
  int32_t* long_rows = nullptr;
  int64_t n_long = 0;
  int64_t max_row = 0;
  double mean_row = 0;
  char desc[256] = {0};
};

namespace bzg {
namespace {

// Products and sums with explicit rounding per operation (no contraction into FMA).
// (Plain operators compiled under `fp contract(off)` above.)
__device__ __forceinline__ double mul_rn(double a, double b) { return a * b; }
__device__ __forceinline__ double add_rn(double a, double b) { return a + b; }
__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }
__device__ __forceinline__ int32_t mul_rn(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a * (uint32_t)b);
}
__device__ __forceinline__ int32_t add_rn(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}
template <typename T>
__device__ __forceinline__ T neg_zero() { return T(0); }
template <>
__device__ __forceinline__ double neg_zero<double>() { return -0.0; }
template <>
__device__ __forceinline__ float neg_zero<float>() { return -0.0f; }

template <typename T, int G>
__global__ __launch_bounds__(256) void csr_vector_kernel(int64_t n_rows, const int32_t* __restrict__ rows,
                                                         const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ colidx,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ x,
                                                         T* __restrict__ y) {
  const int lig = threadIdx.x % G;
  const int group_base = (threadIdx.x % 64) - lig;  // first lane of this group in the wave
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (g >= n_rows) return;  // whole groups exit together
  const int64_t row = rows ? rows[g] : g;
  const int32_t start = rowptr[row], end = rowptr[row + 1];
  // Starting from -0 reproduces sp = p0 exactly (-0 + p == p for every p, also p = +-0).
  T s = neg_zero<T>();
  for (int32_t base = start; base < end; base += G) {
    const int32_t e = base + lig;
    T p = T(0);
    if (e < end) {
      const int32_t c = __builtin_nontemporal_load(colidx + e);
      p = mul_rn(__builtin_nontemporal_load(val + e), x[c]);
    }
    const int cnt = min(G, end - base);
#pragma unroll
    for (int t = 0; t < G; ++t) {
      const T q = (G == 1) ? p : __shfl(p, group_base + t, 64);
      if (t < cnt) s = add_rn(s, q);
    }
  }
  if (lig == 0) y[row] = (end == start) ? T(0) : s;
}

template <typename T>
__global__ __launch_bounds__(256) void csr_wave_kernel(int64_t n_rows, const int32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colidx,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ x,
                                                       T* __restrict__ y) {
  const int lane = threadIdx.x % 64;
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (w >= n_rows) return;
  const int64_t row = rows ? rows[w] : w;
  const int32_t start = rowptr[row], end = rowptr[row + 1];
  T s0 = T(0), s1 = T(0);
  int32_t e = start + lane;
  for (; e + 64 < end; e += 128) {
    const int32_t c0 = __builtin_nontemporal_load(colidx + e);
    const int32_t c1 = __builtin_nontemporal_load(colidx + e + 64);
    const T v0 = __builtin_nontemporal_load(val + e);
    const T v1 = __builtin_nontemporal_load(val + e + 64);
    s0 = add_rn(s0, mul_rn(v0, x[c0]));
    s1 = add_rn(s1, mul_rn(v1, x[c1]));
  }
  if (e < end) s0 = add_rn(s0, mul_rn(__builtin_nontemporal_load(val + e), x[__builtin_nontemporal_load(colidx + e)]));
  T s = add_rn(s0, s1);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s = add_rn(s, __shfl_xor(s, off, 64));
  if (lane == 0) y[row] = s;
}

template <typename T>
bzg_status csrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                 const int32_t* colidx, const T* val, const T* x, T* y) {
  BZG_REQUIRE(ctx && plan, BZG_ERR_INVALID_ARGUMENT, "csrmv: null context or plan");
  if (plan->n_rows == 0) return BZG_OK;
  BZG_REQUIRE(rowptr && y, BZG_ERR_INVALID_ARGUMENT, "csrmv: null rowptr or y");
  if (plan->nnz > 0) BZG_REQUIRE(colidx && val && x, BZG_ERR_INVALID_ARGUMENT, "csrmv: null CSR arrays or x");
  ProfScope prof(ctx, "csrmv");
  if (plan->n_short > 0) {
    const int G = plan->lanes;
    const int64_t blocks = (plan->n_short * G + 255) / 256;
    BZG_REQUIRE(blocks < (1LL << 31), BZG_ERR_UNSUPPORTED, "csrmv: too many rows");
    dim3 grid((unsigned)blocks), block(256);
#define BZG_VEC(GG)                                                                        \
  case GG:                                                                                 \
    hipLaunchKernelGGL((csr_vector_kernel<T, GG>), grid, block, 0, ctx->stream, plan->n_short, \
                       plan->short_rows, rowptr, colidx, val, x, y);                       \
    break;
    switch (G) {
      BZG_VEC(1) BZG_VEC(2) BZG_VEC(4) BZG_VEC(8) BZG_VEC(16) BZG_VEC(32)
      default:
        set_error("csrmv: bad lanes per row in plan");
        return BZG_ERR_INVALID_ARGUMENT;
    }
#undef BZG_VEC
    BZG_LAUNCH_CHECK();
  }
  if (plan->n_long > 0) {
    const int64_t blocks = (plan->n_long * 64 + 255) / 256;
    hipLaunchKernelGGL((csr_wave_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       plan->n_long, plan->long_rows, rowptr, colidx, val, x, y);
    BZG_LAUNCH_CHECK();
  }
  return BZG_OK;
}

}  // namespace
}  // namespace bzg

extern "C" {

bzg_status bzg_csr_plan_create(bzg_context* ctx, int64_t n_rows, int64_t n_cols,
                               const int32_t* rowptr_host, bzg_csr_plan** out) {
  BZG_REQUIRE(ctx && out, BZG_ERR_INVALID_ARGUMENT, "csr_plan_create: null argument");
  BZG_REQUIRE(n_rows >= 0 && n_cols >= 0, BZG_ERR_INVALID_ARGUMENT, "csr_plan_create: negative size");
  BZG_REQUIRE(n_rows < (1LL << 31), BZG_ERR_UNSUPPORTED, "csr_plan_create: n_rows >= 2^31");
  BZG_REQUIRE(n_rows == 0 || rowptr_host, BZG_ERR_INVALID_ARGUMENT, "csr_plan_create: null rowptr");
  *out = nullptr;
  auto* p = new bzg_csr_plan();
  p->n_rows = n_rows;
  p->n_cols = n_cols;
  if (n_rows == 0) {
    std::snprintf(p->desc, sizeof p->desc, "empty");
    *out = p;
    return BZG_OK;
  }
  if (rowptr_host[0] != 0) {
    delete p;
    bzg::set_error("csr_plan_create: rowptr[0] != 0");
    return BZG_ERR_INVALID_ARGUMENT;
  }
  // Row-length histogram in powers of two.
  int64_t hist[33] = {0};
  int64_t max_len = 0, total = 0;
  for (int64_t r = 0; r < n_rows; ++r) {
    const int64_t len = (int64_t)rowptr_host[r + 1] - rowptr_host[r];
    if (len < 0) {
      delete p;
      bzg::set_error("csr_plan_create: rowptr not non-decreasing");
      return BZG_ERR_INVALID_ARGUMENT;
    }
    int b = 0;
    while ((1LL << b) < len) ++b;
    hist[b]++;
    max_len = std::max(max_len, len);
    total += len;
  }
  p->nnz = total;
  p->max_row = max_len;
  p->mean_row = (double)total / (double)n_rows;
  // Rows longer than 4 wavefront-widths of work for a 32-lane group go to the wavefront
  // kernel; the remaining rows pick the lane count nearest their mean length.
  const int64_t long_thr = 64;
  int64_t n_long = 0, short_nnz = 0;
  for (int b = 0; b <= 32; ++b)
    if ((1LL << b) > long_thr) n_long += hist[b];
  for (int64_t r = 0; r < n_rows; ++r) {
    const int64_t len = (int64_t)rowptr_host[r + 1] - rowptr_host[r];
    if (len <= long_thr) short_nnz += len;
  }
  const int64_t n_short = n_rows - n_long;
  const double mean_short = n_short ? (double)short_nnz / (double)n_short : 0.0;
  int G = 1;
  while (G < 32 && G < mean_short) G *= 2;
  p->lanes = G;
  p->long_threshold = (int)long_thr;
  p->n_short = n_short;
  p->n_long = n_long;
  if (n_long > 0) {
    std::vector<int32_t> sr, lr;
    sr.reserve(n_short);
    lr.reserve(n_long);
    for (int64_t r = 0; r < n_rows; ++r) {
      const int64_t len = (int64_t)rowptr_host[r + 1] - rowptr_host[r];
      (len > long_thr ? lr : sr).push_back((int32_t)r);
    }
    if (n_short) BZG_HIP_CHECK(hipMalloc(&p->short_rows, n_short * sizeof(int32_t)));
    BZG_HIP_CHECK(hipMalloc(&p->long_rows, n_long * sizeof(int32_t)));
    if (n_short)
      BZG_HIP_CHECK(hipMemcpyAsync(p->short_rows, sr.data(), n_short * sizeof(int32_t),
                                   hipMemcpyHostToDevice, ctx->stream));
    BZG_HIP_CHECK(hipMemcpyAsync(p->long_rows, lr.data(), n_long * sizeof(int32_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  std::snprintf(p->desc, sizeof p->desc,
                "rows=%lld nnz=%lld mean=%.2f max=%lld vector-csr lanes=%d rows=%lld; "
                "wavefront-per-row rows=%lld (len>%lld)",
                (long long)n_rows, (long long)total, p->mean_row, (long long)max_len, G,
                (long long)n_short, (long long)n_long, (long long)long_thr);
  *out = p;
  return BZG_OK;
}

bzg_status bzg_csr_plan_destroy(bzg_csr_plan* p) {
  if (!p) return BZG_OK;
  if (p->short_rows) hipFree(p->short_rows);
  if (p->long_rows) hipFree(p->long_rows);
  delete p;
  return BZG_OK;
}

bzg_status bzg_csr_plan_describe(const bzg_csr_plan* p, char* buf, size_t len) {
  BZG_REQUIRE(p && buf && len > 0, BZG_ERR_INVALID_ARGUMENT, "csr_plan_describe: null argument");
  std::snprintf(buf, len, "%s", p->desc);
  return BZG_OK;
}

bzg_status bzg_dcsrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                      const int32_t* colidx, const double* val, const double* x, double* y) {
  return bzg::csrmv<double>(ctx, plan, rowptr, colidx, val, x, y);
}
bzg_status bzg_scsrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                      const int32_t* colidx, const float* val, const float* x, float* y) {
  return bzg::csrmv<float>(ctx, plan, rowptr, colidx, val, x, y);
}
bzg_status bzg_icsrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                      const int32_t* colidx, const int32_t* val, const int32_t* x, int32_t* y) {
  return bzg::csrmv<int32_t>(ctx, plan, rowptr, colidx, val, x, y);
}

}  // extern "C"
