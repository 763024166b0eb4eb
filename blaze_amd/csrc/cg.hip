// Conjugate gradients on a resident CSR matrix: the loop of the reference's CG benchmark
// (blazemark/src/blaze/CG.cpp:114-135) with every vector in HBM:
//   r = A x - b;  delta = r.r;  d = -r;
//   repeat: h = A d;  alpha = delta / (d.h);  x += alpha d;  r += alpha h;
//           beta = r.r;  d = (beta / delta) d - r;  delta = beta.
// Per iteration: the SpMV (bzg_dcsrmv, same plan and kernels as y = S * x; the plan's
// sliced-ELL copy through bzg_dsellmv when it holds one, same sums), then three streaming
// kernels.  Dot products are two-stage and deterministic: each workgroup writes the
// partial sum of its fixed slice (fixed butterfly inside), and every consumer adds the
// partials in index order itself, so alpha and beta are computed redundantly but identically
// by every workgroup that needs them and no scalar ever leaves the device.
#include "common.hpp"

#include <algorithm>

namespace bzg {
namespace {

constexpr int kCgThreads = 256;
constexpr int kCgPerThread = 4;                         // elements per thread and pass
constexpr int kCgSlice = kCgThreads * kCgPerThread;     // elements per workgroup slice
constexpr int kCgMaxParts = 2048;                       // partial sums per dot product

// Sum of the first `parts` partials in index order (every caller gets the same bits).
__device__ double sum_parts(const double* __restrict__ part, int parts) {
  __shared__ double red[kCgThreads];
  double s = 0.0;
  for (int q = threadIdx.x; q < parts; q += kCgThreads) s += part[q];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kCgThreads / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double t = red[0];
  __syncthreads();
  return t;
}

// Workgroup sum of v (fixed order), written by thread 0 to part[blockIdx.x].
__device__ void block_partial(double v, double* __restrict__ part) {
  __shared__ double red[kCgThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kCgThreads / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Grid-stride over slices: workgroup g owns slices g, g + G, ...  (G = gridDim.x <= kCgMaxParts)
#define CG_FOR(i)                                                                        \
  for (int64_t base = (int64_t)blockIdx.x * kCgSlice; base < n; base += (int64_t)gridDim.x * kCgSlice) \
    for (int u = 0; u < kCgPerThread; ++u)                                               \
      if (const int64_t i = base + u * kCgThreads + threadIdx.x; i < n)

// r = h - b (h = A x), d = -r, partial r.r
__global__ __launch_bounds__(kCgThreads) void cg_init_kernel(int64_t n, const double* __restrict__ h,
                                                              const double* __restrict__ b,
                                                              double* __restrict__ r, double* __restrict__ d,
                                                              double* __restrict__ part_rr) {
  double s = 0.0;
  CG_FOR(i) {
    const double ri = h[i] - b[i];
    r[i] = ri;
    d[i] = -ri;
    s = fma(ri, ri, s);
  }
  block_partial(s, part_rr);
}

// partial d.h
__global__ __launch_bounds__(kCgThreads) void cg_dot_kernel(int64_t n, const double* __restrict__ d,
                                                             const double* __restrict__ h,
                                                             double* __restrict__ part) {
  double s = 0.0;
  CG_FOR(i) s = fma(d[i], h[i], s);
  block_partial(s, part);
}

// alpha = delta / (d.h);  x += alpha d;  r += alpha h;  partial r.r
__global__ __launch_bounds__(kCgThreads) void cg_update_kernel(int64_t n, int parts,
                                                                const double* __restrict__ part_rr_old,
                                                                const double* __restrict__ part_dh,
                                                                double* __restrict__ x, double* __restrict__ r,
                                                                const double* __restrict__ d,
                                                                const double* __restrict__ h,
                                                                double* __restrict__ part_rr) {
  const double delta = sum_parts(part_rr_old, parts);
  const double dh = sum_parts(part_dh, parts);
  const double alpha = delta / dh;
  double s = 0.0;
  CG_FOR(i) {
    x[i] = fma(alpha, d[i], x[i]);
    const double ri = fma(alpha, h[i], r[i]);
    r[i] = ri;
    s = fma(ri, ri, s);
  }
  block_partial(s, part_rr);
}

// d = (beta / delta) d - r
__global__ __launch_bounds__(kCgThreads) void cg_direction_kernel(int64_t n, int parts,
                                                                   const double* __restrict__ part_rr_old,
                                                                   const double* __restrict__ part_rr,
                                                                   double* __restrict__ d,
                                                                   const double* __restrict__ r) {
  const double delta = sum_parts(part_rr_old, parts);
  const double beta = sum_parts(part_rr, parts);
  const double f = beta / delta;
  CG_FOR(i) d[i] = fma(f, d[i], -r[i]);
}

#undef CG_FOR

}  // namespace
}  // namespace bzg

extern "C" {

bzg_status bzg_dcg(bzg_context* ctx, const bzg_csr_plan* plan, int64_t n, const int32_t* rowptr,
                   const int32_t* colidx, const double* val, const double* b, double* x, int32_t iterations,
                   double* rr_out) {
  BZG_REQUIRE(ctx && plan, BZG_ERR_INVALID_ARGUMENT, "cg: null context or plan");
  BZG_REQUIRE(n >= 0 && iterations >= 0, BZG_ERR_INVALID_ARGUMENT, "cg: negative size or iteration count");
  if (n == 0) return BZG_OK;
  BZG_REQUIRE(rowptr && colidx && val && b && x, BZG_ERR_INVALID_ARGUMENT, "cg: null array");
  // workspace slot 10: r, d, h (n each) and three rotating partial-sum arrays
  const int parts = (int)std::min<int64_t>(bzg::kCgMaxParts, (n + bzg::kCgSlice - 1) / bzg::kCgSlice);
  void* w = nullptr;
  bzg_status st = bzg::workspace(ctx, 10, (size_t)3 * n * sizeof(double) + 3 * bzg::kCgMaxParts * sizeof(double),
                                 &w);
  if (st) return st;
  double* r = static_cast<double*>(w);
  double* d = r + n;
  double* h = d + n;
  double* pa = h + n;                    // r.r of the current delta
  double* pb = pa + bzg::kCgMaxParts;    // d.h
  double* pc = pb + bzg::kCgMaxParts;    // r.r of the next delta
  bzg::ProfScope prof(ctx, "cg");
  const dim3 grid((unsigned)parts), block(bzg::kCgThreads);
  const bool sliced = bzg::plan_has_sell(plan, 0);
  auto spmv = [&](const double* in, double* out) {
    return sliced ? bzg_dsellmv(ctx, plan, in, out) : bzg_dcsrmv(ctx, plan, rowptr, colidx, val, in, out);
  };
  // r = A x - b, delta = r.r, d = -r
  st = spmv(x, h);
  if (st) return st;
  hipLaunchKernelGGL(bzg::cg_init_kernel, grid, block, 0, ctx->stream, n, h, b, r, d, pa);
  BZG_LAUNCH_CHECK();
  for (int it = 0; it < iterations; ++it) {
    st = spmv(d, h);  // h = A d
    if (st) return st;
    hipLaunchKernelGGL(bzg::cg_dot_kernel, grid, block, 0, ctx->stream, n, d, h, pb);
    hipLaunchKernelGGL(bzg::cg_update_kernel, grid, block, 0, ctx->stream, n, parts, pa, pb, x, r, d, h, pc);
    hipLaunchKernelGGL(bzg::cg_direction_kernel, grid, block, 0, ctx->stream, n, parts, pa, pc, d, r);
    BZG_LAUNCH_CHECK();
    std::swap(pa, pc);  // delta = beta
  }
  if (rr_out) {
    // r.r of the last residual, summed in the same order as the kernels do
    double hp[bzg::kCgMaxParts];
    BZG_HIP_CHECK(hipMemcpyAsync(hp, pa, (size_t)parts * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double red[bzg::kCgThreads] = {0.0};
    for (int t = 0; t < bzg::kCgThreads; ++t)
      for (int q = t; q < parts; q += bzg::kCgThreads) red[t] += hp[q];
    for (int w2 = bzg::kCgThreads / 2; w2 >= 1; w2 >>= 1)
      for (int t = 0; t < w2; ++t) red[t] += red[t + w2];
    *rr_out = red[0];
  }
  return BZG_OK;
}

}  // extern "C"
