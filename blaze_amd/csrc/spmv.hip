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
// Nonzero values and column indices use ordinary cached loads: a row's chunks and neighbouring
// rows share cache lines, which non-temporal loads would fetch from HBM once per access.
#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

// Every product is rounded before it is added: never contract v*x + s into an FMA here.
#pragma clang fp contract(off)

struct bzg_csr_plan {
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int lanes = 1;            // lanes per row in the vector kernel
  int rows_per_group = 1;   // rows per lane group in the vector kernel
  int cap = 8;              // nonzeros per lane per batch in the vector kernel
  int mode = 0;             // short rows: 0 = row blocks through LDS, 1 = lane groups
  int4* blocks = nullptr;   // device: (row0, row1, nz0, nz1) of each LDS row block
  int64_t n_blocks = 0;
  int block_nnz = 2048;     // nonzeros per row block (2048 or 4096)
  int block_nt = 0;         // non-temporal loads of values / column indices
  int long_threshold = 0;   // rows with more nonzeros leave the sequential (short-row) path
  int32_t* short_rows = nullptr;  // device row list; nullptr = all rows, contiguous
  int64_t n_short = 0;
  int32_t* long_rows = nullptr;   // rows handled by one wavefront each
  int64_t n_long = 0;
  int32_t* huge_rows = nullptr;   // rows handled by one workgroup each
  int64_t n_huge = 0;
  int64_t max_row = 0;
  double mean_row = 0;
  int plus_zero = 0;        // sums start from +0 (column-major product semantics)
  // sliced-ELL copy of an all-short-row matrix (bzg_csr_plan_pack_sell): slices of 64
  // consecutive rows stored column by column (slot k of the slice's rows, then slot k + 1),
  // width = the slice's longest row; padding slots hold column -1 and value 0
  int64_t n_slices = 0;
  int64_t* sell_off = nullptr;     // device: n_slices + 1 element offsets (multiples of 64)
  int32_t* sell_col = nullptr;
  void* sell_val = nullptr;
  int sell_dtype = -1;             // 0 double, 1 float, 2 int32; -1 not packed
  std::vector<int32_t> rowptr_host;  // kept for re-planning by the autotuner
  char desc[256] = {0};
};

namespace bzg {
namespace {

// Workgroup ids are dispatched round-robin over the 8 XCDs.  Remap them so each XCD walks a
// contiguous eighth of the rows: neighbouring rows (and, for stencil-like matrices, the x
// entries they gather) then stay in one XCD's L2.  The grid is padded to a multiple of 8.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  return (b % 8) * (nb / 8) + b / 8;
}

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

// G lanes per row, R rows per lane group, CAP nonzeros per lane per batch.  A batch first
// issues the loads of all R*CAP (value, column) pairs of a lane, then all their x gathers,
// and only then sums, so a row costs one dependent round trip per G*CAP nonzeros (three in
// total for a short row: rowptr -> column -> x) instead of one per element.  The products of
// a batch are added in storage order (lane t of chunk c holds element c*G + t), which keeps
// every row's sum equal to the reference's sequential sum.
template <typename T, int G, int R, int CAP>
__global__ __launch_bounds__(256) void csr_vector_kernel(int64_t n_rows, const int32_t* __restrict__ rows,
                                                         const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ colidx,
                                                         const T* __restrict__ val,
                                                         const T* __restrict__ x,
                                                         T* __restrict__ y, T s0) {
  const int lig = threadIdx.x % G;
  const int group_base = (threadIdx.x % 64) - lig;  // first lane of this group in the wave
  const int64_t g = (xcd_block() * blockDim.x + threadIdx.x) / G;
  const int64_t first = g * R;
  if (first >= n_rows) return;  // whole groups exit together
  constexpr int BATCH = G * CAP;
  int64_t row[R];
  int32_t start[R], len[R];
  int32_t batches = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t q = first + r;
    row[r] = -1;
    start[r] = 0;
    len[r] = 0;
    if (q < n_rows) {
      row[r] = rows ? rows[q] : q;
      start[r] = rowptr[row[r]];
      len[r] = rowptr[row[r] + 1] - start[r];
    }
    batches = max(batches, (len[r] + BATCH - 1) / BATCH);
  }
  // s0 = -0 reproduces sp = p0 exactly (-0 + p == p for every p, also p = +-0); s0 = +0 is the
  // reset-then-accumulate of the column-major product.
  T s[R];
#pragma unroll
  for (int r = 0; r < R; ++r) s[r] = s0;
  for (int32_t b = 0; b < batches; ++b) {
    T v[R][CAP];
    int32_t c[R][CAP];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < CAP; ++k) {
        const int32_t o = b * BATCH + k * G + lig;
        const bool ok = o < len[r];
        v[r][k] = ok ? val[start[r] + o] : T(0);
        c[r][k] = ok ? colidx[start[r] + o] : 0;
      }
    T p[R][CAP];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < CAP; ++k) p[r][k] = mul_rn(v[r][k], x[c[r][k]]);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < CAP; ++k) {
        const int32_t base = b * BATCH + k * G;
        const int cnt = min(G, len[r] - base);
#pragma unroll
        for (int t = 0; t < G; ++t) {
          const T q = (G == 1) ? p[r][k] : __shfl(p[r][k], group_base + t, 64);
          if (t < cnt) s[r] = add_rn(s[r], q);
        }
      }
  }
  if (lig == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row[r] >= 0) y[row[r]] = (len[r] == 0) ? T(0) : s[r];
  }
}

// Row blocks through LDS (for short rows): a block is a run of consecutive short rows holding at
// most BNNZ nonzeros (at most 512 rows); its nonzeros form one contiguous range, read with
// fully coalesced loads.  The products go to LDS and thread t then adds the products of rows
// t and t+256 in storage order: the reference's sequential sum, bit for bit.  A block
// descriptor holds (first row, end row, first nonzero, end nonzero).
// One workgroup per block, in plain dispatch order: the workgroups in flight then cover one
// narrow window of the nonzero stream, which keeps few DRAM pages open at a time (measured on
// the 7-point Laplacians: 5.2 TB/s at 256^3 against 4.0-4.4 TB/s for a persistent grid whose
// workgroups each walk their own range), and the x entries that neighbouring blocks gather are
// shared through the last-level cache.  One LDS buffer of BNNZ products per workgroup.
constexpr int kBlockRowsMax = 512;
constexpr int kShortRow = 32;  // rows up to this length are summed sequentially

template <typename T, int BNNZ, bool NT>
__global__ __launch_bounds__(256) void csr_block_kernel(int64_t n_blocks, const int4* __restrict__ blocks,
                                                        const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colidx,
                                                        const T* __restrict__ val,
                                                        const T* __restrict__ x, T* __restrict__ y,
                                                        T s0) {
  __shared__ T prod[BNNZ];
  constexpr int PER = BNNZ / 256;
  const int tid = threadIdx.x;
  const int64_t b = blockIdx.x;
  if (b >= n_blocks) return;  // uniform over the workgroup
  const int4 cur = blocks[b];
  const int cnt = cur.w - cur.z;
  T p[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + 256 * i;
    const bool ok = e < cnt;
    T v;
    int32_t c;
    if (NT) {
      v = ok ? __builtin_nontemporal_load(val + cur.z + e) : T(0);
      c = ok ? __builtin_nontemporal_load(colidx + cur.z + e) : 0;
    } else {
      v = ok ? val[cur.z + e] : T(0);
      c = ok ? colidx[cur.z + e] : 0;
    }
    p[i] = mul_rn(v, x[c]);  // padding lanes gather x[0] (the host never launches this with nnz = 0)
  }
  const int nrow = cur.y - cur.x;
  int32_t rs[2], re[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = tid + 256 * h;
    rs[h] = re[h] = 0;
    if (t < nrow) {
      rs[h] = rowptr[cur.x + t] - cur.z;
      re[h] = rowptr[cur.x + t + 1] - cur.z;
    }
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + 256 * i;
    if (e < cnt) prod[e] = p[i];
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int t = tid + 256 * h;
    if (t < nrow) {
      T s = s0;
      for (int32_t k = rs[h]; k < re[h]; ++k) s = add_rn(s, prod[k]);
      y[cur.x + t] = (re[h] == rs[h]) ? T(0) : s;
    }
  }
}

// Sliced ELL (SELL-64, rows in their original order): lane l of wavefront s sums row 64 s + l
// from the slice's column-major slots, eight slots' loads and gathers in flight at a time, in
// slot order = storage order (padding slots, column -1, are skipped, not multiplied), so every
// row is the reference's sequential sum bit for bit.  No LDS, no barrier, no row pointers:
// value and column loads are 64 consecutive elements per wavefront instruction.
template <typename T>
__global__ __launch_bounds__(256) void sell_kernel(int64_t n_rows, int64_t n_slices, const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ col, const T* __restrict__ val,
                                                   const T* __restrict__ x, T* __restrict__ y, T s0) {
  const int lane = threadIdx.x & 63;
  const int64_t sl = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (sl >= n_slices) return;  // uniform over the wavefront
  const int64_t base = off[sl];
  const int w = (int)((off[sl + 1] - base) >> 6);
  T sum = s0;
  bool any = false;
  for (int k0 = 0; k0 < w; k0 += 8) {
    int32_t c[8];
    T v[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      // slots past the width re-read the last slot (in bounds, no branch) and are dropped
      const bool in = k0 + u < w;
      const int64_t e = base + (int64_t)(in ? k0 + u : w - 1) * 64 + lane;
      const int32_t cl = __builtin_nontemporal_load(col + e);
      const T vl = __builtin_nontemporal_load(val + e);
      c[u] = in ? cl : -1;
      v[u] = in ? vl : T(0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = c[u] >= 0 ? x[c[u]] : T(0);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (c[u] >= 0) {
        sum = add_rn(sum, mul_rn(v[u], xv[u]));
        any = true;
      }
    }
  }
  const int64_t row = sl * 64 + lane;
  if (row < n_rows) y[row] = any ? sum : T(0);
}

// CSR -> SELL-64: thread per row writes the row's slots (its nonzeros, then padding).
template <typename E>
__global__ __launch_bounds__(256) void sell_pack_kernel(int64_t n_rows, int64_t n_slices, const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colidx, const E* __restrict__ val,
                                                        int32_t* __restrict__ scol, E* __restrict__ sval) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t sl = r >> 6;
  if (sl >= n_slices) return;
  const int lane = (int)(r & 63);
  const int64_t base = off[sl];
  const int w = (int)((off[sl + 1] - base) >> 6);
  const int64_t b = r < n_rows ? rowptr[r] : 0;
  const int len = r < n_rows ? (int)(rowptr[r + 1] - b) : 0;
  for (int k = 0; k < w; ++k) {
    const int64_t e = base + (int64_t)k * 64 + lane;
    scol[e] = k < len ? colidx[b + k] : -1;
    sval[e] = k < len ? val[b + k] : E(0);
  }
}

template <typename T>
bzg_status sellmv(bzg_context* ctx, const bzg_csr_plan* plan, int dtype, const T* x, T* y) {
  BZG_REQUIRE(ctx && plan, BZG_ERR_INVALID_ARGUMENT, "sellmv: null context or plan");
  BZG_REQUIRE(plan->sell_dtype == dtype, BZG_ERR_INVALID_ARGUMENT,
              "sellmv: the plan holds no sliced-ELL copy of this element type (bzg_csr_plan_pack_sell)");
  if (plan->n_rows == 0) return BZG_OK;
  BZG_REQUIRE(x && y, BZG_ERR_INVALID_ARGUMENT, "sellmv: null x or y");
  ProfScope prof(ctx, "sellmv");
  const T s0 = plan->plus_zero ? T(0) : T(-T(0));
  const int64_t wgs = (plan->n_slices + 3) / 4;
  BZG_REQUIRE(wgs < (1LL << 31), BZG_ERR_UNSUPPORTED, "sellmv: too many rows");
  hipLaunchKernelGGL(sell_kernel<T>, dim3((unsigned)wgs), dim3(256), 0, ctx->stream, plan->n_rows, plan->n_slices,
                     plan->sell_off, plan->sell_col, static_cast<const T*>(plan->sell_val), x, y, s0);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

// Long rows: LANES threads per row (64 = one wavefront, 256 = one workgroup).  Each thread
// issues the loads of up to PER nonzeros at once, then their x gathers, keeping PER partial
// sums; the partials are combined in a fixed order (lane butterfly, then the 4 wavefronts in
// order), so the result is deterministic and within gamma_len * sum|v||x| of the sequential
// sum.
template <typename T, int LANES, int PER>
__global__ __launch_bounds__(256) void csr_rows_kernel(int64_t n_rows, const int32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colidx,
                                                       const T* __restrict__ val,
                                                       const T* __restrict__ x,
                                                       T* __restrict__ y) {
  __shared__ T part[4];
  const int tid = threadIdx.x;
  const int t = tid % LANES;
  const int64_t w = (xcd_block() * blockDim.x + tid) / LANES;
  if (w >= n_rows) return;  // uniform per row group (LANES = 256: per workgroup)
  const int64_t row = rows ? rows[w] : w;
  const int32_t start = rowptr[row], end = rowptr[row + 1];
  T acc[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) acc[k] = T(0);
  for (int32_t base = start; base < end; base += LANES * PER) {
    T v[PER];
    int32_t c[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int32_t e = base + t + LANES * k;
      const bool ok = e < end;
      v[k] = ok ? val[e] : T(0);
      c[k] = ok ? colidx[e] : 0;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) acc[k] = add_rn(acc[k], mul_rn(v[k], x[c[k]]));
  }
  T s = acc[0];
#pragma unroll
  for (int k = 1; k < PER; ++k) s = add_rn(s, acc[k]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s = add_rn(s, __shfl_xor(s, off, 64));
  if (LANES == 64) {
    if (t == 0) y[row] = s;
  } else {
    if ((tid & 63) == 0) part[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) y[row] = add_rn(add_rn(part[0], part[1]), add_rn(part[2], part[3]));
  }
}

// y = 0 for a matrix without nonzeros (x may then be null, so no SpMV kernel runs).
template <typename T>
__global__ __launch_bounds__(256) void zero_rows_kernel(int64_t n, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = T(0);
}

template <typename T>
bzg_status csrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                 const int32_t* colidx, const T* val, const T* x, T* y) {
  BZG_REQUIRE(ctx && plan, BZG_ERR_INVALID_ARGUMENT, "csrmv: null context or plan");
  if (plan->n_rows == 0) return BZG_OK;
  BZG_REQUIRE(rowptr && y, BZG_ERR_INVALID_ARGUMENT, "csrmv: null rowptr or y");
  if (plan->nnz > 0) BZG_REQUIRE(colidx && val && x, BZG_ERR_INVALID_ARGUMENT, "csrmv: null CSR arrays or x");
  ProfScope prof(ctx, "csrmv");
  if (plan->nnz == 0) {
    hipLaunchKernelGGL(zero_rows_kernel<T>, dim3((unsigned)((plan->n_rows + 255) / 256)), dim3(256), 0,
                       ctx->stream, plan->n_rows, y);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  // start of the sequential row sums: -0 (SMatDVecMultExpr: the first product itself) or +0
  // (a column-major product: y reset, then accumulated)
  const T s0 = plan->plus_zero ? T(0) : T(-T(0));
  if (plan->n_short > 0 && plan->mode == 0) {
    const int64_t blocks = plan->n_blocks;  // one workgroup per row block
    BZG_REQUIRE(blocks < (1LL << 31), BZG_ERR_UNSUPPORTED, "csrmv: too many row blocks");
    if (plan->block_nnz == 4096) {
      if (plan->block_nt)
        hipLaunchKernelGGL((csr_block_kernel<T, 4096, true>), dim3((unsigned)blocks), dim3(256), 0,
                           ctx->stream, plan->n_blocks, plan->blocks, rowptr, colidx, val, x, y, s0);
      else
        hipLaunchKernelGGL((csr_block_kernel<T, 4096, false>), dim3((unsigned)blocks), dim3(256), 0,
                           ctx->stream, plan->n_blocks, plan->blocks, rowptr, colidx, val, x, y, s0);
    } else {
      if (plan->block_nt)
        hipLaunchKernelGGL((csr_block_kernel<T, 2048, true>), dim3((unsigned)blocks), dim3(256), 0,
                           ctx->stream, plan->n_blocks, plan->blocks, rowptr, colidx, val, x, y, s0);
      else
        hipLaunchKernelGGL((csr_block_kernel<T, 2048, false>), dim3((unsigned)blocks), dim3(256), 0,
                           ctx->stream, plan->n_blocks, plan->blocks, rowptr, colidx, val, x, y, s0);
    }
    BZG_LAUNCH_CHECK();
  } else if (plan->n_short > 0) {
    const int G = plan->lanes;
    const int R = plan->rows_per_group;
    const int CP = plan->cap;
    const int64_t groups = (plan->n_short + R - 1) / R;
    const int64_t blocks = ((groups * G + 255) / 256 + 7) / 8 * 8;
    BZG_REQUIRE(blocks < (1LL << 31), BZG_ERR_UNSUPPORTED, "csrmv: too many rows");
    dim3 grid((unsigned)blocks), block(256);
#define BZG_VEC(GG, RR, CC)                                                                    \
  if (G == GG && R == RR && CP == CC) {                                                        \
    hipLaunchKernelGGL((csr_vector_kernel<T, GG, RR, CC>), grid, block, 0, ctx->stream,         \
                       plan->n_short, plan->short_rows, rowptr, colidx, val, x, y, s0);         \
    launched = true;                                                                           \
  }
#define BZG_VEC_C(GG, RR) BZG_VEC(GG, RR, 1) BZG_VEC(GG, RR, 2) BZG_VEC(GG, RR, 4) BZG_VEC(GG, RR, 8)
#define BZG_VEC_R(GG) BZG_VEC_C(GG, 1) BZG_VEC_C(GG, 2)
    bool launched = false;
    BZG_VEC_R(1) BZG_VEC_R(2) BZG_VEC_R(4) BZG_VEC_R(8) BZG_VEC_R(16) BZG_VEC_R(32)
#undef BZG_VEC_R
#undef BZG_VEC_C
#undef BZG_VEC
    if (!launched) {
      set_error("csrmv: bad lanes/rows per group in plan");
      return BZG_ERR_INVALID_ARGUMENT;
    }
    BZG_LAUNCH_CHECK();
  }
  if (plan->n_long > 0) {
    const int64_t blocks = ((plan->n_long * 64 + 255) / 256 + 7) / 8 * 8;
    hipLaunchKernelGGL((csr_rows_kernel<T, 64, 4>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       plan->n_long, plan->long_rows, rowptr, colidx, val, x, y);
    BZG_LAUNCH_CHECK();
  }
  if (plan->n_huge > 0) {
    const int64_t blocks = (plan->n_huge + 7) / 8 * 8;
    hipLaunchKernelGGL((csr_rows_kernel<T, 256, 8>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       plan->n_huge, plan->huge_rows, rowptr, colidx, val, x, y);
    BZG_LAUNCH_CHECK();
  }
  return BZG_OK;
}

}  // namespace
}  // namespace bzg

namespace bzg {
bool plan_has_sell(const bzg_csr_plan* plan, int dtype) { return plan && plan->sell_dtype == dtype; }

// Groups consecutive short rows into LDS row blocks of at most `bnnz` nonzeros and `brows`
// rows (a long row ends the current block) and uploads the descriptors.
static bzg_status build_row_blocks(bzg_context* ctx, bzg_csr_plan* p, const int32_t* rowptr_host,
                                   int bnnz, int brows) {
  std::vector<int4> bl;
  const int64_t n_rows = p->n_rows, long_thr = p->long_threshold;
  int64_t r = 0;
  while (r < n_rows) {
    const int64_t len = (int64_t)rowptr_host[r + 1] - rowptr_host[r];
    if (len > long_thr) { ++r; continue; }
    const int64_t r0 = r;
    int64_t nz = 0;
    while (r < n_rows && r - r0 < brows) {
      const int64_t l = (int64_t)rowptr_host[r + 1] - rowptr_host[r];
      if (l > long_thr || nz + l > bnnz) break;
      nz += l;
      ++r;
    }
    bl.push_back(make_int4((int)r0, (int)r, rowptr_host[r0], rowptr_host[r]));
  }
  if (p->blocks) hipFree(p->blocks);
  p->blocks = nullptr;
  p->n_blocks = (int64_t)bl.size();
  p->block_nnz = bnnz;
  if (p->n_blocks) {
    BZG_HIP_CHECK(hipMalloc(&p->blocks, bl.size() * sizeof(int4)));
    hipStream_t st = ctx ? ctx->stream : nullptr;
    BZG_HIP_CHECK(hipMemcpyAsync(p->blocks, bl.data(), bl.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    BZG_HIP_CHECK(hipStreamSynchronize(st));
  }
  return BZG_OK;
}
}  // namespace bzg

extern "C" {

static bzg_status csr_plan_create_impl(bzg_context* ctx, int64_t n_rows, int64_t n_cols,
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
  p->rowptr_host.assign(rowptr_host, rowptr_host + n_rows + 1);
  p->mean_row = (double)total / (double)n_rows;
  // Bins by row length: short rows (<= kShortRow) are summed sequentially (bit-identical to the
  // reference order) in LDS row blocks or lane groups; longer rows get a wavefront each, rows
  // beyond 256 nonzeros a whole workgroup each (fixed-order tree sums).
  const int64_t long_thr = bzg::kShortRow, huge_thr = 256;
  int64_t n_long = 0, n_huge = 0;
  std::vector<int32_t> sr, lr, hr;
  for (int64_t r = 0; r < n_rows; ++r) {
    const int64_t len = (int64_t)rowptr_host[r + 1] - rowptr_host[r];
    if (len > huge_thr) hr.push_back((int32_t)r);
    else if (len > long_thr) lr.push_back((int32_t)r);
    else sr.push_back((int32_t)r);
  }
  n_long = (int64_t)lr.size();
  n_huge = (int64_t)hr.size();
  const int64_t n_short = (int64_t)sr.size();
  // 90th percentile of the short rows' lengths (power-of-two histogram bins)
  int64_t acc90 = 0, L90 = 1;
  for (int b = 0; b <= 32 && (1LL << b) <= long_thr; ++b) {
    acc90 += hist[b];
    if (acc90 * 10 >= n_short * 9) { L90 = 1LL << b; break; }
  }
  // lane-group alternative (selectable by the autotuner): G * 8 covers the typical short row
  int G = 1;
  while (G < 32 && G * 8 < L90) G *= 2;
  int CP = 1;
  while (CP < 8 && G * CP < L90) CP *= 2;
  p->lanes = G;
  p->cap = CP;
  p->rows_per_group = 1;
  p->long_threshold = (int)long_thr;
  p->n_short = n_short;
  p->n_long = n_long;
  p->n_huge = n_huge;
  {
    bzg_status bst = bzg::build_row_blocks(ctx, p, rowptr_host, 2048, 256);
    if (bst) { delete p; return bst; }
  }
  // Row blocks pay off when short rows come in long runs (a stencil); when long rows break the
  // short ones into short runs (power-law rows), lane groups over the short-row list do better.
  p->mode = (p->n_blocks == 0 || n_short >= 32 * p->n_blocks) ? 0 : 1;
  auto upload_list = [&](const std::vector<int32_t>& v, int32_t** dst) -> bzg_status {
    BZG_HIP_CHECK(hipMalloc(dst, v.size() * sizeof(int32_t)));
    BZG_HIP_CHECK(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                 ctx->stream));
    return BZG_OK;
  };
  bzg_status ust = BZG_OK;
  if (!ust && n_long + n_huge > 0 && n_short > 0) ust = upload_list(sr, &p->short_rows);
  if (!ust && n_long > 0) ust = upload_list(lr, &p->long_rows);
  if (!ust && n_huge > 0) ust = upload_list(hr, &p->huge_rows);
  if (ust) {
    bzg_csr_plan_destroy(p);
    return ust;
  }
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::snprintf(p->desc, sizeof p->desc,
                "rows=%lld nnz=%lld mean=%.2f max=%lld %s rows=%lld; "
                "wavefront-per-row rows=%lld (%lld<len<=%lld); workgroup-per-row rows=%lld (len>%lld)",
                (long long)n_rows, (long long)total, p->mean_row, (long long)max_len,
                p->mode == 0 ? (std::string("row-blocks=") + std::to_string(p->n_blocks) + " (lds)").c_str()
                             : (std::string("vector-csr lanes=") + std::to_string(G) + " rows/group=1 cap=" +
                                std::to_string(CP)).c_str(),
                (long long)n_short, (long long)n_long, (long long)long_thr, (long long)huge_thr,
                (long long)n_huge, (long long)huge_thr);
  *out = p;
  return BZG_OK;
}

bzg_status bzg_csr_plan_create(bzg_context* ctx, int64_t n_rows, int64_t n_cols,
                               const int32_t* rowptr_host, bzg_csr_plan** out) {
  // no C++ exception leaves the C ABI: a host allocation failure becomes a status
  try {
    return csr_plan_create_impl(ctx, n_rows, n_cols, rowptr_host, out);
  } catch (const std::bad_alloc&) {
    bzg::set_error("csr_plan_create: out of host memory");
    return BZG_ERR_OUT_OF_MEMORY;
  }
}

bzg_status bzg_csr_plan_tune(bzg_csr_plan* p, int32_t lanes, int32_t rows_per_group, int32_t cap) {
  BZG_REQUIRE(p != nullptr, BZG_ERR_INVALID_ARGUMENT, "csr_plan_tune: null plan");
  if (lanes == 0) {  // row blocks through LDS: rows_per_group 1 -> 2048 nnz / 256 rows,
                     // 2 -> 4096 / 512; cap != 0 -> non-temporal nonzero loads
    BZG_REQUIRE(p->rowptr_host.size() == (size_t)p->n_rows + 1, BZG_ERR_INVALID_ARGUMENT,
                "csr_plan_tune: plan has no host row pointers");
    p->mode = 0;
    p->block_nt = cap != 0;
    return bzg::build_row_blocks(nullptr, p, p->rowptr_host.data(), rows_per_group == 2 ? 4096 : 2048,
                            rows_per_group == 2 ? 512 : 256);
  }
  p->mode = 1;
  BZG_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32,
              BZG_ERR_INVALID_ARGUMENT, "csr_plan_tune: lanes must be 1, 2, 4, 8, 16 or 32");
  BZG_REQUIRE(rows_per_group == 1 || rows_per_group == 2, BZG_ERR_INVALID_ARGUMENT,
              "csr_plan_tune: rows_per_group must be 1 or 2");
  BZG_REQUIRE(cap == 1 || cap == 2 || cap == 4 || cap == 8, BZG_ERR_INVALID_ARGUMENT,
              "csr_plan_tune: cap must be 1, 2, 4 or 8");
  p->lanes = lanes;
  p->rows_per_group = rows_per_group;
  p->cap = cap;
  const char* mark = std::strstr(p->desc, " row-blocks=");
  if (mark) {
    char tail[256];
    const char* semi = std::strchr(mark, ';');
    std::snprintf(tail, sizeof tail, "%s", semi ? semi : "");
    std::snprintf(p->desc + (mark - p->desc), sizeof p->desc - (mark - p->desc),
                  " vector-csr lanes=%d rows/group=%d cap=%d rows=%lld%s", lanes, rows_per_group,
                  cap, (long long)p->n_short, tail);
  }
  return BZG_OK;
}

bzg_status bzg_csr_plan_set_sum_start(bzg_csr_plan* p, int32_t positive_zero) {
  BZG_REQUIRE(p != nullptr, BZG_ERR_INVALID_ARGUMENT, "csr_plan_set_sum_start: null plan");
  p->plus_zero = positive_zero != 0;
  return BZG_OK;
}

bzg_status bzg_csc_to_csr_host(int64_t n_rows, int64_t n_cols, const int32_t* colptr, const int32_t* rowidx,
                               const void* val, int32_t elem_bytes, int32_t* rowptr, int32_t* colidx,
                               void* val_out) {
  BZG_REQUIRE(n_rows >= 0 && n_cols >= 0 && n_rows < (1LL << 31), BZG_ERR_INVALID_ARGUMENT,
              "csc_to_csr_host: bad size");
  BZG_REQUIRE(colptr && rowptr && (elem_bytes == 4 || elem_bytes == 8), BZG_ERR_INVALID_ARGUMENT,
              "csc_to_csr_host: bad argument");
  BZG_REQUIRE(colptr[0] == 0, BZG_ERR_INVALID_ARGUMENT, "csc_to_csr_host: colptr[0] != 0");
  const int64_t nnz = colptr[n_cols];
  BZG_REQUIRE(nnz >= 0 && (nnz == 0 || (rowidx && val && colidx && val_out)), BZG_ERR_INVALID_ARGUMENT,
              "csc_to_csr_host: null array");
  // counting sort by row; columns are visited in order, so every row lists its columns ascending
  std::vector<int64_t> cnt;
  try {
    cnt.assign((size_t)n_rows + 1, 0);
  } catch (const std::bad_alloc&) {
    bzg::set_error("csc_to_csr_host: out of host memory");
    return BZG_ERR_OUT_OF_MEMORY;
  }
  for (int64_t c = 0; c < n_cols; ++c) {
    BZG_REQUIRE(colptr[c + 1] >= colptr[c], BZG_ERR_INVALID_ARGUMENT, "csc_to_csr_host: colptr decreasing");
    for (int64_t e = colptr[c]; e < colptr[c + 1]; ++e) {
      BZG_REQUIRE(rowidx[e] >= 0 && rowidx[e] < n_rows, BZG_ERR_INVALID_ARGUMENT, "csc_to_csr_host: row index");
      ++cnt[(size_t)rowidx[e] + 1];
    }
  }
  for (int64_t r = 0; r < n_rows; ++r) cnt[(size_t)r + 1] += cnt[(size_t)r];
  for (int64_t r = 0; r <= n_rows; ++r) rowptr[r] = (int32_t)cnt[(size_t)r];
  const char* src = static_cast<const char*>(val);
  char* dst = static_cast<char*>(val_out);
  for (int64_t c = 0; c < n_cols; ++c)
    for (int64_t e = colptr[c]; e < colptr[c + 1]; ++e) {
      const int64_t q = cnt[(size_t)rowidx[e]]++;
      colidx[q] = (int32_t)c;
      std::memcpy(dst + q * elem_bytes, src + e * elem_bytes, (size_t)elem_bytes);
    }
  return BZG_OK;
}

bzg_status bzg_csr_plan_destroy(bzg_csr_plan* p) {
  if (!p) return BZG_OK;
  if (p->sell_off) hipFree(p->sell_off);
  if (p->sell_col) hipFree(p->sell_col);
  if (p->sell_val) hipFree(p->sell_val);
  if (p->short_rows) hipFree(p->short_rows);
  if (p->blocks) hipFree(p->blocks);
  if (p->long_rows) hipFree(p->long_rows);
  if (p->huge_rows) hipFree(p->huge_rows);
  delete p;
  return BZG_OK;
}

bzg_status bzg_csr_plan_describe(const bzg_csr_plan* p, char* buf, size_t len) {
  BZG_REQUIRE(p && buf && len > 0, BZG_ERR_INVALID_ARGUMENT, "csr_plan_describe: null argument");
  std::snprintf(buf, len, "%s", p->desc);
  return BZG_OK;
}

bzg_status bzg_csr_plan_pack_sell(bzg_context* ctx, bzg_csr_plan* p, const int32_t* rowptr, const int32_t* colidx,
                                  const void* val, int32_t dtype) {
  BZG_REQUIRE(ctx && p, BZG_ERR_INVALID_ARGUMENT, "csr_plan_pack_sell: null context or plan");
  BZG_REQUIRE(dtype >= 0 && dtype <= 2, BZG_ERR_INVALID_ARGUMENT, "csr_plan_pack_sell: dtype must be 0, 1 or 2");
  BZG_REQUIRE(p->n_long == 0 && p->n_huge == 0, BZG_ERR_UNSUPPORTED,
              "csr_plan_pack_sell: rows longer than 32 nonzeros (the sliced layout is for short rows)");
  BZG_REQUIRE(p->rowptr_host.size() == (size_t)p->n_rows + 1, BZG_ERR_INVALID_ARGUMENT,
              "csr_plan_pack_sell: plan has no host row pointers");
  if (p->sell_off) hipFree(p->sell_off);
  if (p->sell_col) hipFree(p->sell_col);
  if (p->sell_val) hipFree(p->sell_val);
  p->sell_off = nullptr;
  p->sell_col = nullptr;
  p->sell_val = nullptr;
  p->sell_dtype = -1;
  p->n_slices = (p->n_rows + 63) / 64;
  if (p->n_rows == 0) {
    p->sell_dtype = dtype;
    return BZG_OK;
  }
  BZG_REQUIRE(rowptr && (p->nnz == 0 || (colidx && val)), BZG_ERR_INVALID_ARGUMENT,
              "csr_plan_pack_sell: null CSR arrays");
  std::vector<int64_t> off((size_t)p->n_slices + 1, 0);
  const std::vector<int32_t>& rp = p->rowptr_host;
  for (int64_t sl = 0; sl < p->n_slices; ++sl) {
    int w = 0;
    for (int64_t r = sl * 64; r < std::min<int64_t>(p->n_rows, sl * 64 + 64); ++r) w = std::max(w, rp[r + 1] - rp[r]);
    off[(size_t)sl + 1] = off[(size_t)sl] + 64 * (int64_t)w;
  }
  const int64_t elems = off.back();
  const size_t esz = dtype == 0 ? 8 : 4;
  BZG_HIP_CHECK(hipMalloc(&p->sell_off, off.size() * sizeof(int64_t)));
  BZG_HIP_CHECK(hipMalloc(&p->sell_col, (size_t)std::max<int64_t>(elems, 1) * sizeof(int32_t)));
  BZG_HIP_CHECK(hipMalloc(&p->sell_val, (size_t)std::max<int64_t>(elems, 1) * esz));
  BZG_HIP_CHECK(hipMemcpyAsync(p->sell_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                               ctx->stream));
  const unsigned grid = (unsigned)((p->n_slices * 64 + 255) / 256);
  if (esz == 8)
    hipLaunchKernelGGL(bzg::sell_pack_kernel<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream, p->n_rows, p->n_slices,
                       p->sell_off, rowptr, colidx, static_cast<const uint64_t*>(val), p->sell_col,
                       static_cast<uint64_t*>(p->sell_val));
  else
    hipLaunchKernelGGL(bzg::sell_pack_kernel<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream, p->n_rows, p->n_slices,
                       p->sell_off, rowptr, colidx, static_cast<const uint32_t*>(val), p->sell_col,
                       static_cast<uint32_t*>(p->sell_val));
  BZG_LAUNCH_CHECK();
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host offsets go out of scope
  p->sell_dtype = dtype;
  return BZG_OK;
}

bzg_status bzg_dsellmv(bzg_context* ctx, const bzg_csr_plan* plan, const double* x, double* y) {
  return bzg::sellmv<double>(ctx, plan, 0, x, y);
}
bzg_status bzg_ssellmv(bzg_context* ctx, const bzg_csr_plan* plan, const float* x, float* y) {
  return bzg::sellmv<float>(ctx, plan, 1, x, y);
}
bzg_status bzg_isellmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* x, int32_t* y) {
  return bzg::sellmv<int32_t>(ctx, plan, 2, x, y);
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

bzg_status bzg_csr_plan_autotune(bzg_context* ctx, bzg_csr_plan* plan, int32_t fp32, const int32_t* rowptr,
                                 const int32_t* colidx, const void* val, const void* x, void* y, float* ms_out) {
  BZG_REQUIRE(ctx && plan && rowptr && x && y, BZG_ERR_INVALID_ARGUMENT, "csr_plan_autotune: null argument");
  // candidates: row blocks (2048 / 4096 nonzeros, cached / non-temporal loads) and lane groups
  struct Cand { int lanes, rpg, cap; };
  std::vector<Cand> cands = {{0, 1, 0}, {0, 2, 0}, {0, 1, 1}, {0, 2, 1}};
  for (int lanes : {2, 4, 8, 16})
    for (int rpg : {1, 2})
      for (int cap : {4, 8}) cands.push_back({lanes, rpg, cap});
  hipEvent_t e0, e1;
  BZG_HIP_CHECK(hipEventCreate(&e0));
  BZG_HIP_CHECK(hipEventCreate(&e1));
  const bool prof = ctx->profile;
  ctx->profile = false;
  auto run = [&]() -> bzg_status {
    return fp32 ? bzg::csrmv<float>(ctx, plan, rowptr, colidx, static_cast<const float*>(val),
                                    static_cast<const float*>(x), static_cast<float*>(y))
                : bzg::csrmv<double>(ctx, plan, rowptr, colidx, static_cast<const double*>(val),
                                     static_cast<const double*>(x), static_cast<double*>(y));
  };
  int best = -1;
  float best_ms = 0.f;
  bzg_status st = BZG_OK;
  for (int i = 0; i < (int)cands.size() && !st; ++i) {
    st = bzg_csr_plan_tune(plan, cands[i].lanes, cands[i].rpg, cands[i].cap);
    if (st) break;
    st = run();  // warm-up
    const int reps = 5;
    hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps && !st; ++r) st = run();
    hipEventRecord(e1, ctx->stream);
    if (hipEventSynchronize(e1) != hipSuccess) st = BZG_ERR_HIP;
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    if (!st && (best < 0 || ms < best_ms)) {
      best = i;
      best_ms = ms;
    }
  }
  ctx->profile = prof;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st) return st;
  if (ms_out) *ms_out = best_ms;
  return bzg_csr_plan_tune(plan, cands[best].lanes, cands[best].rpg, cands[best].cap);
}

}  // extern "C"
