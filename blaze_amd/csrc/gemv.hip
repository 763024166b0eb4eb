// Dense GEMV for gfx950:  y = alpha * A * x + beta * y, row-major A.
//
// Replaces the assignment kernels of DMatDVecMultExpr (blaze/math/expressions/
// DMatDVecMultExpr.h:408-458 dispatch, :528-850 small kernel, :887-1114 large kernel).
// HBM-bound: A is streamed exactly once with 16-byte loads; x is re-read from L2.
//
// A group of G lanes owns R consecutive rows.  Every lane walks the row with a stride of
// G*VW elements (VW = 2 for 16-byte loads of double, 4 for float/int32), keeping one partial
// sum per row; the G partials are combined by a fixed butterfly, so the result is identical
// run to run.  One load of x is reused for the R rows, which cuts L2 traffic for x by R.
#include "common.hpp"

#include <type_traits>

#include <algorithm>

namespace bzg {
namespace {

typedef int i4v __attribute__((ext_vector_type(4)));

template <typename T>
struct Arith {
  __device__ static T mul(T a, T b) { return a * b; }
  __device__ static T add(T a, T b) { return a + b; }
};
template <>
struct Arith<int32_t> {  // wrap-around arithmetic, like the unsigned reinterpretation
  __device__ static int32_t mul(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a * (uint32_t)b);
  }
  __device__ static int32_t add(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a + (uint32_t)b);
  }
};

template <typename T, int G, int R, bool VEC>
__global__ __launch_bounds__(256) void gemv_kernel(int M, int N, int KC, T alpha, const T* __restrict__ A,
                                                   long lda, const T* __restrict__ x, T beta,
                                                   T* __restrict__ y, T* __restrict__ part) {
  // Columns [c0, c1) of the rows: the whole row when the grid has one column chunk, else
  // chunk blockIdx.y of KC columns (KC a multiple of VW), whose sums go to part[chunk][row].
  using AR = Arith<T>;
  constexpr int VW = 16 / sizeof(T);
  const int lane_in_group = threadIdx.x % G;
  const long group = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const long row0 = group * R;
  if (row0 >= M) return;  // the whole group exits together: no partial shuffles
  const long c0l = (long)blockIdx.y * KC;
  const int c0 = (int)c0l;
  const int c1 = (int)min((long)N, c0l + KC);
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(0);

  int cs = c0;  // first column of the scalar pass
  if (VEC) {
    const int vend = c1 / VW;  // c0 is a multiple of VW
    for (int v = c0 / VW + lane_in_group; v < vend; v += G) {
      T xv[VW];
      *reinterpret_cast<i4v*>(xv) = *reinterpret_cast<const i4v*>(x + (long)v * VW);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (row0 + r < M) {
          T av[VW];
          *reinterpret_cast<i4v*>(av) =
              __builtin_nontemporal_load(reinterpret_cast<const i4v*>(A + (row0 + r) * lda + (long)v * VW));
#pragma unroll
          for (int e = 0; e < VW; ++e) acc[r] = AR::add(acc[r], AR::mul(av[e], xv[e]));
        }
      }
    }
    cs = vend * VW;  // tail columns (N not a multiple of VW)
  }
  for (int c = cs + lane_in_group; c < c1; c += G) {
    const T xc = x[c];
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row0 + r < M) acc[r] = AR::add(acc[r], AR::mul(A[(row0 + r) * lda + c], xc));
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) acc[r] = AR::add(acc[r], __shfl_xor(acc[r], off, 64));
  }
  if (lane_in_group == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long row = row0 + r;
      if (row < M) {
        if (gridDim.y > 1) {
          part[(long)blockIdx.y * M + row] = acc[r];
        } else {
          T v = AR::mul(alpha, acc[r]);
          if (beta != T(0)) v = AR::add(v, AR::mul(beta, y[row]));
          y[row] = v;
        }
      }
    }
  }
}

// y[j] = alpha * (sum of the P partials part[r][j]) + beta * y[j]; beta == 0 never reads y
// (BLAS's convention).  A workgroup is CW columns by L = 256 / CW lanes: lane l adds the
// partials r = l, l + L, ... in order, then the L lane sums meet in a fixed pairwise tree
// through LDS.  The shape (CW, L) depends on N and P only, so results are identical run to run.
template <typename T>
__global__ __launch_bounds__(256) void partials_reduce_kernel(int N, int P, int cw_log2, T alpha,
                                                              const T* __restrict__ part, T beta,
                                                              T* __restrict__ y) {
  using AR = Arith<T>;
  __shared__ T red[256];
  const int CW = 1 << cw_log2, L = 256 >> cw_log2;
  const int c = threadIdx.x & (CW - 1), l = threadIdx.x >> cw_log2;
  const long j = (long)blockIdx.x * CW + c;
  T s = T(0);
  if (j < N) {
#pragma unroll 8
    for (int r = l; r < P; r += L) s = AR::add(s, part[(long)r * N + j]);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = L / 2; h >= 1; h >>= 1) {
    if (l < h) red[threadIdx.x] = AR::add(red[threadIdx.x], red[threadIdx.x + h * CW]);
    __syncthreads();
  }
  if (l == 0 && j < N) {
    s = red[c];
    y[j] = (beta == T(0)) ? AR::mul(alpha, s) : AR::add(AR::mul(alpha, s), AR::mul(beta, y[j]));
  }
}

// Launch of partials_reduce_kernel: about 32 partials per lane.
template <typename T>
void launch_partials_reduce(bzg_context* ctx, int N, int P, T alpha, const T* part, T beta, T* y) {
  int l_log2 = 0;
  while (l_log2 < 8 && (32 << l_log2) < P) ++l_log2;
  const int cw_log2 = 8 - l_log2;
  const int cw = 1 << cw_log2;
  hipLaunchKernelGGL(partials_reduce_kernel<T>, dim3((unsigned)((N + cw - 1) / cw)), dim3(256), 0, ctx->stream, N,
                     P, cw_log2, alpha, part, beta, y);
}

constexpr int kGemvTargetGroups = 2048;  // workgroups that keep every CU streaming
constexpr int kGemvMaxParts = 2048;      // partial sums per output entry

template <typename T, int G, int R>
bzg_status launch_gemv(bzg_context* ctx, int m, int n, T alpha, const T* A, long lda, const T* x,
                       T beta, T* y) {
  constexpr int VW = 16 / sizeof(T);
  const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)x % 16 == 0) && (lda % VW == 0);
  const long groups = (m + R - 1) / R;
  const long threads = groups * G;
  const long blocks = (threads + 255) / 256;
  BZG_REQUIRE(blocks < (1L << 31), BZG_ERR_UNSUPPORTED, "gemv: too many rows");
  // Few rows, long rows: split the columns into chunks of at least 2048 (a multiple of a
  // group's G*VW stride) so the grid still fills the chip; the chunk sums are then added in
  // chunk order.  The split depends on m and n only, so results are identical run to run.
  int kc = n, ks = 1;
  if (blocks < kGemvTargetGroups / 2 && n >= 4096) {
    const long want = std::min<long>({(long)kGemvMaxParts, (kGemvTargetGroups + blocks - 1) / blocks, n / 2048L});
    if (want > 1) {
      const long q = (long)G * VW;
      kc = (int)(((n + want - 1) / want + q - 1) / q * q);
      ks = (n + kc - 1) / kc;
    }
  }
  T* part = nullptr;
  if (ks > 1) {
    void* w = nullptr;
    bzg_status st = workspace(ctx, 11, (size_t)ks * m * sizeof(T), &w);
    if (st) return st;
    part = static_cast<T*>(w);
  }
  const dim3 grid((unsigned)blocks, (unsigned)ks);
  if (vec)
    hipLaunchKernelGGL((gemv_kernel<T, G, R, true>), grid, dim3(256), 0, ctx->stream, m, n, kc, alpha, A, lda, x,
                       beta, y, part);
  else
    hipLaunchKernelGGL((gemv_kernel<T, G, R, false>), grid, dim3(256), 0, ctx->stream, m, n, kc, alpha, A, lda, x,
                       beta, y, part);
  if (ks > 1) launch_partials_reduce<T>(ctx, m, ks, alpha, part, beta, y);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

template <typename T>
__global__ void gemv_scale_kernel(int M, T beta, T* y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) y[i] = (beta == T(0)) ? T(0) : Arith<T>::mul(beta, y[i]);
}

template <typename T>
bzg_status gemv_dispatch(bzg_context* ctx, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                         const T* x, T beta, T* y) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemv: null context");
  BZG_REQUIRE(m >= 0 && n >= 0, BZG_ERR_INVALID_ARGUMENT, "gemv: negative size");
  BZG_REQUIRE(m < (1LL << 31) && n < (1LL << 31), BZG_ERR_UNSUPPORTED, "gemv: dimension >= 2^31");
  BZG_REQUIRE(lda >= (n > 0 ? n : 1), BZG_ERR_INVALID_ARGUMENT, "gemv: lda < n");
  if (m == 0) return BZG_OK;
  BZG_REQUIRE(y != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemv: null y");
  if (n == 0) {
    hipLaunchKernelGGL(gemv_scale_kernel<T>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                       ctx->stream, (int)m, beta, y);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  BZG_REQUIRE(A != nullptr && x != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemv: null A or x");
  // Lanes per row from the row length: a full wavefront for long rows, fewer lanes (and
  // more rows per wavefront) for short ones, so no lane idles on a short row.  Long rows of
  // double run two rows per wavefront (twice the workgroups; 6.9 against 6.2 TB/s with four
  // at 16384^2), 4-byte types four (wide FP32 matrices lose up to 12 % with two).
  constexpr int VW = 16 / sizeof(T);
  constexpr int RLONG = sizeof(T) == 8 ? 2 : 4;
  if (n >= 64 * VW * 2) return launch_gemv<T, 64, RLONG>(ctx, m, n, alpha, A, lda, x, beta, y);
  if (n >= 16 * VW * 2) return launch_gemv<T, 16, 2>(ctx, m, n, alpha, A, lda, x, beta, y);
  if (n >= 4 * VW) return launch_gemv<T, 4, 1>(ctx, m, n, alpha, A, lda, x, beta, y);
  return launch_gemv<T, 1, 1>(ctx, m, n, alpha, A, lda, x, beta, y);
}

// Transposed GEMV, y = alpha * A^T * x + beta * y for a row-major m x n A (y has n entries):
// the column-major form of the product (TDMatDVecMultExpr).  Stage 1: a workgroup is CT
// column threads by RL = 256 / CT row lanes; column thread c owns VW adjacent columns (one
// 16-byte load per row when aligned), workgroup (cx, ry) covers columns
// [cx CT VW, (cx + 1) CT VW) and, in row lane l, the rows (ry RL + l) + k RY RL.  x[i] is the
// same for every column thread of a row lane.  Each thread keeps its column sums in row order;
// the RL row lanes are added in lane order through LDS and written to part[ry][j].  Stage 2
// adds the RY partials of column j in index order, so the result is identical run to run.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void gemv_t_partial_kernel(int M, int N, const T* __restrict__ A, long lda,
                                                             const T* __restrict__ x, T* __restrict__ part,
                                                             int ct_log2) {
  using AR = Arith<T>;
  constexpr int VW = 16 / sizeof(T);
  __shared__ T red[256 * VW];
  const int CT = 1 << ct_log2, RL = 256 >> ct_log2;
  const int ct = threadIdx.x & (CT - 1), rl = threadIdx.x >> ct_log2;
  const long j = ((long)blockIdx.x * CT + ct) * VW;
  const long step = (long)gridDim.y * RL;
  T s[VW];
#pragma unroll
  for (int e = 0; e < VW; ++e) s[e] = T(0);
  if (j < N) {
    if (VEC && j + VW <= N) {
#pragma unroll 8
      for (long i = (long)blockIdx.y * RL + rl; i < M; i += step) {
        const T xi = x[i];
        T av[VW];
        *reinterpret_cast<i4v*>(av) = __builtin_nontemporal_load(reinterpret_cast<const i4v*>(A + i * lda + j));
#pragma unroll
        for (int e = 0; e < VW; ++e) s[e] = AR::add(s[e], AR::mul(av[e], xi));
      }
    } else {
      const int w = (int)min((long)VW, N - j);
#pragma unroll 4
      for (long i = (long)blockIdx.y * RL + rl; i < M; i += step) {
        const T xi = x[i];
        const T* row = A + i * lda + j;
#pragma unroll
        for (int e = 0; e < VW; ++e)
          if (e < w) s[e] = AR::add(s[e], AR::mul(row[e], xi));
      }
    }
  }
  if (RL > 1) {
#pragma unroll
    for (int e = 0; e < VW; ++e) red[threadIdx.x * VW + e] = s[e];
    __syncthreads();
    if (rl == 0) {
      for (int l = 1; l < RL; ++l)
#pragma unroll
        for (int e = 0; e < VW; ++e) s[e] = AR::add(s[e], red[(l * CT + ct) * VW + e]);
    }
  }
  if (rl == 0) {
#pragma unroll
    for (int e = 0; e < VW; ++e)
      if (j + e < N) part[(long)blockIdx.y * N + j + e] = s[e];
  }
}

template <typename T>
bzg_status gemv_t_dispatch(bzg_context* ctx, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                           const T* x, T beta, T* y) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemv: null context");
  BZG_REQUIRE(m >= 0 && n >= 0, BZG_ERR_INVALID_ARGUMENT, "gemv: negative size");
  BZG_REQUIRE(m < (1LL << 31) && n < (1LL << 31), BZG_ERR_UNSUPPORTED, "gemv: dimension >= 2^31");
  BZG_REQUIRE(lda >= (n > 0 ? n : 1), BZG_ERR_INVALID_ARGUMENT, "gemv: lda < n");
  if (n == 0) return BZG_OK;
  BZG_REQUIRE(y != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemv: null y");
  if (m == 0) {
    hipLaunchKernelGGL(gemv_scale_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, (int)n, beta, y);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  BZG_REQUIRE(A != nullptr && x != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemv: null A or x");
  constexpr int VW = 16 / sizeof(T);
  const bool vec = ((uintptr_t)A % 16 == 0) && (lda % VW == 0);
  // column threads: the next power of two covering the row, at most 256 (the workgroup)
  int ct_log2 = 0;
  while (ct_log2 < 8 && ((int64_t)VW << ct_log2) < n) ++ct_log2;
  const int64_t ct = 1 << ct_log2, rl = 256 / ct;
  const int64_t cx = (n + ct * VW - 1) / (ct * VW);
  // row partitions: enough workgroups to fill the chip, at least 16 rows per row lane
  const int64_t ry = std::max<int64_t>(1, std::min<int64_t>({(int64_t)kGemvMaxParts, (m + 16 * rl - 1) / (16 * rl),
                                                            (kGemvTargetGroups + cx - 1) / cx}));
  BZG_REQUIRE(cx < (1LL << 31), BZG_ERR_UNSUPPORTED, "gemv: too many columns");
  void* w = nullptr;
  bzg_status st = workspace(ctx, 11, (size_t)ry * n * sizeof(T), &w);
  if (st) return st;
  T* part = static_cast<T*>(w);
  const dim3 grid((unsigned)cx, (unsigned)ry);
  if (vec)
    hipLaunchKernelGGL((gemv_t_partial_kernel<T, true>), grid, dim3(256), 0, ctx->stream, (int)m, (int)n, A,
                       (long)lda, x, part, ct_log2);
  else
    hipLaunchKernelGGL((gemv_t_partial_kernel<T, false>), grid, dim3(256), 0, ctx->stream, (int)m, (int)n, A,
                       (long)lda, x, part, ct_log2);
  launch_partials_reduce<T>(ctx, (int)n, (int)ry, alpha, part, beta, y);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

// B = A^T for a row-major m x n A (lda) into n x m B (ldb): 64 x 64 tiles through LDS (one
// padding element per row), reads and writes both along rows.
__global__ __launch_bounds__(256) void transpose_kernel(int64_t m, int64_t n, const double* __restrict__ A,
                                                        int64_t lda, double* __restrict__ B, int64_t ldb) {
  __shared__ double tile[64][65];
  const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  for (int r = ty; r < 64; r += 4) {
    const int64_t i = i0 + r, j = j0 + tx;
    if (i < m && j < n) tile[r][tx] = A[i * lda + j];
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int64_t j = j0 + r, i = i0 + tx;
    if (j < n && i < m) B[j * ldb + i] = tile[tx][r];
  }
}

}  // namespace
}  // namespace bzg

namespace bzg {
namespace {
// y += alpha x, element by element (integers wrap modulo 2^32, as the INT32 products do)
template <typename T>
__global__ __launch_bounds__(256) void axpy_kernel(long n, T alpha, const T* __restrict__ x, T* __restrict__ y) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    if constexpr (std::is_integral<T>::value)
      y[i] = (T)((uint32_t)y[i] + (uint32_t)alpha * (uint32_t)x[i]);
    else
      y[i] = y[i] + alpha * x[i];
  }
}

template <typename T>
bzg_status axpy(bzg_context* ctx, int64_t n, T alpha, const T* x, T* y) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "axpy: null context");
  BZG_REQUIRE(n >= 0, BZG_ERR_INVALID_ARGUMENT, "axpy: negative size");
  if (n == 0) return BZG_OK;
  BZG_REQUIRE(x && y, BZG_ERR_INVALID_ARGUMENT, "axpy: null pointer");
  const long blocks = std::min<long>((n + 255) / 256, (long)ctx->num_cus * 8);
  hipLaunchKernelGGL(axpy_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (long)n, alpha, x, y);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}
}  // namespace
}  // namespace bzg

extern "C" {

bzg_status bzg_dtranspose(bzg_context* ctx, int64_t m, int64_t n, const double* A, int64_t lda, double* B,
                          int64_t ldb) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "transpose: null context");
  BZG_REQUIRE(m >= 0 && n >= 0, BZG_ERR_INVALID_ARGUMENT, "transpose: negative size");
  if (m == 0 || n == 0) return BZG_OK;
  BZG_REQUIRE(A && B, BZG_ERR_INVALID_ARGUMENT, "transpose: null pointer");
  BZG_REQUIRE(lda >= n && ldb >= m, BZG_ERR_INVALID_ARGUMENT, "transpose: leading dimension too small");
  BZG_REQUIRE((m + 63) / 64 < 65536, BZG_ERR_UNSUPPORTED, "transpose: too many rows");
  BZG_REQUIRE(static_cast<const void*>(A) != static_cast<const void*>(B), BZG_ERR_INVALID_ARGUMENT,
              "transpose: in-place transposition is not supported");
  hipLaunchKernelGGL(bzg::transpose_kernel, dim3((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64)), dim3(256),
                     0, ctx->stream, m, n, A, lda, B, ldb);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

bzg_status bzg_dgemv_t(bzg_context* ctx, int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                       const double* x, double beta, double* y) {
  return bzg::gemv_t_dispatch<double>(ctx, m, n, alpha, A, lda, x, beta, y);
}
bzg_status bzg_sgemv_t(bzg_context* ctx, int64_t m, int64_t n, float alpha, const float* A, int64_t lda,
                       const float* x, float beta, float* y) {
  return bzg::gemv_t_dispatch<float>(ctx, m, n, alpha, A, lda, x, beta, y);
}
bzg_status bzg_igemv_t(bzg_context* ctx, int64_t m, int64_t n, int32_t alpha, const int32_t* A, int64_t lda,
                       const int32_t* x, int32_t beta, int32_t* y) {
  return bzg::gemv_t_dispatch<int32_t>(ctx, m, n, alpha, A, lda, x, beta, y);
}

bzg_status bzg_dgemv(bzg_context* ctx, int64_t m, int64_t n, double alpha, const double* A,
                     int64_t lda, const double* x, double beta, double* y) {
  return bzg::gemv_dispatch<double>(ctx, m, n, alpha, A, lda, x, beta, y);
}
bzg_status bzg_sgemv(bzg_context* ctx, int64_t m, int64_t n, float alpha, const float* A,
                     int64_t lda, const float* x, float beta, float* y) {
  return bzg::gemv_dispatch<float>(ctx, m, n, alpha, A, lda, x, beta, y);
}
bzg_status bzg_igemv(bzg_context* ctx, int64_t m, int64_t n, int32_t alpha, const int32_t* A,
                     int64_t lda, const int32_t* x, int32_t beta, int32_t* y) {
  return bzg::gemv_dispatch<int32_t>(ctx, m, n, alpha, A, lda, x, beta, y);
}

bzg_status bzg_daxpy(bzg_context* ctx, int64_t n, double alpha, const double* x, double* y) {
  return bzg::axpy<double>(ctx, n, alpha, x, y);
}
bzg_status bzg_saxpy(bzg_context* ctx, int64_t n, float alpha, const float* x, float* y) {
  return bzg::axpy<float>(ctx, n, alpha, x, y);
}
bzg_status bzg_iaxpy(bzg_context* ctx, int64_t n, int32_t alpha, const int32_t* x, int32_t* y) {
  return bzg::axpy<int32_t>(ctx, n, alpha, x, y);
}

}  // extern "C"
