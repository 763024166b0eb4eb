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
__global__ __launch_bounds__(256) void gemv_kernel(int M, int N, T alpha, const T* __restrict__ A,
                                                   long lda, const T* __restrict__ x, T beta,
                                                   T* __restrict__ y) {
  using AR = Arith<T>;
  constexpr int VW = 16 / sizeof(T);
  constexpr int STEP = VW;  // elements per lane per load
  const int lane_in_group = threadIdx.x % G;
  const long group = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const long row0 = group * R;
  if (row0 >= M) return;  // the whole group exits together: no partial shuffles
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(0);

  if (VEC) {
    const int nvec = N / VW;
    for (int v = lane_in_group; v < nvec; v += G) {
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
    // Tail columns (N not a multiple of VW).
    for (int c = nvec * VW + lane_in_group; c < N; c += G) {
      const T xc = x[c];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row0 + r < M) acc[r] = AR::add(acc[r], AR::mul(A[(row0 + r) * lda + c], xc));
    }
  } else {
    for (int c = lane_in_group; c < N; c += G) {
      const T xc = x[c];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (row0 + r < M) acc[r] = AR::add(acc[r], AR::mul(A[(row0 + r) * lda + c], xc));
    }
  }
  (void)STEP;
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
        T v = AR::mul(alpha, acc[r]);
        if (beta != T(0)) v = AR::add(v, AR::mul(beta, y[row]));
        y[row] = v;
      }
    }
  }
}

template <typename T, int G, int R>
bzg_status launch_gemv(bzg_context* ctx, int m, int n, T alpha, const T* A, long lda, const T* x,
                       T beta, T* y) {
  constexpr int VW = 16 / sizeof(T);
  const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)x % 16 == 0) && (lda % VW == 0);
  const long groups = (m + R - 1) / R;
  const long threads = groups * G;
  const long blocks = (threads + 255) / 256;
  BZG_REQUIRE(blocks < (1L << 31), BZG_ERR_UNSUPPORTED, "gemv: too many rows");
  if (vec)
    hipLaunchKernelGGL((gemv_kernel<T, G, R, true>), dim3((unsigned)blocks), dim3(256), 0,
                       ctx->stream, m, n, alpha, A, lda, x, beta, y);
  else
    hipLaunchKernelGGL((gemv_kernel<T, G, R, false>), dim3((unsigned)blocks), dim3(256), 0,
                       ctx->stream, m, n, alpha, A, lda, x, beta, y);
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
  // more rows per wavefront) for short ones, so no lane idles on a short row.
  constexpr int VW = 16 / sizeof(T);
  if (n >= 64 * VW * 2) return launch_gemv<T, 64, 4>(ctx, m, n, alpha, A, lda, x, beta, y);
  if (n >= 16 * VW * 2) return launch_gemv<T, 16, 2>(ctx, m, n, alpha, A, lda, x, beta, y);
  if (n >= 4 * VW) return launch_gemv<T, 4, 1>(ctx, m, n, alpha, A, lda, x, beta, y);
  return launch_gemv<T, 1, 1>(ctx, m, n, alpha, A, lda, x, beta, y);
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

}  // extern "C"
