// Complex element types for gfx950: GEMM, GEMV and dense solves on interleaved (re, im) storage
// (the layout of std::complex<T> and of C99 T _Complex).
//
// Replaces the complex instantiations of the products and the solver:
//   DMatDMatMultExpr with complex<float> / complex<double> elements, whose BLAS path is
//     cblas_cgemm / cblas_zgemm (blaze/math/blas/cblas/gemm.h:197-283);
//   DMatDVecMultExpr with complex elements -> cblas_cgemv / cblas_zgemv
//     (blaze/math/blas/cblas/gemv.h:186-245);
//   solve() of a complex<double> system -> gesv -> zgesv_ (blaze/math/lapack/clapack/gesv.h:317-333).
//
// GEMM runs on the real MFMA GEMMs (gemm.hip) through the 2 x 2 real form of a complex
// product.  With A' the m x 2k real view of A (columns re, im interleaved: no copy) and
// B'' the 2k x 2n real matrix whose 2 x 2 block (p, j) is [[br, bi], [-bi, br]] for
// b = alpha * op(B)(p, j), the real product A' B'' is exactly the m x 2n real view of
// alpha * A * op(B): C'(i, 2j) = sum_p ar br - ai bi, C'(i, 2j + 1) = sum_p ar bi + ai br.
// That is 8 m n k flops, the flop count of the complex product itself, at the real GEMM's
// rate; the expansion reads op(B) once and writes twice its size.  op(A) = A^T or A^H is
// copied once into an m x k complex buffer (a 32 x 32 tiled transpose through LDS).
//
// GEMV is HBM-bound and has its own kernels: y = alpha A x + beta y with G lanes per row and a
// fixed butterfly; y = alpha op(A) x + beta y for op = ^T, ^H with one column per thread over
// row chunks, then a fixed-order sum of the chunks.
//
// zgesv solves the 2n x 2n real system [[Ar, -Ai], [Ai, Ar]] [xr; xi] = [br; bi] with the
// single-system FP64 LU (getrf.hip).  The embedding is singular exactly when A is (its
// determinant is |det A|^2); partial pivoting on it is backward stable for the real system,
// which bounds the complex residual the same way.  The pivots are those of the embedding,
// not LAPACK zgetrf's, so no ipiv is returned.
#include "common.hpp"

#include <algorithm>

namespace bzg {
namespace {

template <typename R>
struct Cx {
  R re, im;
};
template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

constexpr int kTile = 32;

// Tile (p0, j0) of X = alpha * op(S) into LDS, op(S) being k x n: op 0 reads S (k x n, ld),
// op 1 / 2 read S^T / S^H with S stored n x k.  Loads are coalesced along S's rows either way.
template <typename R>
__device__ __forceinline__ void load_op_tile(R (&tre)[kTile][kTile + 1], R (&tim)[kTile][kTile + 1], int64_t k,
                                             int64_t n, const R* __restrict__ S, int64_t lds, int op, int64_t p0,
                                             int64_t j0) {
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int r = ty; r < kTile; r += 8) {
    if (op == 0) {
      const int64_t p = p0 + r, j = j0 + tx;
      R re = 0, im = 0;
      if (p < k && j < n) {
        re = S[2 * (p * lds + j)];
        im = S[2 * (p * lds + j) + 1];
      }
      tre[r][tx] = re;
      tim[r][tx] = im;
    } else {
      const int64_t j = j0 + r, p = p0 + tx;  // S(j, p) = op(S)(p, j)
      R re = 0, im = 0;
      if (p < k && j < n) {
        re = S[2 * (j * lds + p)];
        im = S[2 * (j * lds + p) + 1];
        if (op == 2) im = -im;
      }
      tre[tx][r] = re;
      tim[tx][r] = im;
    }
  }
}

// B'' (2k x 2n real, row-major, leading dimension 2n) from alpha * op(B).
template <typename R>
__global__ __launch_bounds__(256) void expand_b_kernel(int64_t k, int64_t n, const R* __restrict__ B, int64_t ldb,
                                                       int op, R ar, R ai, R* __restrict__ out) {
  __shared__ R tre[kTile][kTile + 1], tim[kTile][kTile + 1];
  const int64_t p0 = (int64_t)blockIdx.y * kTile, j0 = (int64_t)blockIdx.x * kTile;
  load_op_tile<R>(tre, tim, k, n, B, ldb, op, p0, j0);
  __syncthreads();
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t ld = 2 * n;
  for (int r = ty; r < kTile; r += 8) {
    const int64_t p = p0 + r, j = j0 + tx;
    if (p < k && j < n) {
      const Cx<R> b = cmul<R>({ar, ai}, {tre[r][tx], tim[r][tx]});
      R* row0 = out + (2 * p) * ld + 2 * j;
      R* row1 = row0 + ld;
      row0[0] = b.re;
      row0[1] = b.im;
      row1[0] = -b.im;
      row1[1] = b.re;
    }
  }
}

// D (k x n complex, leading dimension n) = op(S) for op 1 / 2 (S stored n x k).
template <typename R>
__global__ __launch_bounds__(256) void op_copy_kernel(int64_t k, int64_t n, const R* __restrict__ S, int64_t lds,
                                                      int op, R* __restrict__ D) {
  __shared__ R tre[kTile][kTile + 1], tim[kTile][kTile + 1];
  const int64_t p0 = (int64_t)blockIdx.y * kTile, j0 = (int64_t)blockIdx.x * kTile;
  load_op_tile<R>(tre, tim, k, n, S, lds, op, p0, j0);
  __syncthreads();
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < kTile; r += 8) {
    const int64_t p = p0 + r, j = j0 + tx;
    if (p < k && j < n) {
      D[2 * (p * n + j)] = tre[r][tx];
      D[2 * (p * n + j) + 1] = tim[r][tx];
    }
  }
}

// C := beta * C (m x n complex, leading dimension ldc); beta = 0 writes zeros without reading C.
template <typename R>
__global__ __launch_bounds__(256) void scale_kernel(int64_t m, int64_t n, R br, R bi, R* __restrict__ C,
                                                    int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t - i * n;
    R* c = C + 2 * (i * ldc + j);
    Cx<R> v = {0, 0};
    if (br != R(0) || bi != R(0)) v = cmul<R>({br, bi}, {c[0], c[1]});
    c[0] = v.re;
    c[1] = v.im;
  }
}

// y = alpha A x + beta y, A m x n (lda); G lanes per row, each walking the row with stride G,
// partial sums combined by a fixed butterfly.  beta = 0 never reads y.
template <typename R, int G>
__global__ __launch_bounds__(256) void zgemv_n_kernel(int64_t m, int64_t n, R ar, R ai, const R* __restrict__ A,
                                                      int64_t lda, const R* __restrict__ x, R br, R bi,
                                                      R* __restrict__ y) {
  const int lig = threadIdx.x % G;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (row >= m) return;  // the whole group leaves together
  const R* a = A + 2 * row * lda;
  // four elements per lane and step, loads issued before the arithmetic, one partial sum per
  // slot (added in slot order afterwards)
  R pre[4] = {0, 0, 0, 0}, pim[4] = {0, 0, 0, 0};
  int64_t j = lig;
  for (; j + 3 * G < n; j += 4 * G) {
    R xr[4], xi[4], are[4], aim[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t c = j + u * G;
      are[u] = a[2 * c];
      aim[u] = a[2 * c + 1];
      xr[u] = x[2 * c];
      xi[u] = x[2 * c + 1];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      pre[u] += are[u] * xr[u] - aim[u] * xi[u];
      pim[u] += are[u] * xi[u] + aim[u] * xr[u];
    }
  }
  for (int u = 0; j < n; j += G, ++u) {
    const R xr = x[2 * j], xi = x[2 * j + 1];
    const R are = a[2 * j], aim = a[2 * j + 1];
    pre[u] += are * xr - aim * xi;
    pim[u] += are * xi + aim * xr;
  }
  R sre = (pre[0] + pre[1]) + (pre[2] + pre[3]), sim = (pim[0] + pim[1]) + (pim[2] + pim[3]);
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) {
    sre += __shfl_xor(sre, off, 64);
    sim += __shfl_xor(sim, off, 64);
  }
  if (lig == 0) {
    Cx<R> v = cmul<R>({ar, ai}, {sre, sim});
    if (br != R(0) || bi != R(0)) {
      const Cx<R> t = cmul<R>({br, bi}, {y[2 * row], y[2 * row + 1]});
      v.re += t.re;
      v.im += t.im;
    }
    y[2 * row] = v.re;
    y[2 * row + 1] = v.im;
  }
}

// Partial sums of op(A) x for op = ^T / ^H over rows [c RC, c RC + RC): part[c][j], one
// column j per thread (reads of a row are contiguous across the workgroup).
constexpr int kRowChunk = 256;

template <typename R>
__global__ __launch_bounds__(256) void zgemv_t_part_kernel(int64_t m, int64_t n, const R* __restrict__ A,
                                                           int64_t lda, const R* __restrict__ x, int conj,
                                                           R* __restrict__ part) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t i0 = (int64_t)blockIdx.y * kRowChunk, i1 = (m < i0 + kRowChunk) ? m : i0 + kRowChunk;
  const R sg = conj ? R(-1) : R(1);
  R sre = 0, sim = 0;
#pragma unroll 8
  for (int64_t i = i0; i < i1; ++i) {
    const R are = A[2 * (i * lda + j)], aim = sg * A[2 * (i * lda + j) + 1];
    const R xr = x[2 * i], xi = x[2 * i + 1];
    sre += are * xr - aim * xi;
    sim += are * xi + aim * xr;
  }
  part[2 * ((int64_t)blockIdx.y * n + j)] = sre;
  part[2 * ((int64_t)blockIdx.y * n + j) + 1] = sim;
}

template <typename R>
__global__ __launch_bounds__(256) void zgemv_t_sum_kernel(int64_t n, int64_t chunks, R ar, R ai,
                                                          const R* __restrict__ part, R br, R bi,
                                                          R* __restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  R sre = 0, sim = 0;
  for (int64_t c = 0; c < chunks; ++c) {
    sre += part[2 * (c * n + j)];
    sim += part[2 * (c * n + j) + 1];
  }
  Cx<R> v = cmul<R>({ar, ai}, {sre, sim});
  if (br != R(0) || bi != R(0)) {
    const Cx<R> t = cmul<R>({br, bi}, {y[2 * j], y[2 * j + 1]});
    v.re += t.re;
    v.im += t.im;
  }
  y[2 * j] = v.re;
  y[2 * j + 1] = v.im;
}

// The real 2n x 2n embedding M = [[Ar, -Ai], [Ai, Ar]] (row-major, leading dimension 2n) of
// the n x n complex A (lda), and [br; bi] (2n x nrhs, leading dimension nrhs) of B (ldb).
__global__ __launch_bounds__(256) void embed_a_kernel(int64_t n, const double* __restrict__ A, int64_t lda,
                                                      double* __restrict__ M) {
  const int64_t total = n * n, ld = 2 * n;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / n, j = t - i * n;
    const double re = A[2 * (i * lda + j)], im = A[2 * (i * lda + j) + 1];
    M[i * ld + j] = re;
    M[i * ld + n + j] = -im;
    M[(n + i) * ld + j] = im;
    M[(n + i) * ld + n + j] = re;
  }
}

__global__ __launch_bounds__(256) void embed_b_kernel(int64_t n, int64_t nrhs, const double* __restrict__ B,
                                                      int64_t ldb, double* __restrict__ Bm) {
  const int64_t total = n * nrhs;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / nrhs, r = t - i * nrhs;
    Bm[i * nrhs + r] = B[2 * (i * ldb + r)];
    Bm[(n + i) * nrhs + r] = B[2 * (i * ldb + r) + 1];
  }
}

__global__ __launch_bounds__(256) void unembed_b_kernel(int64_t n, int64_t nrhs, const double* __restrict__ Bm,
                                                        double* __restrict__ B, int64_t ldb) {
  const int64_t total = n * nrhs;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / nrhs, r = t - i * nrhs;
    B[2 * (i * ldb + r)] = Bm[i * nrhs + r];
    B[2 * (i * ldb + r) + 1] = Bm[(n + i) * nrhs + r];
  }
}

__global__ void set_info_kernel(int32_t* info, int32_t v) { *info = v; }

unsigned grid_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
}

// Workspace slots of this file (distinct from every other file's)
constexpr int kWsExpand = 15;  // B'' / the embedding of a solve
constexpr int kWsOpA = 1;      // op(A) copy / GEMV partial sums / the embedding's pivots

template <typename R>
bzg_status real_gemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, R alpha, const R* A, int64_t lda,
                     const R* B, int64_t ldb, R beta, R* C, int64_t ldc);
template <>
bzg_status real_gemm<double>(bzg_context* ctx, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  return bzg_dgemm(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
template <>
bzg_status real_gemm<float>(bzg_context* ctx, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                            int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc) {
  return bzg_sgemm(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

template <typename R>
bzg_status complex_gemm(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        const R* alpha, const R* A, int64_t lda, const R* B, int64_t ldb, const R* beta, R* C,
                        int64_t ldc) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "complex gemm: null context");
  BZG_REQUIRE(transa >= 0 && transa <= 2 && transb >= 0 && transb <= 2, BZG_ERR_INVALID_ARGUMENT,
              "complex gemm: trans must be 0 (N), 1 (T) or 2 (C)");
  BZG_REQUIRE(m >= 0 && n >= 0 && k >= 0, BZG_ERR_INVALID_ARGUMENT, "complex gemm: negative size");
  BZG_REQUIRE(alpha && beta, BZG_ERR_INVALID_ARGUMENT, "complex gemm: null alpha or beta");
  if (m == 0 || n == 0) return BZG_OK;
  BZG_REQUIRE(C != nullptr && ldc >= n, BZG_ERR_INVALID_ARGUMENT, "complex gemm: bad C or ldc");
  const R ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
  const bool no_product = k == 0 || (ar == R(0) && ai == R(0));
  if (!no_product) {
    BZG_REQUIRE(A && B, BZG_ERR_INVALID_ARGUMENT, "complex gemm: null A or B");
    BZG_REQUIRE(lda >= (transa ? m : k) && ldb >= (transb ? k : n), BZG_ERR_INVALID_ARGUMENT,
                "complex gemm: leading dimension smaller than the stored rows");
    BZG_REQUIRE((k + kTile - 1) / kTile < 65536 && (m + kTile - 1) / kTile < 65536, BZG_ERR_UNSUPPORTED,
                "complex gemm: k or m beyond 2,097,120");
  }
  ProfScope prof(ctx, "complex_gemm");
  // C := beta C first when beta is not real (or no product is added); a real beta goes to the
  // real GEMM, which scales both halves of every entry by it
  R rbeta = br;
  if (no_product || bi != R(0)) {
    hipLaunchKernelGGL(scale_kernel<R>, dim3(grid_for(m * n)), dim3(256), 0, ctx->stream, m, n, br, bi, C, ldc);
    BZG_LAUNCH_CHECK();
    if (no_product) return BZG_OK;
    rbeta = R(1);
  }
  void* ws = nullptr;
  bzg_status st = workspace(ctx, kWsExpand, (size_t)4 * k * n * sizeof(R), &ws);
  if (st) return st;
  R* Bx = static_cast<R*>(ws);
  hipLaunchKernelGGL(expand_b_kernel<R>, dim3((unsigned)((n + kTile - 1) / kTile), (unsigned)((k + kTile - 1) / kTile)),
                     dim3(256), 0, ctx->stream, k, n, B, ldb, (int)transb, ar, ai, Bx);
  BZG_LAUNCH_CHECK();
  const R* Ap = A;
  int64_t ldap = 2 * lda;
  if (transa) {
    void* wa = nullptr;
    st = workspace(ctx, kWsOpA, (size_t)2 * m * k * sizeof(R), &wa);
    if (st) return st;
    hipLaunchKernelGGL(op_copy_kernel<R>, dim3((unsigned)((k + kTile - 1) / kTile), (unsigned)((m + kTile - 1) / kTile)),
                       dim3(256), 0, ctx->stream, m, k, A, lda, (int)transa, static_cast<R*>(wa));
    BZG_LAUNCH_CHECK();
    Ap = static_cast<const R*>(wa);
    ldap = 2 * k;
  }
  return real_gemm<R>(ctx, m, 2 * n, 2 * k, R(1), Ap, ldap, Bx, 2 * n, rbeta, C, 2 * ldc);
}

template <typename R>
bzg_status complex_gemv(bzg_context* ctx, int32_t trans, int64_t m, int64_t n, const R* alpha, const R* A,
                        int64_t lda, const R* x, const R* beta, R* y) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "complex gemv: null context");
  BZG_REQUIRE(trans >= 0 && trans <= 2, BZG_ERR_INVALID_ARGUMENT, "complex gemv: trans must be 0, 1 or 2");
  BZG_REQUIRE(m >= 0 && n >= 0, BZG_ERR_INVALID_ARGUMENT, "complex gemv: negative size");
  BZG_REQUIRE(alpha && beta, BZG_ERR_INVALID_ARGUMENT, "complex gemv: null alpha or beta");
  const int64_t ylen = trans ? n : m, xlen = trans ? m : n;
  if (ylen == 0) return BZG_OK;
  BZG_REQUIRE(y != nullptr, BZG_ERR_INVALID_ARGUMENT, "complex gemv: null y");
  const R ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
  ProfScope prof(ctx, "complex_gemv");
  if (xlen == 0 || (ar == R(0) && ai == R(0))) {
    hipLaunchKernelGGL(scale_kernel<R>, dim3(grid_for(ylen)), dim3(256), 0, ctx->stream, ylen, (int64_t)1, br, bi, y,
                       (int64_t)1);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  BZG_REQUIRE(A && x && lda >= n, BZG_ERR_INVALID_ARGUMENT, "complex gemv: null A or x, or lda < n");
  if (!trans) {
    if (n >= 256) {
      const int64_t blocks = (m * 64 + 255) / 256;
      hipLaunchKernelGGL((zgemv_n_kernel<R, 64>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, m, n, ar, ai, A,
                         lda, x, br, bi, y);
    } else {
      const int64_t blocks = (m * 8 + 255) / 256;
      hipLaunchKernelGGL((zgemv_n_kernel<R, 8>), dim3((unsigned)blocks), dim3(256), 0, ctx->stream, m, n, ar, ai, A,
                         lda, x, br, bi, y);
    }
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  const int64_t chunks = (m + kRowChunk - 1) / kRowChunk;
  BZG_REQUIRE(chunks < 65536, BZG_ERR_UNSUPPORTED, "complex gemv: more than 16,776,960 rows");
  void* ws = nullptr;
  bzg_status st = workspace(ctx, kWsOpA, (size_t)2 * chunks * n * sizeof(R), &ws);
  if (st) return st;
  R* part = static_cast<R*>(ws);
  hipLaunchKernelGGL(zgemv_t_part_kernel<R>, dim3((unsigned)((n + 255) / 256), (unsigned)chunks), dim3(256), 0,
                     ctx->stream, m, n, A, lda, x, trans == 2 ? 1 : 0, part);
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(zgemv_t_sum_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n, chunks,
                     ar, ai, part, br, bi, y);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

}  // namespace
}  // namespace bzg

extern "C" {

bzg_status bzg_zgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        const double* alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                        const double* beta, double* C, int64_t ldc) {
  return bzg::complex_gemm<double>(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

bzg_status bzg_cgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        const float* alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                        const float* beta, float* C, int64_t ldc) {
  return bzg::complex_gemm<float>(ctx, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

bzg_status bzg_zgemv_ex(bzg_context* ctx, int32_t trans, int64_t m, int64_t n, const double* alpha,
                        const double* A, int64_t lda, const double* x, const double* beta, double* y) {
  return bzg::complex_gemv<double>(ctx, trans, m, n, alpha, A, lda, x, beta, y);
}

bzg_status bzg_cgemv_ex(bzg_context* ctx, int32_t trans, int64_t m, int64_t n, const float* alpha,
                        const float* A, int64_t lda, const float* x, const float* beta, float* y) {
  return bzg::complex_gemv<float>(ctx, trans, m, n, alpha, A, lda, x, beta, y);
}

bzg_status bzg_ztranspose(bzg_context* ctx, int64_t m, int64_t n, const double* A, int64_t lda, double* B,
                          int64_t ldb) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "ztranspose: null context");
  BZG_REQUIRE(m >= 0 && n >= 0, BZG_ERR_INVALID_ARGUMENT, "ztranspose: negative size");
  if (m == 0 || n == 0) return BZG_OK;
  BZG_REQUIRE(A && B && lda >= n && ldb >= m, BZG_ERR_INVALID_ARGUMENT, "ztranspose: bad pointer or leading dimension");
  BZG_REQUIRE(A != B, BZG_ERR_INVALID_ARGUMENT, "ztranspose: not in place");
  BZG_REQUIRE((n + bzg::kTile - 1) / bzg::kTile < 65536, BZG_ERR_UNSUPPORTED, "ztranspose: n beyond 2,097,120");
  // B (n x m, ldb) = A^T: op_copy writes op(S) with S = A stored m x n, as an n x m matrix of
  // leading dimension equal to its width; copy row blocks when ldb != m
  if (ldb == m) {
    hipLaunchKernelGGL(bzg::op_copy_kernel<double>, dim3((unsigned)((m + bzg::kTile - 1) / bzg::kTile),
                       (unsigned)((n + bzg::kTile - 1) / bzg::kTile)), dim3(256), 0, ctx->stream, n, m, A, lda, 1, B);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  void* ws = nullptr;
  bzg_status st = bzg::workspace(ctx, bzg::kWsOpA, (size_t)2 * m * n * sizeof(double), &ws);
  if (st) return st;
  hipLaunchKernelGGL(bzg::op_copy_kernel<double>, dim3((unsigned)((m + bzg::kTile - 1) / bzg::kTile),
                     (unsigned)((n + bzg::kTile - 1) / bzg::kTile)), dim3(256), 0, ctx->stream, n, m, A, lda, 1,
                     static_cast<double*>(ws));
  BZG_LAUNCH_CHECK();
  BZG_HIP_CHECK(hipMemcpy2DAsync(B, (size_t)ldb * 16, ws, (size_t)m * 16, (size_t)m * 16, (size_t)n,
                                 hipMemcpyDeviceToDevice, ctx->stream));
  return BZG_OK;
}

bzg_status bzg_zgesv(bzg_context* ctx, int64_t n, int64_t nrhs, const double* A, int64_t lda, double* B,
                     int64_t ldb, int32_t* info) {
  BZG_REQUIRE(ctx != nullptr && info != nullptr, BZG_ERR_INVALID_ARGUMENT, "zgesv: null context or info");
  BZG_REQUIRE(n >= 0 && nrhs >= 0, BZG_ERR_INVALID_ARGUMENT, "zgesv: negative size");
  BZG_REQUIRE(n <= 16384, BZG_ERR_UNSUPPORTED, "zgesv: order beyond 16384 (embedding beyond 32768)");
  if (n == 0) {
    hipLaunchKernelGGL(bzg::set_info_kernel, dim3(1), dim3(1), 0, ctx->stream, info, 0);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  BZG_REQUIRE(A && lda >= n, BZG_ERR_INVALID_ARGUMENT, "zgesv: null A or lda < n");
  BZG_REQUIRE(nrhs == 0 || (B && ldb >= nrhs), BZG_ERR_INVALID_ARGUMENT, "zgesv: null B or ldb < nrhs");
  const int64_t n2 = 2 * n;
  void* ws = nullptr;
  bzg_status st = bzg::workspace(ctx, bzg::kWsExpand, (size_t)(n2 * n2 + n2 * std::max<int64_t>(nrhs, 1)) * sizeof(double), &ws);
  if (st) return st;
  double* M = static_cast<double*>(ws);
  double* Bm = M + n2 * n2;
  void* wp = nullptr;
  st = bzg::workspace(ctx, bzg::kWsOpA, (size_t)n2 * sizeof(int32_t), &wp);
  if (st) return st;
  int32_t* ipiv = static_cast<int32_t*>(wp);
  bzg::ProfScope prof(ctx, "zgesv");
  hipLaunchKernelGGL(bzg::embed_a_kernel, dim3(bzg::grid_for(n * n)), dim3(256), 0, ctx->stream, n, A, lda, M);
  BZG_LAUNCH_CHECK();
  if (nrhs > 0) {
    hipLaunchKernelGGL(bzg::embed_b_kernel, dim3(bzg::grid_for(n * nrhs)), dim3(256), 0, ctx->stream, n, nrhs, B, ldb,
                       Bm);
    BZG_LAUNCH_CHECK();
  }
  st = bzg_dgesv(ctx, n2, nrhs, M, n2, ipiv, Bm, std::max<int64_t>(nrhs, 1), info);
  if (st) return st;
  if (nrhs > 0) {
    // B keeps its values when the system is singular (dgesv leaves [br; bi] unchanged then)
    hipLaunchKernelGGL(bzg::unembed_b_kernel, dim3(bzg::grid_for(n * nrhs)), dim3(256), 0, ctx->stream, n, nrhs, Bm, B,
                       ldb);
    BZG_LAUNCH_CHECK();
  }
  return BZG_OK;
}

}  // extern "C"
