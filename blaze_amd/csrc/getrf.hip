// Single-system dense LU with partial pivoting for any order (n <= 32768), its solves, the
// determinant and the inverse from the factors.
//
// Replaces, for one system of any size:
//   solve(A, x, b) / solve(A, X, B)  (blaze/math/dense/LSE.h:6343-6373 -> solveNxN, LSE.h:5229-5257
//     -> gesv -> dgesv_, blaze/math/lapack/clapack/gesv.h:207)       -> bzg_dgesv
//   getrf (blaze/math/lapack/clapack/getrf.h:190-197), lu() (blaze/math/dense/LU.h:98-140)
//                                                                     -> bzg_dgetrf
//   getrs (blaze/math/lapack/clapack/getrs.h:207-215)                 -> bzg_dgetrs
//   det() for N > 6 (blaze/math/expressions/DMatDetExpr.h:324-357)    -> bzg_ddet_lu
//   invert()/inv() by LU (blaze/math/dense/Inversion.h:3337-3347)     -> bzg_dgetri
// The batched kernels (lu.hip) keep serving orders <= 512 with <= 8 right-hand sides.
//
// Matrices are row-major (the math matrix A, leading dimension lda); P A = L U with L unit
// lower, U upper, ipiv 1-based (row i was interchanged with row ipiv[i] - 1, in order) and
// info the 1-based column of the first exactly-zero pivot: the results of LAPACK dgetrf on A.
//
// Factorisation: the recursive column split of LAPACK dgetrf2 (left half, U12 = L11^-1 A12,
// A22 -= L21 U12 through bzg_dgemm on the FP64 matrix cores, right half).  The leaves are
// column blocks of width <= 8 on the rows from the diagonal down, factored by one 512-thread
// workgroup that holds the block in registers (row t + 512 i in thread t): per column a
// wavefront max-reduction of (|a|, -row), one workgroup barrier, the interchange and the
// rank-1 update in registers.  A leaf's interchanges then move the whole rows (all other
// columns: the at most 16 affected rows, one column per thread over the device), which is the
// net effect of dgetrf2's row swaps on both sides, so no per-level swap pass runs.  The triangular solves (dtrsm
// semantics, Left/NoTrans) split recursively down to 32 x 32 diagonal blocks solved by one
// thread per right-hand-side column; the off-diagonal blocks go through bzg_dgemm.
#include "common.hpp"

#include <cfloat>
#include <climits>
#include <mutex>

namespace bzg {
namespace big {

constexpr int kLeafThreads = 512;
constexpr int kTrsmLeaf = 32;
constexpr int64_t kMaxOrder = 32768;

// ---- leaf panel factorisation ----------------------------------------------------------------
// Columns d0 .. d0 + w - 1 (w <= WL) on rows d0 .. n - 1.  R = rows per thread (m <= 512 R).
template <int R, int WL>
__global__ __launch_bounds__(kLeafThreads) void getf2_leaf_kernel(double* __restrict__ A, long lda, int n, int d0,
                                                                   int w, int32_t* __restrict__ ipiv,
                                                                   int32_t* __restrict__ info,
                                                                   int32_t* __restrict__ swp) {
  constexpr int NW = kLeafThreads / 64;
  __shared__ double s_v[2][NW];
  __shared__ int s_r[2][NW];
  __shared__ double s_row[2][NW][WL];
  __shared__ double s_jrow[2][WL];
  __shared__ int s_piv[WL];
  __shared__ long s_pos[2 * WL], s_src[2 * WL];
  __shared__ int s_cnt;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = n - d0;
  double* P = A + (long)d0 * lda + d0;
  // a full-width row of 4 or 8 columns moves as 32-byte vectors (fewer load instructions: what
  // limits one workgroup's share of the memory bandwidth), others element by element
  double a[R][WL];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int row = t + kLeafThreads * i;
    if (WL % 4 == 0 && row < m && w == WL) {
#pragma unroll
      for (int c4 = 0; c4 < WL / 4; ++c4) {
        const dvec4u v = *reinterpret_cast<const dvec4u*>(P + (long)row * lda + 4 * c4);
        a[i][4 * c4] = v[0]; a[i][4 * c4 + 1] = v[1]; a[i][4 * c4 + 2] = v[2]; a[i][4 * c4 + 3] = v[3];
      }
    } else {
#pragma unroll
      for (int c = 0; c < WL; ++c) a[i][c] = (row < m && c < w) ? P[(long)row * lda + c] : 0.0;
    }
  }
  int first_zero = 0;
#pragma unroll
  for (int j = 0; j < WL; ++j) {
    if (j < w) {
      const int par = j & 1;
      // the thread's best candidate (rows >= j): |a|, a NaN below every number; rows grow with
      // i, so the strict comparison keeps the first
      double bv = -2.0;
      int br = INT_MAX;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int row = t + kLeafThreads * i;
        if (row >= j && row < m) {
          const double x = a[i][j];
          const double v = (x != x) ? -1.0 : fabs(x);
          if (v > bv) { bv = v; br = row; }
        }
      }
      // wavefront maximum of (v, -row): every lane ends with the same pair
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int orow = __shfl_xor(br, off, 64);
        if (ov > bv || (ov == bv && orow < br)) { bv = ov; br = orow; }
      }
      if (br < m && br % kLeafThreads == t) {  // the wavefront's best row: its values to LDS
        const int ib = br / kLeafThreads;
#pragma unroll
        for (int i = 0; i < R; ++i)
          if (i == ib) {
#pragma unroll
            for (int c = 0; c < WL; ++c) s_row[par][wave][c] = a[i][c];
          }
      }
      if (lane == 0) { s_v[par][wave] = bv; s_r[par][wave] = br; }
      if (t == j) {  // row j is held by thread j (first slot)
#pragma unroll
        for (int c = 0; c < WL; ++c) s_jrow[par][c] = a[0][c];
      }
      __syncthreads();
      double gv = s_v[par][0];
      int gr = s_r[par][0], gw = 0;
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        const double v = s_v[par][q];
        const int r = s_r[par][q];
        if (v > gv || (v == gv && r < gr)) { gv = v; gr = r; gw = q; }
      }
      // idamax keeps a NaN at the diagonal position as the pivot
      const bool jnan = s_jrow[par][j] != s_jrow[par][j];
      const int prow = jnan ? j : gr;
      double pv[WL];
#pragma unroll
      for (int c = 0; c < WL; ++c) pv[c] = jnan ? s_jrow[par][c] : s_row[par][gw][c];
      const double piv = pv[j];
      const bool nz = piv != 0.0;
      if (t == 0) s_piv[j] = prow;
      if (nz && prow != j) {  // interchange rows j and prow (panel columns, in registers)
        if (t == j) {
#pragma unroll
          for (int c = 0; c < WL; ++c) a[0][c] = pv[c];
        }
        if (prow % kLeafThreads == t) {
          const int ip = prow / kLeafThreads;
#pragma unroll
          for (int i = 0; i < R; ++i)
            if (i == ip) {
#pragma unroll
              for (int c = 0; c < WL; ++c) a[i][c] = s_jrow[par][c];
            }
        }
      }
      // scale by the reciprocal (divide when |pivot| < DBL_MIN, dgetf2's sfmin rule), then the
      // rank-1 update of the columns to the right (applied for a zero pivot too, as dger is)
      const double rinv = 1.0 / piv;
      const bool tiny = fabs(piv) < DBL_MIN;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const int row = t + kLeafThreads * i;
        if (row > j && row < m) {
          double l = a[i][j];
          if (nz) {
            l = tiny ? l / piv : l * rinv;
            a[i][j] = l;
          }
#pragma unroll
          for (int c = j + 1; c < WL; ++c) a[i][c] -= l * pv[c];
        }
      }
      if (!nz && first_zero == 0) first_zero = d0 + j + 1;
    }
  }
  __syncthreads();
  // net permutation of the block's interchanges over the affected rows (<= 2 w of them)
  if (t == 0) {
    int cnt = 0;
    for (int j = 0; j < w; ++j) {
      ipiv[d0 + j] = d0 + s_piv[j] + 1;
      const long ra = d0 + j, rb = d0 + s_piv[j];
      if (ra == rb) continue;
      int ia = -1, ib = -1;
      for (int k = 0; k < cnt; ++k) {
        if (s_pos[k] == ra) ia = k;
        if (s_pos[k] == rb) ib = k;
      }
      if (ia < 0) { s_pos[cnt] = ra; s_src[cnt] = ra; ia = cnt++; }
      if (ib < 0) { s_pos[cnt] = rb; s_src[cnt] = rb; ib = cnt++; }
      const long tmp = s_src[ia];
      s_src[ia] = s_src[ib];
      s_src[ib] = tmp;
    }
    s_cnt = cnt;
    if (first_zero && *info == 0) *info = first_zero;
  }
  __syncthreads();
  // the other columns of the affected rows (row s_pos[k] takes the old row s_src[k]) are moved
  // by row_swap_kernel, spread over the device: the list goes out here
  if (t == 0) swp[0] = s_cnt;
  if (t < s_cnt) {
    swp[1 + 2 * t] = (int32_t)s_pos[t];
    swp[2 + 2 * t] = (int32_t)s_src[t];
  }
  // the block's rows, at their final positions
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int row = t + kLeafThreads * i;
    if (WL % 4 == 0 && row < m && w == WL) {
#pragma unroll
      for (int c4 = 0; c4 < WL / 4; ++c4)
        *reinterpret_cast<dvec4u*>(P + (long)row * lda + 4 * c4) =
            dvec4u{a[i][4 * c4], a[i][4 * c4 + 1], a[i][4 * c4 + 2], a[i][4 * c4 + 3]};
    } else if (row < m) {
#pragma unroll
      for (int c = 0; c < WL; ++c)
        if (c < w) P[(long)row * lda + c] = a[i][c];
    }
  }
}

// The leaf's interchanges on every column outside its block: row pos[k] takes the old row
// src[k] (k < cnt <= 2 WL), one column per thread (the rows' loads are contiguous across a
// wavefront).  This is the net effect of dgetrf2's row swaps on both sides of the block.
template <int WL>
__global__ __launch_bounds__(256) void row_swap_kernel(double* __restrict__ A, long lda, int n, int d0, int w,
                                                       const int32_t* __restrict__ swp) {
  const int cnt = swp[0];
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cnt == 0 || col >= n - w) return;
  const long c = col < d0 ? col : col + w;
  double v[2 * WL];
#pragma unroll
  for (int k = 0; k < 2 * WL; ++k)
    if (k < cnt) v[k] = A[(long)swp[2 + 2 * k] * lda + c];
#pragma unroll
  for (int k = 0; k < 2 * WL; ++k)
    if (k < cnt) A[(long)swp[1 + 2 * k] * lda + c] = v[k];
}

// ---- triangular solves on 32 x 32 diagonal blocks ---------------------------------------------
// B (nb x ncols, nb <= 32) := T^-1 B, T lower unit (UPPER = false) or upper non-unit, one
// thread per column; the operation order of reference dtrsm (Left, NoTrans): for each k in
// order, skip a zero B(k), scale by the diagonal (upper), subtract B(k) T(i, k) from the rest.
template <bool UPPER>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(const double* __restrict__ T, long ldt, int nb,
                                                        double* __restrict__ B, long ldb, long ncols) {
  __shared__ double sT[kTrsmLeaf][kTrsmLeaf + 1];
  for (int e = threadIdx.x; e < nb * nb; e += 256) sT[e / nb][e % nb] = T[(long)(e / nb) * ldt + e % nb];
  __syncthreads();
  const long col = (long)blockIdx.x * 256 + threadIdx.x;
  if (col >= ncols) return;
  double x[kTrsmLeaf];
#pragma unroll
  for (int i = 0; i < kTrsmLeaf; ++i) x[i] = i < nb ? B[(long)i * ldb + col] : 0.0;
  if (!UPPER) {
#pragma unroll
    for (int k = 0; k < kTrsmLeaf; ++k) {
      if (k < nb && x[k] != 0.0) {
#pragma unroll
        for (int i = k + 1; i < kTrsmLeaf; ++i)
          if (i < nb) x[i] -= x[k] * sT[i][k];
      }
    }
  } else {
#pragma unroll
    for (int k = kTrsmLeaf - 1; k >= 0; --k) {
      if (k < nb && x[k] != 0.0) {
        x[k] = x[k] / sT[k][k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] -= x[k] * sT[i][k];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kTrsmLeaf; ++i)
    if (i < nb) B[(long)i * ldb + col] = x[i];
}

// ---- row permutation of the right-hand sides ---------------------------------------------------
// perm[i] = the original row that ends at position i after the interchanges ipiv[0..n) in
// order (one workgroup; the sequence runs in LDS, n <= kMaxOrder).
__global__ __launch_bounds__(1024) void ipiv_to_perm_kernel(int n, const int32_t* __restrict__ ipiv,
                                                            int32_t* __restrict__ perm) {
  extern __shared__ int sp[];
  for (int i = threadIdx.x; i < n; i += 1024) sp[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) {
      const int p = ipiv[i] - 1;
      if (p != i) {
        const int tmp = sp[i];
        sp[i] = sp[p];
        sp[p] = tmp;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 1024) perm[i] = sp[i];
}

// W[i][c] = B[perm[i]][c] (identity = true: W = P I, the permuted identity, n x n)
__global__ __launch_bounds__(256) void gather_rows_kernel(int n, long ncols, const int32_t* __restrict__ perm,
                                                          const double* __restrict__ B, long ldb,
                                                          double* __restrict__ W, long ldw, bool identity) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)n * ncols) return;
  const long i = e / ncols, c = e % ncols;
  const int src = perm[i];
  W[i * ldw + c] = identity ? (src == c ? 1.0 : 0.0) : B[(long)src * ldb + c];
}

__global__ __launch_bounds__(256) void copy_rows_kernel(int n, long ncols, const double* __restrict__ W, long ldw,
                                                        double* __restrict__ B, long ldb) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)n * ncols) return;
  const long i = e / ncols, c = e % ncols;
  B[i * ldb + c] = W[i * ldw + c];
}

__global__ void zero_info_kernel(int32_t* info) { *info = 0; }

// detNxN's rule from the factors: 0 for a zero pivot, else the product, left to right from 1,
// of U's diagonal with the sign flipped at every interchange.
__global__ void det_kernel(int n, const double* __restrict__ A, long lda, const int32_t* __restrict__ ipiv,
                           const int32_t* __restrict__ info, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (*info > 0) {
    *out = 0.0;
    return;
  }
  double d = 1.0;
  for (int i = 0; i < n; ++i) {
    d *= A[(long)i * lda + i];
    if (ipiv[i] != i + 1) d = -d;
  }
  *out = d;
}

// ---- host recursion ------------------------------------------------------------------------
// rows per thread (a power of two) and leaf width for a panel of m rows
static void leaf_shape(int64_t m, int* R, int* WL) {
  int r = 1;
  while ((int64_t)r * kLeafThreads < m) r *= 2;
  *R = r;
  *WL = r <= 8 ? 8 : 64 / r;
}

static bzg_status launch_leaf(bzg_context* ctx, double* A, long lda, int n, int d0, int w, int32_t* ipiv,
                              int32_t* info) {
  int R, WL;
  leaf_shape(n - d0, &R, &WL);
  // the leaf's interchange list (count, then (row, source row) pairs) in workspace slot 16
  void* sw = nullptr;
  bzg_status st = workspace(ctx, 16, 256, &sw);
  if (st) return st;
  int32_t* swp = static_cast<int32_t*>(sw);
  const unsigned sgrid = (unsigned)std::max<long>(1, ((long)n - w + 255) / 256);
#define BZG_LEAF(RR, WW)                                                                                   \
  if (R == RR) {                                                                                           \
    hipLaunchKernelGGL((getf2_leaf_kernel<RR, WW>), dim3(1), dim3(kLeafThreads), 0, ctx->stream, A, lda, n, d0, \
                       w, ipiv, info, swp);                                                                \
    BZG_LAUNCH_CHECK();                                                                                    \
    if (n > w) {                                                                                           \
      hipLaunchKernelGGL((row_swap_kernel<WW>), dim3(sgrid), dim3(256), 0, ctx->stream, A, lda, n, d0, w, swp); \
      BZG_LAUNCH_CHECK();                                                                                  \
    }                                                                                                      \
    return BZG_OK;                                                                                         \
  }
  BZG_LEAF(1, 8)
  BZG_LEAF(2, 8)
  BZG_LEAF(4, 8)
  BZG_LEAF(8, 8)
  BZG_LEAF(16, 4)
  BZG_LEAF(32, 2)
  BZG_LEAF(64, 1)
#undef BZG_LEAF
  set_error("getrf: panel taller than 32768 rows");
  return BZG_ERR_UNSUPPORTED;
}

// B (nb x ncols) := T^-1 B with T the nb x nb diagonal block at T (lower unit or upper).
static bzg_status trsm(bzg_context* ctx, bool upper, const double* T, long ldt, int nb, double* B, long ldb,
                       long ncols) {
  if (nb <= 0 || ncols <= 0) return BZG_OK;
  if (nb <= kTrsmLeaf) {
    const unsigned grid = (unsigned)((ncols + 255) / 256);
    if (upper)
      hipLaunchKernelGGL(trsm_leaf_kernel<true>, dim3(grid), dim3(256), 0, ctx->stream, T, ldt, nb, B, ldb, ncols);
    else
      hipLaunchKernelGGL(trsm_leaf_kernel<false>, dim3(grid), dim3(256), 0, ctx->stream, T, ldt, nb, B, ldb, ncols);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  int h = ((nb / 2 + kTrsmLeaf - 1) / kTrsmLeaf) * kTrsmLeaf;
  if (h >= nb) h = nb - kTrsmLeaf;
  bzg_status st;
  if (!upper) {
    // [L11 0; L21 L22]: X1 = L11^-1 B1, B2 -= L21 X1, X2 = L22^-1 B2
    if ((st = trsm(ctx, false, T, ldt, h, B, ldb, ncols))) return st;
    if ((st = bzg_dgemm(ctx, nb - h, ncols, h, -1.0, T + (long)h * ldt, ldt, B, ldb, 1.0, B + (long)h * ldb, ldb)))
      return st;
    return trsm(ctx, false, T + (long)h * ldt + h, ldt, nb - h, B + (long)h * ldb, ldb, ncols);
  }
  // [U11 U12; 0 U22]: X2 = U22^-1 B2, B1 -= U12 X2, X1 = U11^-1 B1
  if ((st = trsm(ctx, true, T + (long)h * ldt + h, ldt, nb - h, B + (long)h * ldb, ldb, ncols))) return st;
  if ((st = bzg_dgemm(ctx, h, ncols, nb - h, -1.0, T + h, ldt, B + (long)h * ldb, ldb, 1.0, B, ldb))) return st;
  return trsm(ctx, true, T, ldt, h, B, ldb, ncols);
}

// Factors columns d0 .. d0 + nc - 1 on rows d0 .. n - 1 (dgetrf2's recursion).
static bzg_status rgetrf(bzg_context* ctx, double* A, long lda, int n, int d0, int nc, int32_t* ipiv,
                         int32_t* info) {
  int R, WL;
  leaf_shape(n - d0, &R, &WL);
  if (nc <= WL) return launch_leaf(ctx, A, lda, n, d0, nc, ipiv, info);
  int n1 = ((nc / 2 + 7) / 8) * 8;  // multiples of 8 keep the sub-blocks 16-byte aligned
  if (n1 >= nc) n1 = nc / 2;
  if (n1 < 1) n1 = 1;
  const int n2 = nc - n1, m = n - d0;
  bzg_status st;
  if ((st = rgetrf(ctx, A, lda, n, d0, n1, ipiv, info))) return st;
  double* A11 = A + (long)d0 * lda + d0;
  // U12 = L11^-1 A12, then A22 -= L21 U12
  if ((st = trsm(ctx, false, A11, lda, n1, A11 + n1, lda, n2))) return st;
  if ((st = bzg_dgemm(ctx, m - n1, n2, n1, -1.0, A11 + (long)n1 * lda, lda, A11 + n1, lda, 1.0,
                      A11 + (long)n1 * lda + n1, lda)))
    return st;
  return rgetrf(ctx, A, lda, n, d0 + n1, n2, ipiv, info);
}

static bzg_status check_square(bzg_context* ctx, int64_t n, const void* A, int64_t lda, const char* what) {
  if (!ctx) {
    set_error(std::string(what) + ": null context");
    return BZG_ERR_INVALID_ARGUMENT;
  }
  if (n < 0 || lda < (n > 0 ? n : 1)) {
    set_error(std::string(what) + ": bad order or leading dimension");
    return BZG_ERR_INVALID_ARGUMENT;
  }
  if (n > kMaxOrder) {
    set_error(std::string(what) + ": order above 32768");
    return BZG_ERR_UNSUPPORTED;
  }
  if (n > 0 && !A) {
    set_error(std::string(what) + ": null matrix");
    return BZG_ERR_INVALID_ARGUMENT;
  }
  return BZG_OK;
}

// The permutation kernel's LDS (4 bytes per row) beyond 64 KB needs the attribute raised once per
// device.
static std::mutex g_perm_mu;
static bool g_perm_attr[64] = {};

static bzg_status perm_attr(bzg_context* ctx) {
  if (ctx->device < 0 || ctx->device >= 64) return BZG_OK;
  std::lock_guard<std::mutex> lock(g_perm_mu);
  if (g_perm_attr[ctx->device]) return BZG_OK;
  BZG_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&ipiv_to_perm_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kMaxOrder * sizeof(int))));
  g_perm_attr[ctx->device] = true;
  return BZG_OK;
}

// perm (n ints) and W (n x ncols doubles, ld = ncols) from the context's workspaces.
static bzg_status solve_scratch(bzg_context* ctx, int n, long ncols, int32_t** perm, double** W) {
  void *p = nullptr, *w = nullptr;
  bzg_status st = workspace(ctx, 13, (size_t)n * sizeof(int32_t), &p);
  if (!st) st = workspace(ctx, 14, (size_t)n * (size_t)ncols * sizeof(double), &w);
  if (st) return st;
  *perm = static_cast<int32_t*>(p);
  *W = static_cast<double*>(w);
  return perm_attr(ctx);
}

// X = A^-1 B from the factors: W = P B (or P I), forward and back substitution in W.
static bzg_status getrs_into_w(bzg_context* ctx, int n, long ncols, const double* LU, long lda,
                               const int32_t* ipiv, const double* B, long ldb, bool identity, double** Wout) {
  int32_t* perm = nullptr;
  double* W = nullptr;
  bzg_status st = solve_scratch(ctx, n, ncols, &perm, &W);
  if (st) return st;
  hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3(1), dim3(1024), (size_t)n * sizeof(int), ctx->stream, n, ipiv, perm);
  BZG_LAUNCH_CHECK();
  const long total = (long)n * ncols;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, n, ncols,
                     perm, B, ldb, W, ncols, identity);
  BZG_LAUNCH_CHECK();
  if ((st = trsm(ctx, false, LU, lda, n, W, ncols, ncols))) return st;
  if ((st = trsm(ctx, true, LU, lda, n, W, ncols, ncols))) return st;
  *Wout = W;
  return BZG_OK;
}

}  // namespace big
}  // namespace bzg

extern "C" {

bzg_status bzg_dgetrf(bzg_context* ctx, int64_t n, double* A, int64_t lda, int32_t* ipiv, int32_t* info) {
  bzg_status st = bzg::big::check_square(ctx, n, A, lda, "dgetrf");
  if (st) return st;
  BZG_REQUIRE(info != nullptr && (n == 0 || ipiv != nullptr), BZG_ERR_INVALID_ARGUMENT, "dgetrf: null ipiv or info");
  bzg::ProfScope prof(ctx, "dgetrf");
  hipLaunchKernelGGL(bzg::big::zero_info_kernel, dim3(1), dim3(1), 0, ctx->stream, info);
  BZG_LAUNCH_CHECK();
  if (n == 0) return BZG_OK;
  return bzg::big::rgetrf(ctx, A, (long)lda, (int)n, 0, (int)n, ipiv, info);
}

bzg_status bzg_dgetrs(bzg_context* ctx, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                      const int32_t* ipiv, double* B, int64_t ldb) {
  bzg_status st = bzg::big::check_square(ctx, n, A, lda, "dgetrs");
  if (st) return st;
  BZG_REQUIRE(nrhs >= 0 && ldb >= (nrhs > 0 ? nrhs : 1), BZG_ERR_INVALID_ARGUMENT, "dgetrs: bad nrhs or ldb");
  if (n == 0 || nrhs == 0) return BZG_OK;
  BZG_REQUIRE(ipiv != nullptr && B != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgetrs: null ipiv or B");
  bzg::ProfScope prof(ctx, "dgetrs");
  double* W = nullptr;
  if ((st = bzg::big::getrs_into_w(ctx, (int)n, (long)nrhs, A, (long)lda, ipiv, B, (long)ldb, false, &W))) return st;
  const long total = (long)n * nrhs;
  hipLaunchKernelGGL(bzg::big::copy_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                     (int)n, (long)nrhs, W, (long)nrhs, B, (long)ldb);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

bzg_status bzg_dgesv(bzg_context* ctx, int64_t n, int64_t nrhs, double* A, int64_t lda, int32_t* ipiv, double* B,
                     int64_t ldb, int32_t* info) {
  bzg_status st = bzg::big::check_square(ctx, n, A, lda, "dgesv");
  if (st) return st;
  BZG_REQUIRE(nrhs >= 0 && ldb >= (nrhs > 0 ? nrhs : 1), BZG_ERR_INVALID_ARGUMENT, "dgesv: bad nrhs or ldb");
  if ((st = bzg_dgetrf(ctx, n, A, lda, ipiv, info))) return st;
  if (n == 0 || nrhs == 0) return BZG_OK;
  // dgesv leaves B unchanged when the factorisation reports a zero pivot
  int32_t h_info = 0;
  BZG_HIP_CHECK(hipMemcpyAsync(&h_info, info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (h_info != 0) return BZG_OK;
  return bzg_dgetrs(ctx, n, nrhs, A, lda, ipiv, B, ldb);
}

bzg_status bzg_dgetri(bzg_context* ctx, int64_t n, double* A, int64_t lda, const int32_t* ipiv) {
  bzg_status st = bzg::big::check_square(ctx, n, A, lda, "dgetri");
  if (st) return st;
  if (n == 0) return BZG_OK;
  BZG_REQUIRE(ipiv != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgetri: null ipiv");
  bzg::ProfScope prof(ctx, "dgetri");
  double* W = nullptr;
  if ((st = bzg::big::getrs_into_w(ctx, (int)n, (long)n, A, (long)lda, ipiv, nullptr, 0, true, &W))) return st;
  const long total = (long)n * n;
  hipLaunchKernelGGL(bzg::big::copy_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                     (int)n, (long)n, W, (long)n, A, (long)lda);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

bzg_status bzg_ddet_lu(bzg_context* ctx, int64_t n, const double* A, int64_t lda, const int32_t* ipiv,
                       const int32_t* info, double* det) {
  bzg_status st = bzg::big::check_square(ctx, n, A, lda, "ddet_lu");
  if (st) return st;
  BZG_REQUIRE(info != nullptr && det != nullptr && (n == 0 || ipiv != nullptr), BZG_ERR_INVALID_ARGUMENT,
              "ddet_lu: null ipiv, info or det");
  hipLaunchKernelGGL(bzg::big::det_kernel, dim3(1), dim3(64), 0, ctx->stream, (int)n, A, (long)lda, ipiv, info, det);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

}  // extern "C"
