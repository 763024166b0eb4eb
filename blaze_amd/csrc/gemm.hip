// Dense GEMM for gfx950:  C = alpha * A * B + beta * C, row-major operands.
//
// Replaces the assignment kernels of DMatDMatMultExpr (blaze/math/expressions/
// DMatDMatMultExpr.h:516-569 dispatch, :864-1887 small kernel, MMM.h:105-600 blocked kernel).
//
// FP64 / FP32: MFMA 16x16x4 tiles (v_mfma_f64_16x16x4f64, v_mfma_f32_16x16x4f32).  A workgroup
// of 4 wavefronts computes a BM x BN tile of C; each wave owns a WM x WN sub-tile held in
// MFMA accumulators.  A and B K-slabs of depth BK are staged through LDS with two buffers:
// global loads of slab t+1 are issued into registers before the MFMAs of slab t and written
// to the idle LDS buffer afterwards, so one barrier per slab suffices.
// Every C element is produced by one wave in a fixed k order (no split-K, no atomics), so the
// result is identical run to run and independent of the launch shape chosen for other calls.
//
// INT32: no integer MFMA takes 32-bit operands; a VALU kernel with LDS tiles and 4x4 register
// micro-tiles in wrap-around uint32 arithmetic (bit-exact against any summation order).
#include "common.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace bzg {
namespace {

constexpr int kThreads = 256;
constexpr int kBK = 16;
constexpr int kPadA = 1;   // A slab stored [BM][BK+1]: conflict-free fragment reads
constexpr int kPadB = 16;  // B slab stored [BK][BN+16]

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  using acc_t = d4;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // Accumulator register i of lane l holds row 4*i + l/16, column l%16 (measured on gfx950).
  __device__ static int row(int i, int lane) { return 4 * i + lane / 16; }
};

template <>
struct Mfma<float> {
  using acc_t = f4;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // Accumulator register i of lane l holds row 4*(l/16) + i, column l%16.
  __device__ static int row(int i, int lane) { return 4 * (lane / 16) + i; }
};

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

// Maps the linear workgroup id to a (tile_m, tile_n) pair.  Consecutive workgroup ids are
// dispatched round-robin over the 8 XCDs; the remap gives each XCD a contiguous range of
// logical tiles, and logical tiles are walked in column-major groups of 8 tile rows so the
// workgroups resident on one XCD share A row panels and B column panels in that XCD's L2.
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  int logical = bid;
  if (nt >= 64 && (nt % 8) == 0) logical = (bid % 8) * (nt / 8) + bid / 8;
  constexpr int G = 8;
  const int group = logical / (G * tiles_n);
  const int first_m = group * G;
  const int gsize = min(tiles_m - first_m, G);
  const int in_group = logical % (G * tiles_n);
  tm = first_m + in_group % gsize;
  tn = in_group / gsize;
}

template <typename T, int BM, int BN>
struct GemmShm {
  T As[2][BM * (kBK + kPadA)];
  T Bs[2][kBK * (BN + kPadB)];
};

// One BM x BN tile (tm, tn) of C = alpha A B + beta C by the 4 wavefronts of the workgroup.
// Every element of C is summed in the same k order (slabs of kBK, MFMA k-steps of 4) whatever
// the tile shape, so all tile shapes give bit-identical results.
// TA / TB: A (B) is given transposed, i.e. stored k x m (n x k) row-major with leading
// dimension lda >= m (ldb >= k); the slabs are then read along m (k) and transposed on the
// way into LDS, so the MFMA loop and its summation order are the same for all four forms.
template <typename T, int BM, int BN, int WM, int WN, bool VEC, bool TA = false, bool TB = false>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, T alpha, const T* __restrict__ A,
                                          long lda, const T* __restrict__ B, long ldb, T beta,
                                          T* __restrict__ C, long ldc, int tm, int tn,
                                          GemmShm<T, BM, BN>& sh) {
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  using V2 = typename Vec2<T>::type;
  constexpr int WAVES_N = BN / WN;
  constexpr int FM = WM / 16;  // MFMA blocks per wave along m
  constexpr int FN = WN / 16;  // MFMA blocks per wave along n
  constexpr int LDA_S = kBK + kPadA;
  constexpr int LDB_S = BN + kPadB;
  constexpr int NA = BM * kBK / (2 * kThreads);  // 2-element vectors per thread, A slab
  constexpr int NB = kBK * BN / (2 * kThreads);  // 2-element vectors per thread, B slab
  constexpr int A_VPR = kBK / 2;                 // vectors per A slab row
  constexpr int A_RPP = kThreads / A_VPR;        // A rows per pass
  constexpr int B_VPR = BN / 2;
  constexpr int B_RPP = kThreads / B_VPR;
  constexpr int AT_VPR = BM / 2;                 // transposed A: vectors per stored k row
  constexpr int AT_RPP = kThreads / AT_VPR;
  constexpr int BT_VPR = kBK / 2;                // transposed B: vectors per stored n row
  constexpr int BT_RPP = kThreads / BT_VPR;
  static_assert(NA >= 1 && NB >= 1, "tile too small for 256 threads");
  static_assert((BM / WM) * (BN / WN) == 4, "four wavefronts per workgroup");

  auto& As = sh.As;
  auto& Bs = sh.Bs;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / WAVES_N) * WM;
  const int wn0 = (wave % WAVES_N) * WN;

  const int m0 = tm * BM;
  const int n0 = tn * BN;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

  V2 ra[NA], rb[NB];

  // thread-local (row, column) of vector i in the stored slab
  auto a_rc = [&](int i, int& r, int& c) {
    if (TA) { r = tid / AT_VPR + i * AT_RPP; c = (tid % AT_VPR) * 2; }  // r: k, c: m
    else    { r = tid / A_VPR + i * A_RPP;   c = (tid % A_VPR) * 2; }   // r: m, c: k
  };
  auto b_rc = [&](int i, int& r, int& c) {
    if (TB) { r = tid / BT_VPR + i * BT_RPP; c = (tid % BT_VPR) * 2; }  // r: n, c: k
    else    { r = tid / B_VPR + i * B_RPP;   c = (tid % B_VPR) * 2; }   // r: k, c: n
  };
  auto load_slab = [&](int k0) {
    const bool full = (m0 + BM <= M) && (n0 + BN <= N) && (k0 + kBK <= K);
    if (VEC && full) {
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        int r, c;
        a_rc(i, r, c);
        ra[i] = TA ? *reinterpret_cast<const V2*>(A + (long)(k0 + r) * lda + m0 + c)
                   : *reinterpret_cast<const V2*>(A + (long)(m0 + r) * lda + k0 + c);
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        int r, c;
        b_rc(i, r, c);
        rb[i] = TB ? *reinterpret_cast<const V2*>(B + (long)(n0 + r) * ldb + k0 + c)
                   : *reinterpret_cast<const V2*>(B + (long)(k0 + r) * ldb + n0 + c);
      }
    } else {
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        int r, c;
        a_rc(i, r, c);
        // stored row gr, stored columns gc, gc + 1 and their bounds
        const int gr = TA ? k0 + r : m0 + r, gc = TA ? m0 + c : k0 + c;
        const bool rin = gr < (TA ? K : M);
        const int cl = TA ? M : K;
        ra[i].x = (rin && gc < cl) ? A[(long)gr * lda + gc] : T(0);
        ra[i].y = (rin && gc + 1 < cl) ? A[(long)gr * lda + gc + 1] : T(0);
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        int r, c;
        b_rc(i, r, c);
        const int gr = TB ? n0 + r : k0 + r, gc = TB ? k0 + c : n0 + c;
        const bool rin = gr < (TB ? N : K);
        const int cl = TB ? K : N;
        rb[i].x = (rin && gc < cl) ? B[(long)gr * ldb + gc] : T(0);
        rb[i].y = (rin && gc + 1 < cl) ? B[(long)gr * ldb + gc + 1] : T(0);
      }
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      int r, c;
      a_rc(i, r, c);
      if (TA) {  // (k = r, m = c, c + 1) -> As[m][k]
        As[buf][c * LDA_S + r] = ra[i].x;
        As[buf][(c + 1) * LDA_S + r] = ra[i].y;
      } else {
        As[buf][r * LDA_S + c] = ra[i].x;
        As[buf][r * LDA_S + c + 1] = ra[i].y;
      }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      int r, c;
      b_rc(i, r, c);
      if (TB) {  // (n = r, k = c, c + 1) -> Bs[k][n]
        Bs[buf][c * LDB_S + r] = rb[i].x;
        Bs[buf][(c + 1) * LDB_S + r] = rb[i].y;
      } else {
        *reinterpret_cast<V2*>(&Bs[buf][r * LDB_S + c]) = rb[i];
      }
    }
  };

  const int nk = (K + kBK - 1) / kBK;
  load_slab(0);
  store_slab(0);
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load_slab((kt + 1) * kBK);
    const T* as = As[buf];
    const T* bs = Bs[buf];
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      T a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = as[(wm0 + i * 16 + lane % 16) * LDA_S + kk + lane / 16];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = bs[(kk + lane / 16) * LDB_S + wn0 + j * 16 + lane % 16];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(a[i], b[j], acc[i][j]);
    }
    if (kt + 1 < nk) store_slab(buf ^ 1);
    __syncthreads();
  }

  // Epilogue: C = alpha*acc + beta*C; beta == 0 never reads C (it may hold NaN/garbage).
  const bool read_c = beta != T(0);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = m0 + wm0 + i * 16 + MF::row(r, lane);
        const int gc = n0 + wn0 + j * 16 + lane % 16;
        if (gr < M && gc < N) {
          T* p = C + (long)gr * ldc + gc;
          T v = alpha * acc[i][j][r];
          if (read_c) v += beta * *p;
          *p = v;
        }
      }
}

template <typename T, int BM, int BN, int WM, int WN, bool VEC, bool TA = false, bool TB = false>
__global__ __launch_bounds__(kThreads, 2) void gemm_mfma_kernel(int M, int N, int K, T alpha,
                                                             const T* __restrict__ A, long lda,
                                                             const T* __restrict__ B, long ldb,
                                                             T beta, T* __restrict__ C,
                                                             long ldc) {
  __shared__ GemmShm<T, BM, BN> sh;
  const int tiles_m = (M + BM - 1) / BM;
  const int tiles_n = (N + BN - 1) / BN;
  int tm, tn;
  tile_coords(tiles_m, tiles_n, tm, tn);
  gemm_tile<T, BM, BN, WM, WN, VEC, TA, TB>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, sh);
}

// ---- grouped GEMM: many independent products of uneven sizes in one persistent launch ----
constexpr int kGT = 64;  // tile of the grouped kernel (64 x 64, waves of 32 x 32)

// Tiles of one problem; a problem that fails the bzg_dgemm argument rules gets none (and
// status -1 when a status array is given), an empty one none (status 0).
__device__ __forceinline__ int64_t grouped_tiles(const bzg_dgemm_problem& q, int32_t* status) {
  const bool ok = q.m >= 0 && q.n >= 0 && q.k >= 0 && q.m < (1LL << 31) && q.n < (1LL << 31) &&
                  q.k < (1LL << 31) && q.lda >= (q.k > 0 ? q.k : 1) && q.ldb >= (q.n > 0 ? q.n : 1) &&
                  q.ldc >= (q.n > 0 ? q.n : 1) && (q.m == 0 || q.n == 0 || q.C != nullptr) &&
                  (q.m == 0 || q.n == 0 || q.k == 0 || (q.A != nullptr && q.B != nullptr));
  if (status) *status = ok ? 0 : -1;
  if (!ok || q.m == 0 || q.n == 0) return 0;
  return ((q.m + kGT - 1) / kGT) * ((q.n + kGT - 1) / kGT);
}

// prefix[p] = number of 64 x 64 tiles of problems 0..p-1 (p = 0..count); one workgroup.
__global__ __launch_bounds__(1024) void grouped_prefix_kernel(int count, const bzg_dgemm_problem* __restrict__ pr,
                                                              int64_t* __restrict__ prefix,
                                                              int32_t* __restrict__ status) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int per = (count + 1023) / 1024;
  const int b = t * per, e = min(count, b + per);
  int64_t s = 0;
  for (int p = b; p < e; ++p) s += grouped_tiles(pr[p], nullptr);
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0;
    for (int i = 0; i < 1024; ++i) {
      const int64_t v = part[i];
      part[i] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int64_t acc = part[t];
  for (int p = b; p < e; ++p) {
    prefix[p] = acc;
    acc += grouped_tiles(pr[p], status ? status + p : nullptr);
  }
  if (t == 1023) prefix[count] = acc;  // the last range ends at count: acc is the total
}

// Persistent: workgroup w takes tiles w, w + grid, ... of the concatenated tile list (a fixed
// assignment, no atomics); the problem of a tile is found by binary search in the prefix.
// Problems with k == 0 are scaled (C = beta C) tile by tile.
__global__ __launch_bounds__(kThreads, 2) void gemm_grouped_kernel(int count, const bzg_dgemm_problem* __restrict__ pr,
                                                                   const int64_t* __restrict__ prefix) {
  __shared__ GemmShm<double, kGT, kGT> sh;
  const int64_t total = prefix[count];
  for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
    int lo = 0, hi = count - 1;  // last p with prefix[p] <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const bzg_dgemm_problem q = pr[lo];
    const int tiles_n = (int)((q.n + kGT - 1) / kGT);
    const int local = (int)(t - prefix[lo]);
    const int tm = local / tiles_n, tn = local % tiles_n;
    const bool vec = ((uintptr_t)q.A % 16 == 0) && ((uintptr_t)q.B % 16 == 0) && (q.lda % 2 == 0) &&
                     (q.ldb % 2 == 0);
    if (q.k == 0) {
      for (int e = threadIdx.x; e < kGT * kGT; e += kThreads) {
        const long r = (long)tm * kGT + e / kGT, c = (long)tn * kGT + e % kGT;
        if (r < q.m && c < q.n) {
          double* p = q.C + r * q.ldc + c;
          *p = (q.beta == 0.0) ? 0.0 : q.beta * *p;
        }
      }
    } else if (vec) {
      gemm_tile<double, kGT, kGT, 32, 32, true>((int)q.m, (int)q.n, (int)q.k, q.alpha, q.A, q.lda, q.B, q.ldb,
                                                q.beta, q.C, q.ldc, tm, tn, sh);
    } else {
      gemm_tile<double, kGT, kGT, 32, 32, false>((int)q.m, (int)q.n, (int)q.k, q.alpha, q.A, q.lda, q.B, q.ldb,
                                                 q.beta, q.C, q.ldc, tm, tn, sh);
    }
    __syncthreads();  // LDS of this tile is free before the next one is staged
  }
}

// C = beta*C for k == 0 (and the empty product).
template <typename T>
__global__ void scale_matrix_kernel(int M, int N, T beta, T* C, long ldc) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)M * N) return;
  const long r = idx / N, c = idx % N;
  T* p = C + r * ldc + c;
  *p = (beta == T(0)) ? T(0) : beta * *p;
}

// INT32 GEMM on VALU: 64x64 tile, 256 threads, 4x4 micro-tile per thread.  TA / TB: the
// operand is stored transposed (A as K x M, B as N x K); the staging loop then walks the tile so
// that neighbouring threads read neighbouring addresses.  Wrap-around integer arithmetic is
// exact, so every form gives the same bits as the product of transposed copies.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void igemm_kernel(int M, int N, int K, uint32_t alpha,
                                                    const int32_t* __restrict__ A, long lda,
                                                    const int32_t* __restrict__ B, long ldb,
                                                    uint32_t beta, int32_t* __restrict__ C,
                                                    long ldc) {
  constexpr int T = 64, BK = 16;
  __shared__ uint32_t As[BK][T + 4];  // transposed: As[k][m]
  __shared__ uint32_t Bs[BK][T + 4];
  const int tid = threadIdx.x;
  const int tiles_n = (N + T - 1) / T;
  const int m0 = (blockIdx.x / tiles_n) * T, n0 = (blockIdx.x % tiles_n) * T;
  const int tr = (tid / 16) * 4, tc = (tid % 16) * 4;
  uint32_t acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += BK) {
    for (int e = tid; e < T * BK; e += 256) {
      // A element (m0+r, k0+c): stored at [m][k], or at [k][m] when TA
      const int r = TA ? e % T : e / BK, c = TA ? e / T : e % BK;
      const bool ain = m0 + r < M && k0 + c < K;
      As[c][r] = ain ? (uint32_t)(TA ? A[(long)(k0 + c) * lda + m0 + r] : A[(long)(m0 + r) * lda + k0 + c]) : 0u;
      // B element (k0+kr, n0+nc): stored at [k][n], or at [n][k] when TB
      const int kr = TB ? e % BK : e / T, nc = TB ? e / BK : e % T;
      const bool bin = k0 + kr < K && n0 + nc < N;
      Bs[kr][nc] = bin ? (uint32_t)(TB ? B[(long)(n0 + nc) * ldb + k0 + kr] : B[(long)(k0 + kr) * ldb + n0 + nc]) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BK; ++k) {
      uint32_t a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[k][tr + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[k][tc + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gr = m0 + tr + i, gc = n0 + tc + j;
      if (gr < M && gc < N) {
        int32_t* p = C + (long)gr * ldc + gc;
        uint32_t v = alpha * acc[i][j];
        if (beta != 0u) v += beta * (uint32_t)*p;
        *p = (int32_t)v;
      }
    }
}

template <typename T>
bzg_status check_gemm_args(bzg_context* ctx, int64_t m, int64_t n, int64_t k, const T* A,
                           int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc, bool ta = false,
                           bool tb = false) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemm: null context");
  BZG_REQUIRE(m >= 0 && n >= 0 && k >= 0, BZG_ERR_INVALID_ARGUMENT, "gemm: negative size");
  BZG_REQUIRE(m < (1LL << 31) && n < (1LL << 31) && k < (1LL << 31), BZG_ERR_UNSUPPORTED,
              "gemm: dimension >= 2^31");
  const int64_t arow = ta ? m : k, brow = tb ? k : n;  // stored row lengths of A and B
  BZG_REQUIRE(lda >= (arow > 0 ? arow : 1) && ldb >= (brow > 0 ? brow : 1) && ldc >= (n > 0 ? n : 1),
              BZG_ERR_INVALID_ARGUMENT, "gemm: leading dimension smaller than row length");
  if (m > 0 && n > 0) BZG_REQUIRE(C != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemm: null C");
  if (m > 0 && n > 0 && k > 0)
    BZG_REQUIRE(A != nullptr && B != nullptr, BZG_ERR_INVALID_ARGUMENT, "gemm: null A or B");
  return BZG_OK;
}

template <typename T, int BM, int BN, int WM, int WN, bool TA = false, bool TB = false>
bzg_status launch_tile(bzg_context* ctx, int m, int n, int k, T alpha, const T* A, long lda,
                       const T* B, long ldb, T beta, T* C, long ldc) {
  const bool vec = ((uintptr_t)A % (2 * sizeof(T)) == 0) && ((uintptr_t)B % (2 * sizeof(T)) == 0) &&
                   (lda % 2 == 0) && (ldb % 2 == 0);
  ProfScope prof(ctx, sizeof(T) == 8 ? "dgemm" : "sgemm");
  const long tiles = (long)((m + BM - 1) / BM) * ((n + BN - 1) / BN);
  BZG_REQUIRE(tiles < (1L << 31), BZG_ERR_UNSUPPORTED, "gemm: too many tiles");
  dim3 grid((unsigned)tiles), block(kThreads);
  if (vec)
    hipLaunchKernelGGL((gemm_mfma_kernel<T, BM, BN, WM, WN, true, TA, TB>), grid, block, 0, ctx->stream,
                       m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  else
    hipLaunchKernelGGL((gemm_mfma_kernel<T, BM, BN, WM, WN, false, TA, TB>), grid, block, 0, ctx->stream,
                       m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

// ---- tile-shape autotuner ------------------------------------------------------------
// Tile shapes: 0 = 32 x 32, 1 = 64 x 64, 2 = 128 x 128.  bzg_gemm_autotune times the three
// shapes on a product of the given size and records the fastest for that size class
// (ceil(log2) of m, n, k and the element type); later products of the class use it.  All
// shapes give bit-identical results, so tuning changes speed only.
struct TuneKey {
  int type, lm, ln, lk;
  bool operator<(const TuneKey& o) const {
    return std::tie(type, lm, ln, lk) < std::tie(o.type, o.lm, o.ln, o.lk);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;

int ceil_log2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}
TuneKey tune_key(int type, int64_t m, int64_t n, int64_t k) {
  return TuneKey{type, ceil_log2(m), ceil_log2(n), ceil_log2(k)};
}

// Shapes: 0 = 32x32, 1 = 64x64, 2 = 128x128 tiles; 3 = 128x256 (64x128 per wavefront, FP32
// only: its accumulators need 128 registers, 256 in FP64) -- 123.8 against 116.1 TFLOP/s at
// 8192^3, 112.0 against 102.7 at 4096^3.
template <typename T>
constexpr int kShapes = sizeof(T) == 4 ? 4 : 3;

template <typename T>
int heuristic_shape(int64_t m, int64_t n) {
  if (sizeof(T) == 4 && ((m + 127) / 128) * ((n + 255) / 256) >= 240) return 3;
  const long t128 = ((m + 127) / 128) * ((n + 127) / 128);
  const long t64 = ((m + 63) / 64) * ((n + 63) / 64);
  if (t128 >= 240) return 2;
  if (t64 >= 200) return 1;
  return 0;
}

template <typename T, bool TA = false, bool TB = false>
bzg_status launch_shape_t(bzg_context* ctx, int shape, int64_t m, int64_t n, int64_t k, T alpha, const T* A,
                          int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc) {
  if constexpr (sizeof(T) == 4) {
    if (shape == 3)
      return launch_tile<T, 128, 256, 64, 128, TA, TB>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  if (shape == 2)
    return launch_tile<T, 128, 128, 64, 64, TA, TB>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (shape == 1)
    return launch_tile<T, 64, 64, 32, 32, TA, TB>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  return launch_tile<T, 32, 32, 16, 16, TA, TB>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

template <typename T>
bzg_status launch_shape(bzg_context* ctx, int shape, int64_t m, int64_t n, int64_t k, T alpha, const T* A,
                        int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc, bool ta = false,
                        bool tb = false) {
  if (ta && tb) return launch_shape_t<T, true, true>(ctx, shape, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (ta) return launch_shape_t<T, true, false>(ctx, shape, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (tb) return launch_shape_t<T, false, true>(ctx, shape, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  return launch_shape_t<T, false, false>(ctx, shape, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Tile choice by size only (never by load), so a product's result never depends on what
// else is running: the tuned shape of the size class if there is one, else a heuristic
// (large tiles cut L2->LDS traffic; small tiles keep >= ~256 workgroups in flight).
template <typename T>
bzg_status gemm_dispatch(bzg_context* ctx, int64_t m, int64_t n, int64_t k, T alpha, const T* A,
                         int64_t lda, const T* B, int64_t ldb, T beta, T* C, int64_t ldc, bool ta = false,
                         bool tb = false) {
  bzg_status st = check_gemm_args(ctx, m, n, k, A, lda, B, ldb, C, ldc, ta, tb);
  if (st != BZG_OK) return st;
  if (m == 0 || n == 0) return BZG_OK;
  if (k == 0) {
    const long total = m * n;
    hipLaunchKernelGGL(scale_matrix_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       ctx->stream, (int)m, (int)n, beta, C, (long)ldc);
    BZG_LAUNCH_CHECK();
    return BZG_OK;
  }
  int shape = heuristic_shape<T>(m, n);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(tune_key(sizeof(T) == 8 ? 0 : 1, m, n, k));
    if (it != g_tune.end()) shape = it->second;
  }
  return launch_shape<T>(ctx, shape, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ta, tb);
}

template <typename T>
bzg_status gemm_autotune(bzg_context* ctx, int64_t m, int64_t n, int64_t k, int32_t* chosen, float* ms_out) {
  BZG_REQUIRE(ctx && m > 0 && n > 0 && k > 0, BZG_ERR_INVALID_ARGUMENT, "gemm_autotune: bad argument");
  BZG_REQUIRE(m < (1LL << 31) && n < (1LL << 31) && k < (1LL << 31), BZG_ERR_UNSUPPORTED,
              "gemm_autotune: dimension >= 2^31");
  const size_t bytes = sizeof(T) * ((size_t)m * k + (size_t)k * n + (size_t)m * n);
  void* ws = nullptr;
  bzg_status st = workspace(ctx, 7, bytes, &ws);
  if (st) return st;
  T* A = static_cast<T*>(ws);
  T* B = A + (size_t)m * k;
  T* C = B + (size_t)k * n;
  // operand values do not change the timing of these kernels: zero-filled scratch
  BZG_HIP_CHECK(hipMemsetAsync(ws, 0, bytes, ctx->stream));
  hipEvent_t e0, e1;
  BZG_HIP_CHECK(hipEventCreate(&e0));
  BZG_HIP_CHECK(hipEventCreate(&e1));
  int best = -1;
  float best_ms = 0.f;
  const bool prof = ctx->profile;
  ctx->profile = false;
  for (int shape = 0; shape < kShapes<T> && !st; ++shape) {
    st = launch_shape<T>(ctx, shape, m, n, k, T(1), A, k, B, n, T(0), C, n);  // warm-up
    if (st) break;
    const int reps = 3;
    hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps && !st; ++r) st = launch_shape<T>(ctx, shape, m, n, k, T(1), A, k, B, n, T(0), C, n);
    hipEventRecord(e1, ctx->stream);
    if (hipEventSynchronize(e1) != hipSuccess) st = BZG_ERR_HIP;
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    if (!st && (best < 0 || ms < best_ms)) {
      best = shape;
      best_ms = ms;
    }
  }
  ctx->profile = prof;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st) return st;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[tune_key(sizeof(T) == 8 ? 0 : 1, m, n, k)] = best;
  }
  if (chosen) *chosen = best;
  if (ms_out) *ms_out = best_ms;
  return BZG_OK;
}

}  // namespace
}  // namespace bzg

extern "C" {

bzg_status bzg_dgemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, double alpha,
                     const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                     double* C, int64_t ldc) {
  return bzg::gemm_dispatch<double>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

bzg_status bzg_dgemm_grouped(bzg_context* ctx, int64_t count, const bzg_dgemm_problem* problems,
                             int32_t* status) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgemm_grouped: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "dgemm_grouped: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(problems != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgemm_grouped: null problem array");
  void* ws = nullptr;
  bzg_status st = bzg::workspace(ctx, 6, (size_t)(count + 1) * sizeof(int64_t), &ws);
  if (st) return st;
  int64_t* prefix = static_cast<int64_t*>(ws);
  bzg::ProfScope prof(ctx, "dgemm_grouped");
  hipLaunchKernelGGL(bzg::grouped_prefix_kernel, dim3(1), dim3(1024), 0, ctx->stream, (int)count, problems, prefix,
                     status);
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(bzg::gemm_grouped_kernel, dim3((unsigned)(ctx->num_cus * 4)), dim3(bzg::kThreads), 0,
                     ctx->stream, (int)count, problems, prefix);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

bzg_status bzg_gemm_autotune(bzg_context* ctx, int32_t fp32, int64_t m, int64_t n, int64_t k,
                             int32_t* shape, float* ms) {
  return fp32 ? bzg::gemm_autotune<float>(ctx, m, n, k, shape, ms)
              : bzg::gemm_autotune<double>(ctx, m, n, k, shape, ms);
}

bzg_status bzg_gemm_tuning_clear(void) {
  std::lock_guard<std::mutex> lk(bzg::g_tune_mu);
  bzg::g_tune.clear();
  return BZG_OK;
}

bzg_status bzg_sgemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, float alpha,
                     const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                     float* C, int64_t ldc) {
  return bzg::gemm_dispatch<float>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

bzg_status bzg_dgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        double alpha, const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc) {
  return bzg::gemm_dispatch<double>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, transa != 0, transb != 0);
}

bzg_status bzg_sgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                        float* C, int64_t ldc) {
  return bzg::gemm_dispatch<float>(ctx, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, transa != 0, transb != 0);
}

bzg_status bzg_igemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        int32_t alpha, const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb,
                        int32_t beta, int32_t* C, int64_t ldc) {
  const bool ta = transa != 0, tb = transb != 0;
  bzg_status st = bzg::check_gemm_args(ctx, m, n, k, A, lda, B, ldb, C, ldc, ta, tb);
  if (st != BZG_OK) return st;
  if (m == 0 || n == 0) return BZG_OK;
  {
    // the INT8 matrix cores where they measured faster than the vector ALUs (MI355X: from
    // 256 x 256 x 4096 up, 2.7x at 1024^3, 4.5x at 8192^3; even at 1024 x 1024 x 256): k >= 512
    // and at least one 256 x 256 tile of output, or k >= 256 and 2048 x 2048; m and n >= 128
    // (the digit GEMM pads both to 256: skinny products stay on the VALU); a digit workspace
    // of at most 8 GiB (one K chunk of the digits and the 4 class planes)
    const int64_t mp = (m + 255) / 256 * 256, np = (n + 255) / 256 * 256;
    const int64_t kp = std::min<int64_t>((k + 127) / 128 * 128, bzg::kIgemmDigitsChunk);
    const double ws_bytes = 16.0 * mp * np + 4.0 * (mp + np) * kp;
    const double mn = (double)m * (double)n;
    if (((k >= 512 && mn >= 65536.0) || (k >= 256 && mn >= 4194304.0)) && std::min(m, n) >= 128 &&
        ws_bytes <= 8.0 * (1L << 30))
      return bzg::igemm_digits_impl(ctx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  const long tiles = ((m + 63) / 64) * ((n + 63) / 64);
  BZG_REQUIRE(tiles < (1L << 31), BZG_ERR_UNSUPPORTED, "igemm: too many tiles");
  const dim3 grid((unsigned)tiles), block(256);
#define BZG_IGEMM(TA_, TB_)                                                                                  \
  hipLaunchKernelGGL((bzg::igemm_kernel<TA_, TB_>), grid, block, 0, ctx->stream, (int)m, (int)n, (int)k,   \
                     (uint32_t)alpha, A, (long)lda, B, (long)ldb, (uint32_t)beta, C, (long)ldc)
  if (ta && tb) BZG_IGEMM(true, true);
  else if (ta) BZG_IGEMM(true, false);
  else if (tb) BZG_IGEMM(false, true);
  else BZG_IGEMM(false, false);
#undef BZG_IGEMM
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

bzg_status bzg_igemm_digits(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                            int32_t alpha, const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb,
                            int32_t beta, int32_t* C, int64_t ldc) {
  const bool ta = transa != 0, tb = transb != 0;
  bzg_status st = bzg::check_gemm_args(ctx, m, n, k, A, lda, B, ldb, C, ldc, ta, tb);
  if (st != BZG_OK) return st;
  if (m == 0 || n == 0) return BZG_OK;
  if (k == 0) return bzg_igemm_ex(ctx, transa, transb, m, n, 0, alpha, A, lda, B, ldb, beta, C, ldc);
  return bzg::igemm_digits_impl(ctx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

bzg_status bzg_igemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, int32_t alpha,
                     const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb,
                     int32_t beta, int32_t* C, int64_t ldc) {
  return bzg_igemm_ex(ctx, 0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

}  // extern "C"
