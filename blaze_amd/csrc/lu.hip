// Batched dense solves for gfx950: LU with partial pivoting (LAPACK dgetrf/dgetrs semantics)
// and the mixed-precision variant (FP32 LU + FP64 iterative refinement, LAPACK dsgesv).
//
// Replaces solve(A,x,b) / solve(A,X,B) (blaze/math/dense/LSE.h:6343-6373, 6452-6482) whose
// N >= 7 branch copies A and calls gesv (LSE.h:5229-5257 -> lapack/gesv.h:136-175 ->
// dgesv_ at lapack/clapack/gesv.h:207).  Here one call solves many independent systems.
//
// Execution model: a persistent grid (a few workgroups per CU) pulls system indices from an
// atomic work counter in longest-first order, so uneven sizes (n = 1..512) balance without a
// host round trip.  One workgroup of one wavefront (64 threads) owns one system at a time:
//   factor: blocked right-looking LU, panel width 8 (FP64) or 16 (FP32).  The panel lives in
//           the wavefront's registers (NM / 64 rows per lane); each column step is a per-lane
//           argmax, a butterfly over the wavefront, the pivot row's broadcast and a rank-1
//           update.  The panel's interchanges are then applied to the other columns in one
//           gather (each lane follows the swaps backwards to find the source row of every
//           touched position), the U12 block row is solved against the unit-lower L11, and the
//           trailing matrix gets A22 -= L21 * U12 with MFMA 16x16x4 tiles (operands from LDS,
//           A22 in place in HBM).
//   solve:  P*B is loaded into LDS; forward/backward substitution by 16-row blocks: the
//           off-diagonal part is a dot product over the factored rows, the 16x16 diagonal block
//           is solved by one lane per right-hand side.
// The mixed-precision entry point (bzg_dsgesv_batched) sends orders 1..256 to the on-chip
// kernels of lu_mixed.hip and keeps the streaming kernel below for orders 257..512; systems
// that do not converge in FP32 are re-solved by the FP64 kernel (gesv_fallback_kernel).
// Pivot rule = dgetf2/idamax: the first row of maximal |a| in the column; an exactly-zero pivot
// sets info (1-based) and the column is left unscaled, as LAPACK does.
#include "common.hpp"

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

namespace bzg {
namespace {

constexpr int NT = 64;        // threads per workgroup: one wavefront per system
constexpr int NW = NT / 64;   // wavefronts per workgroup
constexpr int NB = 16;        // largest panel width
// Panel width per precision: the panel lives in the registers of one wavefront (4 rows x
// kNB values per lane), 8 doubles or 16 floats keep it at 64 registers.
template <typename T>
constexpr int kNB = sizeof(T) == 8 ? 8 : 16;
constexpr int NMAX = 512;     // largest system order
constexpr int NSMALL = 256;   // systems up to this order use the NM = 256 kernels
constexpr int RMAX = 8;       // largest number of right-hand sides
constexpr int UPAD = 16;      // row padding of the U12 buffer (conflict-free MFMA B reads)

// NM: the largest order a kernel instance handles (256 or 512); it sizes the LDS buffers and
// the register panel (NM / 64 rows per lane).
template <typename T, int NM>
struct FactorShm {
  T panel[NM][kNB<T> + 1];      // L11 / L21 of the current panel (rows beyond n are zero)
  T ublk[kNB<T>][NM + UPAD];    // U12 of the current panel
};

template <typename T, int NM>
struct SolveShm {
  T x[NM * RMAX];             // right-hand sides / solution, row-major, ld = RMAX
  T diag[16 * 17];            // diagonal block being solved
};

template <int NM>
struct Misc {
  double pv[2][4];            // per-wave best |a| of the column (double-buffered by step parity)
  int pr[2][4];               // ... and its row
  int piv[NB];                // absolute pivot rows of the current panel (0-based)
  int swsrc[2 * NB];          // interchange sources: rows j0.. (first NB), pivot rows (last NB)
  int swdst[2 * NB];          // ... and their destinations (-1: none)
  int perm[NM];               // row i of the factored system = original row perm[i]
  int pinv[NM];               // original row r is row pinv[r] of the factored system
  int sys;
  unsigned long long cmax[RMAX * 2];  // column maxima (bit patterns of non-negative doubles)
  int flag;
};

template <typename T>
struct MfmaLU;
template <>
struct MfmaLU<double> {
  using acc_t = d4;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int i, int lane) { return 4 * i + lane / 16; }
};
template <>
struct MfmaLU<float> {
  using acc_t = f4;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int i, int lane) { return 4 * (lane / 16) + i; }
};

template <typename T>
__device__ __forceinline__ T tabs(T v) { return v < T(0) ? -v : v; }

// Value of lane `src` (wave-uniform) broadcast to the wavefront.
__device__ __forceinline__ double bcast(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// LU factorisation of the n x n row-major matrix A (leading dimension lda) in place.
// Writes 1-based pivots to ipiv (may be null) and the row permutation to misc.perm.
// Returns info (0 or 1-based index of the first exactly-zero pivot).
template <typename T, int NM>
__device__ __forceinline__ int lu_factor(T* __restrict__ A, int n, int lda, int32_t* __restrict__ ipiv,
                         FactorShm<T, NM>& sh, Misc<NM>& misc) {
  using MF = MfmaLU<T>;
  using acc_t = typename MF::acc_t;
  constexpr int NB = kNB<T>;
  constexpr int Q = NM / 64;  // panel rows per lane
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int info = 0;
  for (int i = tid; i < n; i += NT) misc.perm[i] = i;
  __syncthreads();

  for (int j0 = 0; j0 < n; j0 += NB) {
    const int m = n - j0;
    const int nbe = min(NB, m);

    // ---- panel factorisation (columns j0 .. j0+nbe-1) by wavefront 0 alone ----
    // Lane l holds panel rows l, l+64, l+128, ... (Q of them) in registers.  Rows never move between
    // registers: each keeps its logical position lpos (its row index after the interchanges
    // so far), an interchange just swaps two lpos values, and the rows are written to their
    // final positions at the end.  A column step is a per-lane max over its 4 rows, a
    // butterfly over the wavefront and v_readlane broadcasts of the pivot row: no LDS traffic
    // and no barrier per column.
    if (wave == 0) {
      T P[Q][NB];
      int lpos[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int rr = lane + 64 * q;
        lpos[q] = (rr < m) ? rr : (1 << 20);  // rows beyond the matrix never become candidates
#pragma unroll
        for (int c = 0; c < NB; ++c)
          P[q][c] = (rr < m && c < nbe) ? A[(long)(j0 + rr) * lda + j0 + c] : T(0);
      }
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        if (jj < nbe) {
          // Candidates: logical rows jj..m-1.  A NaN entry ranks below every number (as in
          // idamax, where NaN never compares greater) but above the non-candidates, so the
          // pivot is always a row of the matrix.  Ties: the lower logical row.
          int p;
          if constexpr (sizeof(T) == 4) {
            // one 64-bit key per candidate: (|a| bits + 2, or 1 for NaN, 0 for a non-candidate)
            // above (2^32 - 1 - logical row); the wavefront maximum is the pivot
            unsigned long long key = 0ull;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
              const bool cand = lpos[q] >= jj && lpos[q] < m;
              const float a = (float)P[q][jj];
              const unsigned hi = !cand ? 0u : (a != a ? 1u : (__float_as_uint(a) & 0x7FFFFFFFu) + 2u);
              const unsigned long long k = ((unsigned long long)hi << 32) | (0xFFFFFFFFu - (unsigned)lpos[q]);
              key = k > key ? k : key;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
              const unsigned long long o = __shfl_xor(key, off, 64);
              key = o > key ? o : key;
            }
            p = (int)(0xFFFFFFFFu - (unsigned)__builtin_amdgcn_readfirstlane((int)(key & 0xFFFFFFFFull)));
          } else {
            double v = -1.0;
            int r = 1 << 30;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
              double a = (lpos[q] >= jj && lpos[q] < m) ? (double)tabs(P[q][jj]) : -1.0;
              if (a != a) a = -0.5;
              if (a > v || (a == v && lpos[q] < r)) { v = a; r = lpos[q]; }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
              const double ov = __shfl_xor(v, off, 64);
              const int orr = __shfl_xor(r, off, 64);
              if (ov > v || (ov == v && orr < r)) { v = ov; r = orr; }
            }
            p = __builtin_amdgcn_readfirstlane(r);
          }
          // physical place of logical row p: the lane holding it (ballot) and its register row
          bool mine = false;
          int myq = 0;
#pragma unroll
          for (int q = 0; q < Q; ++q)
            if (lpos[q] == p) { mine = true; myq = q; }
          const unsigned long long hold = __ballot(mine);
          const int pl = __builtin_amdgcn_readfirstlane(__ffsll((long long)hold) - 1);
          const int pq = __builtin_amdgcn_readlane(myq, pl);
          // pivot row values (physical row (pl, pq)), kept in scalar registers
          T pr[NB];
#pragma unroll
          for (int c = 0; c < NB; ++c) {
            T cand = P[0][c];
#pragma unroll
            for (int q = 1; q < Q; ++q) cand = (pq == q) ? P[q][c] : cand;
            pr[c] = bcast(cand, pl);
          }
          const T piv = pr[jj];
          if (piv != T(0)) {
            if (lane == 0) {
              misc.piv[jj] = j0 + p;
              const int t = misc.perm[j0 + jj];
              misc.perm[j0 + jj] = misc.perm[j0 + p];
              misc.perm[j0 + p] = t;
            }
            // interchange logical rows jj and p
#pragma unroll
            for (int q = 0; q < Q; ++q)
              lpos[q] = (lpos[q] == jj) ? p : ((lpos[q] == p) ? jj : lpos[q]);
            const T rinv = T(1) / piv;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
              if (lpos[q] > jj && lpos[q] < m) {
                const T l = P[q][jj] * rinv;
                P[q][jj] = l;
                // constant trip count (c > jj folds after the jj loop is unrolled): a loop
                // starting at jj + 1 would be unrolled first with a run-time trip count and
                // index P dynamically
#pragma unroll
                for (int c = 0; c < NB; ++c)
                  if (c > jj) P[q][c] -= l * pr[c];
              }
            }
          } else {
            // Exactly-zero pivot: no interchange, no scaling (dgetf2); the column below is
            // zero, so the rank-1 update is a no-op.
            if (lane == 0) misc.piv[jj] = j0 + jj;
            if (info == 0) info = j0 + jj + 1;
          }
        }
      }
      // ---- panel to HBM and LDS, each row at its final (logical) position ----
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int rr = lane + 64 * q;
        const int lr = lpos[q];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          if (rr < m) {
            if (c < nbe) A[(long)(j0 + lr) * lda + j0 + c] = P[q][c];
            sh.panel[lr][c] = P[q][c];
          } else {
            sh.panel[rr][c] = T(0);
          }
        }
      }
      if (lane == 0) misc.flag = info;
    }
    __syncthreads();
    info = misc.flag;  // wavefront 0 tracked the first zero pivot
    if (ipiv && tid < nbe) ipiv[j0 + tid] = misc.piv[tid] + 1;

    // ---- interchanges on the other columns, then U12 = L11^-1 A12 ----
    // Source row of every position the panel's interchanges touch (rows j0..j0+nbe-1 and the
    // pivot rows below them), found by following the swaps backwards; computed once by 32
    // threads, then every column gathers with one batch of loads.
    if (tid < 2 * NB) {
      const int q = tid % NB;
      int pos = -1;
      if (q < nbe) pos = (tid < NB) ? j0 + q : misc.piv[q];
      if (tid >= NB && pos < j0 + nbe) pos = -1;  // low list: only pivot rows below the block
      const int dst = pos;
      if (pos >= 0) {
        for (int jj = nbe - 1; jj >= 0; --jj) {
          const int pj = misc.piv[jj];
          if (pos == j0 + jj) pos = pj;
          else if (pos == pj) pos = j0 + jj;
        }
      }
      misc.swsrc[tid] = pos;
      misc.swdst[tid] = dst;
    }
    __syncthreads();
    const int other = n - nbe;
    for (int oc = tid; oc < other; oc += NT) {
      const int c = oc < j0 ? oc : oc + nbe;
      T top[NB], low[NB];
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        const int st = misc.swsrc[jj], sl = misc.swsrc[NB + jj];
        top[jj] = (st >= 0) ? A[(long)st * lda + c] : T(0);
        low[jj] = (sl >= 0) ? A[(long)sl * lda + c] : T(0);
      }
      if (c > j0) {  // right of the panel: U12 column
#pragma unroll
        for (int jj = 1; jj < NB; ++jj)
#pragma unroll
          for (int kk = 0; kk < jj; ++kk) top[jj] -= sh.panel[jj][kk] * top[kk];
        const int uc = c - j0 - NB;
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) sh.ublk[jj][uc] = top[jj];
      }
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        if (jj < nbe) A[(long)(j0 + jj) * lda + c] = top[jj];
        const int dl = misc.swdst[NB + jj];
        if (dl >= 0) A[(long)dl * lda + c] = low[jj];
      }
    }
    // zero the unused tail of the U12 buffer (columns beyond n) for the MFMA tiles
    const int rem = n - j0 - NB;
    if (rem > 0) {
      const int padded = (rem + 15) & ~15;
      for (int e = tid; e < NB * (padded - rem); e += NT) {
        const int rr = e / (padded - rem), cc = rem + e % (padded - rem);
        sh.ublk[rr][cc] = T(0);
      }
    }
    __syncthreads();

    // ---- trailing update A22 -= L21 * U12 (MFMA 16x16x4) ----
    // Each wavefront takes its tiles in batches of TB: all loads of a batch are issued before
    // the MFMAs, so a batch costs one HBM/L2 round trip instead of one per tile.
    if (rem > 0) {
      constexpr int TB = (sizeof(T) == 8 ? 16 : 32) / (NM > 256 ? 2 : 1);  // register budget
      const int tiles = (rem + 15) / 16;
      const int r0 = j0 + NB;
      const int total = tiles * tiles;
      for (int t0 = wave; t0 < total; t0 += NW * TB) {
        acc_t acc[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) {
          const int t = t0 + NW * q;
          const int ti = t / tiles, tj = t % tiles;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int gr = r0 + ti * 16 + MF::row(i, lane), gc = r0 + tj * 16 + lane % 16;
            acc[q][i] = (t < total && gr < n && gc < n) ? A[(long)gr * lda + gc] : T(0);
          }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
          const int t = t0 + NW * q;
          if (t < total) {
            const int ti = t / tiles, tj = t % tiles;
#pragma unroll
            for (int kk = 0; kk < NB; kk += 4) {
              const T a = -sh.panel[NB + ti * 16 + lane % 16][kk + lane / 16];
              const T b = sh.ublk[kk + lane / 16][tj * 16 + lane % 16];
              acc[q] = MF::mma(a, b, acc[q]);
            }
          }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
          const int t = t0 + NW * q;
          const int ti = t / tiles, tj = t % tiles;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int gr = r0 + ti * 16 + MF::row(i, lane), gc = r0 + tj * 16 + lane % 16;
            if (t < total && gr < n && gc < n) A[(long)gr * lda + gc] = acc[q][i];
          }
        }
      }
    }
    __syncthreads();
  }
  return info;
}

// Dot products of up to 16 consecutive rows of M (row0 .. row0+nrows-1) with the columns of
// the LDS matrix X (ld RMAX) over k in [kb, ke):  out(r, c) = sum_k M(row0+r, k) * X(k, c).
// Thread t handles rows t/16 + 4j (j < 4) and k = kb + t%16 (mod 16): the loads of all 16
// rows over U k-steps (16 x 16U values) are in flight together, so a row block costs
// ceil((ke - kb) / 16U) memory round trips.  The 16 partials of a row are combined by a fixed
// butterfly; every lane of a 16-lane group ends with the row's sums.
template <int U, typename TM, typename TX, typename TA>
__device__ __forceinline__ void rows_dot(const TM* __restrict__ M, int lda, int row0, int nrows,
                                         int kb, int ke, const TX* __restrict__ X, int nrhs,
                                         TA (&acc)[4][RMAX]) {
  const int tid = threadIdx.x;
  const int r0 = tid / 16, ks = tid % 16;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < RMAX; ++c) acc[j][c] = TA(0);
  for (int k0 = kb + ks; k0 < ke; k0 += 16 * U) {
    TA a[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = r0 + 4 * j, k = k0 + 16 * u;
        a[u][j] = (r < nrows && k < ke) ? (TA)M[(long)(row0 + r) * lda + k] : TA(0);
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + 16 * u;
      if (k < ke) {
#pragma unroll
        for (int c = 0; c < RMAX; ++c) {
          if (c < nrhs) {
            const TA xv = (TA)X[k * RMAX + c];
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[j][c] += a[u][j] * xv;
          }
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < RMAX; ++c) {
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) acc[j][c] += __shfl_xor(acc[j][c], off, 64);
    }
}

// Solves (L U) X = X in place for the factored n x n matrix LU (row-major, unit-lower L),
// X in LDS (n x nrhs, ld RMAX), already permuted (X = P * B).  By 16-row blocks: the loads of
// the off-diagonal dot products and of the 16 x 16 diagonal block are issued together (one
// round trip per block), the diagonal block goes to LDS (`diag`, 16 x 17), and one thread per
// right-hand side then solves the block from registers.
template <typename T>
__device__ __forceinline__ void lu_solve(const T* __restrict__ LU, int n, int lda, T* X, int nrhs, T* diag) {
  const int tid = threadIdx.x;
  const int r0 = tid / 16, ks = tid % 16;
  auto stage_block = [&](int b0, int nr, int kb, int ke) {
    T dv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + 4 * j;
      dv[j] = (r < nr && ks < nr) ? LU[(long)(b0 + r) * lda + b0 + ks] : T(0);
    }
    T acc[4][RMAX];
    rows_dot<8, T, T, T>(LU, lda, b0, nr, kb, ke, X, nrhs, acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + 4 * j;
      diag[r * 17 + ks] = dv[j];
      if (r < nr && ks == 0 && ke > kb) {
#pragma unroll
        for (int c = 0; c < RMAX; ++c)
          if (c < nrhs) X[(b0 + r) * RMAX + c] -= acc[j][c];
      }
    }
    __syncthreads();
  };
  // forward: L y = b
  for (int b0 = 0; b0 < n; b0 += 16) {
    const int nr = min(16, n - b0);
    stage_block(b0, nr, 0, b0);
    if (tid < nrhs) {
      T xb[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) xb[i] = (i < nr) ? X[(b0 + i) * RMAX + tid] : T(0);
#pragma unroll
      for (int i = 1; i < 16; ++i) {
        T s = xb[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= diag[i * 17 + k] * xb[k];
        xb[i] = s;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < nr) X[(b0 + i) * RMAX + tid] = xb[i];
    }
    __syncthreads();
  }
  // backward: U x = y
  const int last = ((n - 1) / 16) * 16;
  for (int b0 = last; b0 >= 0; b0 -= 16) {
    const int nr = min(16, n - b0);
    stage_block(b0, nr, b0 + 16, n);
    if (tid < nrhs) {
      T xb[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) xb[i] = (i < nr) ? X[(b0 + i) * RMAX + tid] : T(0);
#pragma unroll
      for (int i = 15; i >= 0; --i) {
        if (i < nr) {
          T s = xb[i];
#pragma unroll
          for (int k = i + 1; k < 16; ++k)
            if (k < nr) s -= diag[i * 17 + k] * xb[k];
          xb[i] = s / diag[i * 17 + i];
        }
      }
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < nr) X[(b0 + i) * RMAX + tid] = xb[i];
    }
    __syncthreads();
  }
}

template <int NM>
__device__ __forceinline__ int next_system(int* counter, Misc<NM>& misc) {
  if (threadIdx.x == 0) misc.sys = atomicAdd(counter, 1);
  __syncthreads();
  const int s = misc.sys;
  __syncthreads();
  return s;
}

// ---- plain FP64 (or FP32) dgesv / dgetrf / dgetrs: A and B in place ---------------------
// MODE 0 (gesv): factor A, solve B.  MODE 1 (getrf): factor A only (rv may be null).
// MODE 2 (getrs): A holds factors and ipiv its 1-based pivots (as written by MODE 1); the row
// interchanges are applied to B in pivot order (LAPACK dlaswp) and B is solved in place.
// NM = NSMALL: takes the systems of order <= NSMALL from the (ordered) index range and skips
// the others; NM = NMAX: takes the systems listed in big_list (orders NSMALL+1 .. NMAX).
// Systems outside 1..NMAX / 0..RMAX were marked info = -1 by classify_kernel.
template <typename T, int NM>
__global__ __launch_bounds__(NT) void gesv_batched_kernel(int count, const int32_t* __restrict__ nv,
                                                          const int32_t* __restrict__ rv,
                                                          const int64_t* __restrict__ a_off,
                                                          const int64_t* __restrict__ b_off,
                                                          const int64_t* __restrict__ p_off,
                                                          T* __restrict__ A, T* __restrict__ B,
                                                          int32_t* __restrict__ ipiv,
                                                          int32_t* __restrict__ info,
                                                          const int32_t* __restrict__ order,
                                                          const int32_t* __restrict__ big_list,
                                                          const int* __restrict__ big_count,
                                                          int* counter, int mode) {
  __shared__ union {
    FactorShm<T, NM> f;
    SolveShm<T, NM> s;
  } sh;
  __shared__ Misc<NM> misc;
  const int tid = threadIdx.x;
  if (NM != NSMALL) count = *big_count;
  for (;;) {
    const int idx = next_system(counter, misc);
    if (idx >= count) break;
    const int s = (NM != NSMALL) ? big_list[idx] : (order ? order[idx] : idx);
    const int n = nv[s], nrhs = rv ? rv[s] : 0;
    if (NM == NSMALL && (n < 1 || n > NSMALL || nrhs < 0 || nrhs > RMAX)) continue;
    T* a = A + a_off[s];
    T* b = (mode != 1) ? B + b_off[s] : nullptr;
    int inf = 0;
    if (mode != 2) {
      inf = lu_factor<T, NM>(a, n, n, ipiv ? ipiv + p_off[s] : nullptr, sh.f, misc);
    } else {
      // permutation from the pivots: the interchanges i <-> ipiv[i] - 1 in order
      if (tid == 0) {
        const int32_t* pv = ipiv + p_off[s];
        for (int i = 0; i < n; ++i) misc.perm[i] = i;
        for (int i = 0; i < n; ++i) {
          const int q = pv[i] - 1;
          if (q != i && q >= 0 && q < n) {
            const int t = misc.perm[i];
            misc.perm[i] = misc.perm[q];
            misc.perm[q] = t;
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0 && info) info[s] = inf;
    if (inf == 0 && mode != 1) {
      for (int e = tid; e < n * RMAX; e += NT) {
        const int i = e / RMAX, c = e % RMAX;
        sh.s.x[e] = (c < nrhs) ? b[(long)misc.perm[i] * nrhs + c] : T(0);
      }
      __syncthreads();
      lu_solve<T>(a, n, n, sh.s.x, nrhs, sh.s.diag);
      for (int e = tid; e < n * nrhs; e += NT) b[e] = sh.s.x[(e / nrhs) * RMAX + e % nrhs];
    }
    __syncthreads();
  }
}

// ---- mixed precision: FP32 LU + FP64 refinement (dsgesv) --------------------------------
template <int NM>
struct IRShm {
  double x[NM * RMAX];        // current solution (FP64)
  float d[NM * RMAX];         // P r in FP32, solved in place into the correction
  float diag[16 * 17];
};

// Same work split as gesv_batched_kernel (NM = NSMALL: ordered range, skipping the others;
// NM = NMAX: big_list).
template <int NM>
__global__ __launch_bounds__(NT) void dsgesv_batched_kernel(
    int count, const int32_t* __restrict__ nv, const int32_t* __restrict__ rv,
    const int64_t* __restrict__ a_off, const int64_t* __restrict__ b_off,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ X,
    int32_t* __restrict__ iter, int32_t* __restrict__ info, const int32_t* __restrict__ order,
    const int32_t* __restrict__ big_list, const int* __restrict__ big_count,
    int* counter, float* __restrict__ ws32, int32_t* __restrict__ fallback_list,
    int* fallback_count, double tol, int max_iter) {
  __shared__ union {
    FactorShm<float, NM> f;
    IRShm<NM> ir;
  } sh;
  __shared__ Misc<NM> misc;
  const int tid = threadIdx.x;
  float* W = ws32 + (long)blockIdx.x * NM * NM;
  if (NM != NSMALL) count = *big_count;
  for (;;) {
    const int idx = next_system(counter, misc);
    if (idx >= count) break;
    const int s = (NM != NSMALL) ? big_list[idx] : (order ? order[idx] : idx);
    const int n = nv[s], nrhs = rv[s];
    if (NM == NSMALL && (n < 1 || n > NSMALL || nrhs < 0 || nrhs > RMAX)) continue;
    const double* a = A + a_off[s];
    const double* b = B + b_off[s];
    double* x = X + b_off[s];
    // FP32 copy of A (ld n) and ||A||_inf
    if (tid < 2 * RMAX) misc.cmax[tid] = 0ull;
    __syncthreads();
    {
      // 16 lanes per row, 16 rows per step (rows r0 + 4j), 64 loads per lane in flight per 256
      // columns; the row sums of |a| use a fixed order (lane-strided partial sums, then a
      // butterfly) so ||A||_inf and with it the stopping test are reproducible run to run
      const int r0 = tid / 16, ks = tid % 16;
      double mx = 0.0;
      for (int rb = 0; rb < n; rb += 16) {
        double sum[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 256) {
          double v[4][16];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int r = rb + r0 + 4 * j;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
              const int k = k0 + ks + 16 * u;
              v[j][u] = (r < n && k < n) ? a[(long)r * n + k] : 0.0;
            }
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int r = rb + r0 + 4 * j;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
              const int k = k0 + ks + 16 * u;
              if (r < n && k < n) W[(long)r * n + k] = (float)v[j][u];
              sum[j] += fabs(v[j][u]);
            }
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int off = 8; off >= 1; off >>= 1) sum[j] += __shfl_xor(sum[j], off, 64);
          mx = fmax(mx, sum[j]);
        }
      }
#pragma unroll
      for (int off = 32; off >= 16; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
      if (tid == 0) misc.cmax[2 * RMAX - 1] = (unsigned long long)__double_as_longlong(mx);
    }
    __syncthreads();
    const double anrm = __longlong_as_double((long long)misc.cmax[2 * RMAX - 1]);
    const double eps = tol > 0 ? tol : 0x1.0p-53;
    const double cte = anrm * eps * sqrt((double)n);
    __syncthreads();
    const int inf32 = lu_factor<float, NM>(W, n, n, nullptr, sh.f, misc);
    int it_out = -2;
    if (inf32 == 0) {
      // x = FP32 solve of P b
      for (int e = tid; e < n * RMAX; e += NT) {
        const int i = e / RMAX, c = e % RMAX;
        sh.ir.d[e] = (c < nrhs) ? (float)b[(long)misc.perm[i] * nrhs + c] : 0.0f;
      }
      __syncthreads();
      lu_solve<float>(W, n, n, sh.ir.d, nrhs, sh.ir.diag);
      for (int e = tid; e < n * RMAX; e += NT) sh.ir.x[e] = (double)sh.ir.d[e];
      for (int i = tid; i < n; i += NT) misc.pinv[misc.perm[i]] = i;
      __syncthreads();
      double prev = 1e308;
      int stalls = 0;
      it_out = -3;
      for (int it = 0; it <= max_iter; ++it) {
        // r = b - A x (FP64), column maxima of |r| and |x|
        if (tid < 2 * RMAX - 1) misc.cmax[tid] = 0ull;
        __syncthreads();
        for (int rb = 0; rb < n; rb += 16) {
          const int nr = min(16, n - rb);
          double acc[4][RMAX];
          rows_dot<16, double, double, double>(a, n, rb, nr, 0, n, sh.ir.x, nrhs, acc);
          const int r0 = tid / 16, ks = tid % 16;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int r = r0 + 4 * j;
            if (r < nr && ks < nrhs) {
              // lane ks of the row group writes column ks (all lanes hold every sum)
              double v = acc[j][0];
#pragma unroll
              for (int c = 1; c < RMAX; ++c)
                if (c == ks) v = acc[j][c];
              const double rv2 = b[(long)(rb + r) * nrhs + ks] - v;
              sh.ir.d[misc.pinv[rb + r] * RMAX + ks] = (float)rv2;  // P r for the correction
              atomicMax(&misc.cmax[ks], (unsigned long long)__double_as_longlong(fabs(rv2)));
              atomicMax(&misc.cmax[RMAX + ks],
                        (unsigned long long)__double_as_longlong(fabs(sh.ir.x[(rb + r) * RMAX + ks])));
            }
          }
        }
        __syncthreads();
        bool ok = true;
        double worst = 0.0;
        for (int c = 0; c < nrhs; ++c) {
          const double rn = __longlong_as_double((long long)misc.cmax[c]);
          const double xn = __longlong_as_double((long long)misc.cmax[RMAX + c]);
          if (!(rn <= xn * cte)) ok = false;
          worst = fmax(worst, rn);
        }
        if (ok) { it_out = it; break; }
        if (it == max_iter) break;
        // stagnation: the residual must at least halve every other step
        if (!(worst < 0.5 * prev)) {
          if (++stalls >= 2) break;
        } else {
          stalls = 0;
        }
        prev = worst;
        // d = FP32 solve of P r (written by the residual pass); x += d
        lu_solve<float>(W, n, n, sh.ir.d, nrhs, sh.ir.diag);
        for (int e = tid; e < n * RMAX; e += NT) sh.ir.x[e] += (double)sh.ir.d[e];
        __syncthreads();
      }
      if (it_out >= 0)
        for (int e = tid; e < n * nrhs; e += NT) x[e] = sh.ir.x[(e / nrhs) * RMAX + e % nrhs];
    }
    if (tid == 0) {
      iter[s] = it_out;
      info[s] = 0;
      if (it_out < 0) fallback_list[atomicAdd(fallback_count, 1)] = s;
    }
    __syncthreads();
  }
}

// FP64 fallback for the systems listed by the mixed-precision kernels: copy A to the
// workgroup's FP64 scratch, factor, solve B into X.  The NM = NSMALL instance takes the listed
// systems of order <= NSMALL, the NM = NMAX instance the others (each skips the rest).
template <int NM>
__global__ __launch_bounds__(NT) void gesv_fallback_kernel(
    const int32_t* __restrict__ nv, const int32_t* __restrict__ rv,
    const int64_t* __restrict__ a_off, const int64_t* __restrict__ b_off,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ X,
    int32_t* __restrict__ info, const int32_t* __restrict__ list, const int* list_count,
    int* counter, double* __restrict__ ws64) {
  __shared__ union {
    FactorShm<double, NM> f;
    SolveShm<double, NM> s;
  } sh;
  __shared__ Misc<NM> misc;
  const int tid = threadIdx.x;
  double* W = ws64 + (long)blockIdx.x * NM * NM;
  const int count = *list_count;
  for (;;) {
    const int idx = next_system(counter, misc);
    if (idx >= count) break;
    const int s = list[idx];
    const int n = nv[s], nrhs = rv[s];
    if ((NM == NSMALL) != (n <= NSMALL)) continue;
    const double* a = A + a_off[s];
    for (int e = tid; e < n * n; e += NT) W[e] = a[e];
    __syncthreads();
    const int inf = lu_factor<double, NM>(W, n, n, nullptr, sh.f, misc);
    if (tid == 0) info[s] = inf;
    if (inf == 0) {
      const double* b = B + b_off[s];
      for (int e = tid; e < n * RMAX; e += NT) {
        const int i = e / RMAX, c = e % RMAX;
        sh.s.x[e] = (c < nrhs) ? b[(long)misc.perm[i] * nrhs + c] : 0.0;
      }
      __syncthreads();
      lu_solve<double>(W, n, n, sh.s.x, nrhs, sh.s.diag);
      double* x = X + b_off[s];
      for (int e = tid; e < n * nrhs; e += NT) x[e] = sh.s.x[(e / nrhs) * RMAX + e % nrhs];
    }
    __syncthreads();
  }
}

__global__ void reset_counters_kernel(int* c, int n) {
  if (threadIdx.x < n) c[threadIdx.x] = 0;
}

// Marks systems outside the supported range (info = -1, iter = 0 when iter != null) and lists
// the systems of order NSMALL+1 .. NMAX for the NM = NMAX kernels.
__global__ __launch_bounds__(256) void classify_kernel(int count, const int32_t* __restrict__ nv,
                                                       const int32_t* __restrict__ rv,
                                                       int32_t* __restrict__ info,
                                                       int32_t* __restrict__ iter,
                                                       int32_t* __restrict__ big_list, int* big_count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  const int n = nv[s], nrhs = rv ? rv[s] : 0;
  if (n < 1 || n > NMAX || nrhs < 0 || nrhs > RMAX) {
    if (info) info[s] = -1;
    if (iter) iter[s] = 0;
  } else if (n > NSMALL) {
    big_list[atomicAdd(big_count, 1)] = s;
  }
}

}  // namespace

// Scratch of bzg_dgesv_batched_noalloc: the solver's counters and the list of systems of order
// > NSMALL, in a device array of fixed size that every call on one GPU shares.  Host code never
// passes the array's name to a kernel (in host code it is only a stand-in for the symbol): the
// device address comes from hipGetSymbolAddress, once per device, kept in g_noalloc_ptr.
constexpr int64_t kNoAllocMaxSystems = BZG_NOALLOC_MAX_SYSTEMS;
constexpr size_t kNoAllocCounterBytes = 256;
constexpr size_t kNoAllocBytes = kNoAllocCounterBytes + (size_t)kNoAllocMaxSystems * sizeof(int32_t);
constexpr int kMaxDevices = 64;

}  // namespace bzg

static __device__ __align__(256) unsigned char scratch[bzg::kNoAllocBytes];

namespace bzg {
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kMaxDevices] = {};

// Resolves this device's copy of `scratch` (the caller has made `device` current).  Called by
// bzg_create; idempotent per device.
bzg_status noalloc_scratch_init(int device) {
  BZG_REQUIRE(device >= 0 && device < kMaxDevices, BZG_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return BZG_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return BZG_ERR_HIP;
  }
  g_noalloc_ptr[device] = ptr;
  return BZG_OK;
}

static void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return g_noalloc_ptr[device];
}

// Occupancy of the persistent grids (workgroups per CU); bzg_solver_autotune picks them: the
// FP64 batched LU's, and for the mixed solver the grid of its FP64 fallback kernel (the on-chip
// class kernels always run at full residency, g_onchip_wg_per_cu).
struct SolverTuning {
  int wg_per_cu_fp64 = 8;
  int wg_per_cu_mixed = 8;
};
static SolverTuning g_tuning;
// Workgroups per CU of the on-chip mixed kernels for the order classes 1..32, 33..64, 65..128,
// 129..256 (at most what registers and LDS let reside; extra workgroups find the list drained).
static int g_onchip_wg_per_cu[4] = {12, 6, 2, 1};

}  // namespace bzg

extern "C" {

bzg_status bzg_batched_order(bzg_context* ctx, int64_t count, const int32_t* n_host,
                             int32_t* order_dev) {
  BZG_REQUIRE(ctx && (count == 0 || (n_host && order_dev)), BZG_ERR_INVALID_ARGUMENT,
              "batched_order: null argument");
  std::vector<int32_t> ord(count);
  // counting sort by descending n (stable): longest processing time first
  std::vector<int64_t> cnt(bzg::NMAX + 2, 0);
  for (int64_t i = 0; i < count; ++i) {
    const int v = std::min(std::max(n_host[i], 0), bzg::NMAX + 1);
    cnt[v]++;
  }
  std::vector<int64_t> start(bzg::NMAX + 2, 0);
  int64_t acc = 0;
  for (int v = bzg::NMAX + 1; v >= 0; --v) { start[v] = acc; acc += cnt[v]; }
  for (int64_t i = 0; i < count; ++i) {
    const int v = std::min(std::max(n_host[i], 0), bzg::NMAX + 1);
    ord[start[v]++] = (int32_t)i;
  }
  if (count) {
    BZG_HIP_CHECK(hipMemcpyAsync(order_dev, ord.data(), count * sizeof(int32_t),
                                 hipMemcpyHostToDevice, ctx->stream));
    BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  return BZG_OK;
}

}  // extern "C"

namespace bzg {
namespace {
// Batched LU entry points (MODE as in gesv_batched_kernel).
// Scratch of one call: counters[0] small-kernel work counter, [1] big count, [2] big-kernel
// work counter (at least 8 ints, cleared by a kernel); big = room for `count` system indices
// (the systems of order > NSMALL).  Allocates nothing.
bzg_status lu_batched_scratch(bzg_context* ctx, int mode, int64_t count, const int32_t* n, const int32_t* nrhs,
                              const int64_t* a_off, const int64_t* b_off, const int64_t* p_off, double* A,
                              double* B, int32_t* ipiv, int32_t* info, const int32_t* order, int* counters,
                              int32_t* big) {
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, ctx->stream, counters, 8);
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(classify_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream,
                     (int)count, n, nrhs, info, (int32_t*)nullptr, big, counters + 1);
  BZG_LAUNCH_CHECK();
  // the few systems of order > NSMALL first (their kernel holds 2 workgroups per CU), then
  // the rest at full occupancy; persistent grids never larger than the batch
  const int64_t cap = std::max<int64_t>(count, 1);
  hipLaunchKernelGGL((gesv_batched_kernel<double, NMAX>),
                     dim3((unsigned)std::min<int64_t>(ctx->num_cus * 2, cap)), dim3(NT), 0,
                     ctx->stream, (int)count, n, nrhs, a_off, b_off, p_off, A, B, ipiv, info, order,
                     (const int32_t*)big, (const int*)(counters + 1), counters + 2, mode);
  BZG_LAUNCH_CHECK();
  const int grid = (int)std::min<int64_t>(ctx->num_cus * g_tuning.wg_per_cu_fp64, cap);
  hipLaunchKernelGGL((gesv_batched_kernel<double, NSMALL>), dim3(grid), dim3(NT), 0,
                     ctx->stream, (int)count, n, nrhs, a_off, b_off, p_off, A, B, ipiv, info, order,
                     (const int32_t*)big, (const int*)(counters + 1), counters + 0, mode);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

// The same with the scratch cached in the context: slot 0 = counters, slot 3 = big list.
bzg_status lu_batched(bzg_context* ctx, int mode, int64_t count, const int32_t* n, const int32_t* nrhs,
                      const int64_t* a_off, const int64_t* b_off, const int64_t* p_off, double* A,
                      double* B, int32_t* ipiv, int32_t* info, const int32_t* order) {
  void *cnt = nullptr, *lst = nullptr;
  bzg_status st = workspace(ctx, 0, 256, &cnt);
  if (!st) st = workspace(ctx, 3, (size_t)2 * count * sizeof(int32_t), &lst);
  if (st) return st;
  ProfScope prof(ctx, mode == 0 ? "dgesv" : (mode == 1 ? "dgetrf" : "dgetrs"));
  return lu_batched_scratch(ctx, mode, count, n, nrhs, a_off, b_off, p_off, A, B, ipiv, info, order,
                            static_cast<int*>(cnt), static_cast<int32_t*>(lst));
}

// Determinants from the factors (one thread per system), detNxN's rule
// (blaze/math/expressions/DMatDetExpr.h:345-357): 0 when getrf reported a zero pivot, else the
// product, left to right from 1, of U's diagonal with the sign flipped at every row interchange.
// An empty system has determinant 0 (DMatDetExpr.h:400-402) and info 0; an order outside
// 1..NMAX keeps info -1 and gets NaN.
__global__ __launch_bounds__(256) void det_from_lu_kernel(int64_t count, const int32_t* __restrict__ n,
                                                          const int64_t* __restrict__ a_off,
                                                          const int64_t* __restrict__ p_off,
                                                          const double* __restrict__ LU,
                                                          const int32_t* __restrict__ ipiv,
                                                          int32_t* __restrict__ info, double* __restrict__ det) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const int ni = n[i];
  if (ni == 0) {
    det[i] = 0.0;
    info[i] = 0;
    return;
  }
  if (ni < 0 || ni > NMAX || info[i] < 0) {
    det[i] = __builtin_nan("");
    return;
  }
  if (info[i] > 0) {
    det[i] = 0.0;
    return;
  }
  const double* a = LU + a_off[i];
  const int32_t* p = ipiv + p_off[i];
  double d = 1.0;
  for (int k = 0; k < ni; ++k) {
    const double u = a[(int64_t)k * (ni + 1)];
    d *= (p[k] == k + 1) ? u : -u;
  }
  det[i] = d;
}

// The batched inverse solves A_i X = I one eight-column chunk at a time, every (system, chunk)
// pair a separate "virtual system" of one getrs launch (same factors, pivots and offsets as its
// system, its own right-hand-side block).  inv_rhs_kernel writes chunk v's identity columns
// [col[v], col[v] + nr) as the row-major n x nr block at B + b_off[v], nr = min(8, n - col[v]),
// and sets nrhs[v] = nr, or 0 when its system is singular (info[sys[v]] != 0).
__global__ __launch_bounds__(256) void inv_rhs_kernel(int64_t nv, const int32_t* __restrict__ n,
                                                      const int32_t* __restrict__ col,
                                                      const int32_t* __restrict__ sys,
                                                      const int32_t* __restrict__ info,
                                                      const int64_t* __restrict__ b_off, double* __restrict__ B,
                                                      int32_t* __restrict__ nrhs) {
  const int64_t v = blockIdx.x;
  if (v >= nv) return;
  const int ni = n[v], c0 = col[v];
  const int nr = (info[sys[v]] == 0) ? min(8, ni - c0) : 0;
  if (threadIdx.x == 0) nrhs[v] = nr;
  double* b = B + b_off[v];
  for (int e = threadIdx.x; e < ni * nr; e += 256) {
    const int r = e / nr, q = e - r * nr;
    b[e] = (r == c0 + q) ? 1.0 : 0.0;
  }
}

// info[i] = 0 for an empty system (n[i] = 0), whose factorisation getrf reports out of range.
__global__ __launch_bounds__(256) void empty_info_kernel(int64_t count, const int32_t* __restrict__ n,
                                                         int32_t* __restrict__ info) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count && n[i] == 0) info[i] = 0;
}

// A_i[:, col[v] + q] = X_v[:, q]: the chunk's solution written over the factors.
__global__ __launch_bounds__(256) void inv_scatter_kernel(int64_t nv, const int32_t* __restrict__ n,
                                                          const int32_t* __restrict__ col,
                                                          const int32_t* __restrict__ nrhs,
                                                          const int64_t* __restrict__ a_off,
                                                          const int64_t* __restrict__ b_off,
                                                          const double* __restrict__ B, double* __restrict__ A) {
  const int64_t v = blockIdx.x;
  if (v >= nv) return;
  const int nr = nrhs[v];
  if (nr == 0) return;
  const int ni = n[v], c0 = col[v];
  const double* b = B + b_off[v];
  double* a = A + a_off[v];
  for (int e = threadIdx.x; e < ni * nr; e += 256) {
    const int r = e / nr, q = e - r * nr;
    a[(int64_t)r * ni + c0 + q] = b[e];
  }
}
}  // namespace
}  // namespace bzg

extern "C" {

static bzg_status dinv_batched_impl(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                            const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info, const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dinv_batched: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "dinv_batched: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && a_off && p_off && A && ipiv && info, BZG_ERR_INVALID_ARGUMENT, "dinv_batched: null array");
  // orders and offsets on the host (one synchronisation) to size the scratch
  std::vector<int32_t> hn((size_t)count);
  std::vector<int64_t> ha((size_t)count);
  BZG_HIP_CHECK(hipMemcpyAsync(hn.data(), n, (size_t)count * 4, hipMemcpyDeviceToHost, ctx->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(ha.data(), a_off, (size_t)count * 8, hipMemcpyDeviceToHost, ctx->stream));
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  int64_t extent = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (hn[(size_t)i] >= 1 && hn[(size_t)i] <= bzg::NMAX) {
      BZG_REQUIRE(ha[(size_t)i] >= 0, BZG_ERR_INVALID_ARGUMENT, "dinv_batched: negative offset");
      extent = std::max<int64_t>(extent, ha[(size_t)i] + (int64_t)hn[(size_t)i] * hn[(size_t)i]);
    }
  }
  // A_i := P L U in place (dgetrf); an empty system then gets info 0 (dgetri's n = 0)
  bzg_status st = bzg::lu_batched(ctx, 1, count, n, nullptr, a_off, nullptr, p_off, A, nullptr, ipiv, info, order);
  if (st) return st;
  hipLaunchKernelGGL(bzg::empty_info_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, count,
                     n, info);
  BZG_LAUNCH_CHECK();
  // virtual systems: one per (system, eight-column chunk) of every system in range
  std::vector<int32_t> vn, vcol, vsys;
  std::vector<int64_t> va, vp, vb;
  int64_t bsz_v = 0;
  std::vector<int64_t> hp((size_t)count);
  BZG_HIP_CHECK(hipMemcpyAsync(hp.data(), p_off, (size_t)count * 8, hipMemcpyDeviceToHost, ctx->stream));
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i < count; ++i) {
    const int ni = hn[(size_t)i];
    if (ni < 1 || ni > bzg::NMAX) continue;
    for (int c0 = 0; c0 < ni; c0 += 8) {
      vn.push_back(ni);
      vcol.push_back(c0);
      vsys.push_back((int32_t)i);
      va.push_back(ha[(size_t)i]);
      vp.push_back(hp[(size_t)i]);
      vb.push_back(bsz_v);
      bsz_v += (int64_t)ni * std::min(8, ni - c0);
    }
  }
  const int64_t nv = (int64_t)vn.size();
  if (nv == 0) return BZG_OK;
  BZG_REQUIRE(nv < (1LL << 31), BZG_ERR_UNSUPPORTED, "dinv_batched: too many column chunks");
  // scratch (slot 9): a copy of the factors, the right-hand-side blocks, the virtual systems'
  // n / col / sys / nrhs / getrs info (int32) and a_off / p_off / b_off (int64), their order
  const size_t lu_b = (size_t)extent * 8, rhs_b = (size_t)bsz_v * 8;
  const size_t i32_b = ((size_t)nv * 4 + 7) / 8 * 8, i64_b = (size_t)nv * 8;
  void* w = nullptr;
  st = bzg::workspace(ctx, 9, lu_b + rhs_b + 6 * i32_b + 3 * i64_b, &w);
  if (st) return st;
  char* cur = static_cast<char*>(w);
  auto take = [&](size_t bytes) { char* q = cur; cur += bytes; return q; };
  double* LUw = reinterpret_cast<double*>(take(lu_b));
  double* Bw = reinterpret_cast<double*>(take(rhs_b));
  int32_t* dn_v = reinterpret_cast<int32_t*>(take(i32_b));
  int32_t* dcol = reinterpret_cast<int32_t*>(take(i32_b));
  int32_t* dsys = reinterpret_cast<int32_t*>(take(i32_b));
  int32_t* dnr = reinterpret_cast<int32_t*>(take(i32_b));
  int32_t* dinfo2 = reinterpret_cast<int32_t*>(take(i32_b));
  int32_t* dord = reinterpret_cast<int32_t*>(take(i32_b));
  int64_t* da_v = reinterpret_cast<int64_t*>(take(i64_b));
  int64_t* dp_v = reinterpret_cast<int64_t*>(take(i64_b));
  int64_t* db_v = reinterpret_cast<int64_t*>(take(i64_b));
  BZG_HIP_CHECK(hipMemcpyAsync(dn_v, vn.data(), (size_t)nv * 4, hipMemcpyHostToDevice, ctx->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(dcol, vcol.data(), (size_t)nv * 4, hipMemcpyHostToDevice, ctx->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(dsys, vsys.data(), (size_t)nv * 4, hipMemcpyHostToDevice, ctx->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(da_v, va.data(), i64_b, hipMemcpyHostToDevice, ctx->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(dp_v, vp.data(), i64_b, hipMemcpyHostToDevice, ctx->stream));
  BZG_HIP_CHECK(hipMemcpyAsync(db_v, vb.data(), i64_b, hipMemcpyHostToDevice, ctx->stream));
  st = bzg_batched_order(ctx, nv, vn.data(), dord);  // longest systems' chunks first
  if (st) return st;
  if (extent > 0) BZG_HIP_CHECK(hipMemcpyAsync(LUw, A, lu_b, hipMemcpyDeviceToDevice, ctx->stream));
  // X = A^-1 I: every chunk of every system in one dgetrs launch over the copied factors, then
  // written over the factors in A
  hipLaunchKernelGGL(bzg::inv_rhs_kernel, dim3((unsigned)nv), dim3(256), 0, ctx->stream, nv, dn_v, dcol, dsys,
                     (const int32_t*)info, (const int64_t*)db_v, Bw, dnr);
  BZG_LAUNCH_CHECK();
  st = bzg::lu_batched(ctx, 2, nv, dn_v, dnr, da_v, db_v, dp_v, LUw, Bw, ipiv, dinfo2, dord);
  if (st) return st;
  hipLaunchKernelGGL(bzg::inv_scatter_kernel, dim3((unsigned)nv), dim3(256), 0, ctx->stream, nv, dn_v, dcol,
                     (const int32_t*)dnr, (const int64_t*)da_v, (const int64_t*)db_v, (const double*)Bw, A);
  BZG_LAUNCH_CHECK();
  BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host copy of b_off is released on return
  return BZG_OK;
}

bzg_status bzg_dinv_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                            const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info, const int32_t* order) {
  // no C++ exception leaves the C ABI: a host allocation failure becomes a status
  try {
    return dinv_batched_impl(ctx, count, n, a_off, p_off, A, ipiv, info, order);
  } catch (const std::bad_alloc&) {
    bzg::set_error("dinv_batched: out of host memory");
    return BZG_ERR_OUT_OF_MEMORY;
  }
}

bzg_status bzg_ddet_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                            const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info, double* det,
                            const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "ddet_batched: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "ddet_batched: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && a_off && p_off && A && ipiv && info && det, BZG_ERR_INVALID_ARGUMENT,
              "ddet_batched: null array");
  bzg_status st = bzg::lu_batched(ctx, 1, count, n, nullptr, a_off, nullptr, p_off, A, nullptr, ipiv, info, order);
  if (st) return st;
  hipLaunchKernelGGL(bzg::det_from_lu_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream,
                     count, n, a_off, p_off, A, ipiv, info, det);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}


bzg_status bzg_dgesv_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int32_t* nrhs,
                             const int64_t* a_off, const int64_t* b_off, const int64_t* p_off,
                             double* A, double* B, int32_t* ipiv, int32_t* info,
                             const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgesv_batched: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "dgesv_batched: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && nrhs && a_off && b_off && A && B && info, BZG_ERR_INVALID_ARGUMENT,
              "dgesv_batched: null array");
  BZG_REQUIRE(ipiv == nullptr || p_off != nullptr, BZG_ERR_INVALID_ARGUMENT,
              "dgesv_batched: ipiv without p_off");
  return bzg::lu_batched(ctx, 0, count, n, nrhs, a_off, b_off, p_off, A, B, ipiv, info, order);
}

bzg_status bzg_dgesv_batched_noalloc(bzg_context* ctx, int64_t count, const int32_t* n, const int32_t* nrhs,
                                     const int64_t* a_off, const int64_t* b_off, const int64_t* p_off,
                                     double* A, double* B, int32_t* ipiv, int32_t* info, const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgesv_batched_noalloc: null context");
  BZG_REQUIRE(count >= 0, BZG_ERR_INVALID_ARGUMENT, "dgesv_batched_noalloc: bad count");
  BZG_REQUIRE(count <= bzg::kNoAllocMaxSystems, BZG_ERR_UNSUPPORTED,
              "dgesv_batched_noalloc: more than BZG_NOALLOC_MAX_SYSTEMS (1048576) systems in one call; "
              "split the batch or use bzg_dgesv_batched");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && nrhs && a_off && b_off && A && B && info, BZG_ERR_INVALID_ARGUMENT,
              "dgesv_batched_noalloc: null array");
  BZG_REQUIRE(ipiv == nullptr || p_off != nullptr, BZG_ERR_INVALID_ARGUMENT,
              "dgesv_batched_noalloc: ipiv without p_off");
  char* base = static_cast<char*>(bzg::noalloc_scratch(ctx->device));
  BZG_REQUIRE(base != nullptr, BZG_ERR_HIP, "dgesv_batched_noalloc: scratch not resolved on this device");
  int* counters = reinterpret_cast<int*>(base);
  int32_t* big = reinterpret_cast<int32_t*>(base + bzg::kNoAllocCounterBytes);
  return bzg::lu_batched_scratch(ctx, 0, count, n, nrhs, a_off, b_off, p_off, A, B, ipiv, info, order,
                                 counters, big);
}

bzg_status bzg_dgetrf_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                              const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info,
                              const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgetrf_batched: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "dgetrf_batched: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && a_off && p_off && A && ipiv && info, BZG_ERR_INVALID_ARGUMENT,
              "dgetrf_batched: null array");
  return bzg::lu_batched(ctx, 1, count, n, nullptr, a_off, nullptr, p_off, A, nullptr, ipiv, info, order);
}

bzg_status bzg_dgetrs_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int32_t* nrhs,
                              const int64_t* a_off, const int64_t* b_off, const int64_t* p_off,
                              const double* LU, const int32_t* ipiv, double* B, int32_t* info,
                              const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dgetrs_batched: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "dgetrs_batched: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && nrhs && a_off && b_off && p_off && LU && ipiv && B && info, BZG_ERR_INVALID_ARGUMENT,
              "dgetrs_batched: null array");
  // the kernel only reads the factors and the pivots in this mode
  return bzg::lu_batched(ctx, 2, count, n, nrhs, a_off, b_off, p_off, const_cast<double*>(LU), B,
                         const_cast<int32_t*>(ipiv), info, order);
}

bzg_status bzg_dsgesv_batched(bzg_context* ctx, int64_t count, const int32_t* n,
                              const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                              const double* A, const double* B, double* X, int32_t* iter,
                              int32_t* info, double tol, int32_t max_iter, const int32_t* order) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "dsgesv_batched: null context");
  BZG_REQUIRE(count >= 0 && count < (1LL << 31), BZG_ERR_INVALID_ARGUMENT, "dsgesv_batched: bad count");
  if (count == 0) return BZG_OK;
  BZG_REQUIRE(n && nrhs && a_off && b_off && A && B && X && iter && info, BZG_ERR_INVALID_ARGUMENT,
              "dsgesv_batched: null array");
  BZG_REQUIRE(max_iter >= 0 && max_iter <= 1000, BZG_ERR_INVALID_ARGUMENT, "dsgesv_batched: bad max_iter");
  // Orders 1..256 run on chip (lu_mixed.hip, four order classes), orders 257..512 in the
  // streaming kernel below; systems that do not converge fall back to the FP64 LU.
  // Persistent grids never larger than the batch; the workspaces are sized for the full grids
  // so that they are not regrown as batch sizes vary.
  const int64_t cap = std::max<int64_t>(count, 1);
  const int grid_big_full = ctx->num_cus * 2;
  const int grid_fb_full = ctx->num_cus * bzg::g_tuning.wg_per_cu_mixed;  // FP64 fallback grid
  const int grid_big = (int)std::min<int64_t>(grid_big_full, cap);
  const int grid_fb = (int)std::min<int64_t>(grid_fb_full, cap);
  // scratch: slot 0 counters ([0..3] on-chip class work counters, [4] fallback count, [5] and
  // [6] fallback work counters, [7] order-257..512 work counter, [8..12] class sizes);
  // slot 3 the five class lists (count entries each) and the fallback list; slot 5 FP32
  // matrices of the order-257..512 kernel's workgroups; slot 2 FP64 matrices of the fallback
  // kernels' workgroups
  void *cnt = nullptr, *w32b = nullptr, *w64 = nullptr, *lst = nullptr;
  const size_t nn_small = (size_t)bzg::NSMALL * bzg::NSMALL, nn_big = (size_t)bzg::NMAX * bzg::NMAX;
  const size_t w64_bytes = std::max((size_t)grid_fb_full * nn_small, (size_t)grid_big_full * nn_big) * sizeof(double);
  bzg_status st = bzg::workspace(ctx, 0, 256, &cnt);
  if (!st) st = bzg::workspace(ctx, 5, (size_t)grid_big_full * nn_big * sizeof(float), &w32b);
  if (!st) st = bzg::workspace(ctx, 2, w64_bytes, &w64);
  if (!st) st = bzg::workspace(ctx, 3, (size_t)6 * count * sizeof(int32_t), &lst);
  if (st) return st;
  int* counters = static_cast<int*>(cnt);
  int32_t* lists = static_cast<int32_t*>(lst);
  int32_t* fb_list = lists + 5 * count;
  bzg::ProfScope prof(ctx, "dsgesv");
  hipLaunchKernelGGL(bzg::reset_counters_kernel, dim3(1), dim3(64), 0, ctx->stream, counters, 16);
  BZG_LAUNCH_CHECK();
  st = bzg::launch_partition(ctx, count, n, nrhs, order, info, iter, lists, (int)count, counters + 8);
  if (st) return st;
  hipLaunchKernelGGL(bzg::dsgesv_batched_kernel<bzg::NMAX>, dim3(grid_big), dim3(bzg::NT), 0, ctx->stream,
                     (int)count, n, nrhs, a_off, b_off, A, B, X, iter, info, order,
                     (const int32_t*)(lists + 4 * count), (const int*)(counters + 12), counters + 7,
                     static_cast<float*>(w32b), fb_list, counters + 4, tol, max_iter);
  BZG_LAUNCH_CHECK();
  st = bzg::launch_dsgesv_onchip(ctx, lists, (int)count, counters + 8, n, nrhs, a_off, b_off, A, B, X, iter, info,
                                 counters, fb_list, counters + 4, tol, max_iter, bzg::g_onchip_wg_per_cu, count);
  if (st) return st;
  hipLaunchKernelGGL(bzg::gesv_fallback_kernel<bzg::NSMALL>, dim3(grid_fb), dim3(bzg::NT), 0, ctx->stream, n,
                     nrhs, a_off, b_off, A, B, X, info, static_cast<const int32_t*>(fb_list), counters + 4,
                     counters + 5, static_cast<double*>(w64));
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(bzg::gesv_fallback_kernel<bzg::NMAX>, dim3(grid_big), dim3(bzg::NT), 0, ctx->stream, n,
                     nrhs, a_off, b_off, A, B, X, info, static_cast<const int32_t*>(fb_list), counters + 4,
                     counters + 6, static_cast<double*>(w64));
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

bzg_status bzg_solver_autotune(bzg_context* ctx, int32_t mixed, int64_t count, const int32_t* n,
                               const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                               const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                               const int32_t* order, int32_t* wg_per_cu, float* ms_out) {
  BZG_REQUIRE(ctx && count > 0 && n && nrhs && a_off && b_off && A && B && a_elems > 0 && b_elems > 0,
              BZG_ERR_INVALID_ARGUMENT, "solver_autotune: bad argument");
  // scratch: copies of A and B (the FP64 solver works in place) or X, plus iter / info
  const size_t bytes = (size_t)(mixed ? 0 : a_elems) * 8 + (size_t)b_elems * 8 + (size_t)count * 8 + 256;
  void* ws = nullptr;
  bzg_status st = bzg::workspace(ctx, 7, bytes, &ws);
  if (st) return st;
  double* Aw = static_cast<double*>(ws);
  double* Bw = Aw + (mixed ? 0 : a_elems);
  int32_t* it = reinterpret_cast<int32_t*>(Bw + b_elems);
  int32_t* inf = it + count;
  hipEvent_t e0, e1;
  BZG_HIP_CHECK(hipEventCreate(&e0));
  BZG_HIP_CHECK(hipEventCreate(&e1));
  const bool prof = ctx->profile;
  ctx->profile = false;
  int& knob = mixed ? bzg::g_tuning.wg_per_cu_mixed : bzg::g_tuning.wg_per_cu_fp64;
  const int saved = knob;
  int best = -1;
  float best_ms = 0.f;
  // the FP64 solver overwrites its operands: fresh copies before every run, outside the timing
  auto restore = [&]() -> bzg_status {
    if (mixed) return BZG_OK;
    if (hipMemcpyAsync(Aw, A, (size_t)a_elems * 8, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(Bw, B, (size_t)b_elems * 8, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
      return BZG_ERR_HIP;
    return BZG_OK;
  };
  auto run = [&]() -> bzg_status {
    if (mixed)
      return bzg_dsgesv_batched(ctx, count, n, nrhs, a_off, b_off, A, B, Bw, it, inf, 0.0, 30, order);
    return bzg_dgesv_batched(ctx, count, n, nrhs, a_off, b_off, nullptr, Aw, Bw, nullptr, inf, order);
  };
  for (int cand : {1, 2, 3, 4, 6, 8}) {
    knob = cand;
    st = restore();
    if (!st) st = run();  // warm-up (and workspace growth)
    if (!st) st = restore();
    if (st) break;
    hipEventRecord(e0, ctx->stream);
    st = run();
    hipEventRecord(e1, ctx->stream);
    if (st || hipEventSynchronize(e1) != hipSuccess) {
      if (!st) st = BZG_ERR_HIP;
      break;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (best < 0 || ms < best_ms) {
      best = cand;
      best_ms = ms;
    }
  }
  ctx->profile = prof;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  knob = st ? saved : best;
  if (st) return st;
  if (wg_per_cu) *wg_per_cu = best;
  if (ms_out) *ms_out = best_ms;
  return BZG_OK;
}

}  // extern "C"
