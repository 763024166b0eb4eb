// Batched mixed-precision solves with the FP32 matrix on chip (orders 1..256): FP32 LU with
// partial pivoting + FP64 iterative refinement, LAPACK dsgesv semantics.
//
// Replaces, per system, solve(A, X, B) (blaze/math/dense/LSE.h:6452-6482 -> solveNxN, LSE.h:5733-5770
// -> gesv -> dgesv_, blaze/math/lapack/clapack/gesv.h:207) for the systems of a batch; the
// result is the FP64 solution to the caller's tolerance (the factorisation is FP32, the residuals
// and the solution FP64).
//
// One workgroup owns one system at a time (persistent grid, systems pulled from a work counter).
// Orders are padded to a multiple of 16 (identity below / right of A) and split into 16 x 16
// tiles.  For an order class NM (32, 64, 128, 256) the workgroup has W = max(1, NM/32) wavefronts
// and the whole FP32 matrix lives in their registers (NM/2 floats per lane):
//   factorisation: wavefront w owns tile columns J = w and J = w + W (column-panel ownership),
//     each tile held as the 16 x 16 accumulator of v_mfma_f32_16x16x4f32 (lane l, register r =
//     element (4(l/16) + r, l % 16)).  Right-looking blocked LU, panel = one tile column:
//       - the owner of panel p factors it in registers (row-per-lane layout).  It first tries
//         every pivot on the diagonal, each lane checking that no candidate row beats the
//         pivot; when that holds (always for diagonally dominant matrices) the result is the
//         partial-pivoting one without a search, and the panel has no interchanges.  Otherwise
//         the panel is factored again with the search (a per-lane max and a butterfly; rows
//         keep logical positions and are written in order at the end), and the system's later
//         panels search directly.  The owner writes L21 and L11^-1 to LDS and, after a search,
//         the table of the panel's row interchanges;
//       - one workgroup barrier;
//       - every wavefront applies the interchanges (if any) to its own tile columns (the affected rows
//         go through a per-wavefront LDS staging area), forms U12 = L11^-1 A12 with one MFMA
//         group per tile and updates its trailing tiles A22 -= L21 U12 with MFMAs whose B
//         operand is U12 itself (the accumulator layout of a tile is the B-operand layout of the
//         next MFMA when both operands index K as 4(l/16) + kk);
//       - the owner of panel p+1 updates that panel first, factors it, and only then updates its
//         other tile column (look-ahead: the next panel is factored while the other wavefronts
//         still update); in the classes of 128 and 256 that other update waits until the start
//         of step p+1, where the owner of panel p+2 is on the critical path instead, so the
//         factors of a panel are kept for two steps (three panel buffers).
//   solves: the factors are redistributed through LDS (two rounds) to row-block ownership
//     (wavefront w owns tile rows I = w and w + W), held as MFMA A operands.  Forward and
//     backward substitution run block by block: the owner of block s forms y_s = L_ss^-1 (c_s -
//     sum_J L_sJ y_J) with the explicit 16 x 16 diagonal-block inverse (one MFMA group), the
//     other wavefronts add L_Is y_s to their running sums; one barrier per block.
//   refinement: r = b - A x in FP64 on the FP64 matrix cores (v_mfma_f64_16x16x4f64, A streamed
//     from HBM, x in LDS), written as FP32 P r for the next correction; ||A||_inf from the same
//     pass; LAPACK dsgesv's stopping test per right-hand side.
// Systems whose FP32 factorisation meets an exactly-zero pivot, or whose refinement does not
// converge (or stagnates), are listed for the FP64 LU (lu.hip).
#include "common.hpp"

#include <algorithm>

namespace bzg {
namespace mx {

constexpr int RMAX = 8;  // largest number of right-hand sides

template <int NM>
struct Cfg {
  static constexpr int NT = NM / 16;                                   // tiles per dimension
  static constexpr int W = NM >= 64 ? NM / 32 : 1;                     // wavefronts per workgroup
  static constexpr int NTC = NT / W;                                   // tile columns (rows) per wavefront
  static constexpr int Q = NM >= 64 ? NM / 64 : 1;                     // panel rows per lane
  static constexpr int THREADS = 64 * W;
  // resident wavefronts per SIMD the register budget is sized for (256, 256, 168, 128 registers)
  static constexpr int WAVES_PER_EU = NM >= 128 ? 2 : (NM == 64 ? 3 : 4);
  // The owner of panel p + 1 leaves its other tile column's update with panel p to the next
  // step (DEFER), so panel p's factors must outlive step p: NBUF panel buffers, by p % NBUF.
  // Only for the multi-wavefront classes whose LDS budget allows the third buffer.
  static constexpr bool DEFER = NM >= 128;
  static constexpr int NBUF = DEFER ? 3 : 2;
};

typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mma32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4 mma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// The thread index as a value the compiler cannot treat as loop-invariant: every per-lane
// quantity derived from it is recomputed where it is used instead of being hoisted out of the
// persistent loop (and the panel and refinement loops) and held in registers across them.
__device__ __forceinline__ int opaque_tid(int t) {
  __asm__ volatile("" : "+v"(t));
  return t;
}

// Orders LDS accesses of one wavefront (writes by some lanes, then reads by others).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

template <int NM>
struct Shm {
  static constexpr int NT = Cfg<NM>::NT, W = Cfg<NM>::W;
  union {
    struct {
      float prw[16];           // pivot row being broadcast (panel factorisation)
      float pbuf[Cfg<NM>::NBUF][NM][20];  // factored panel p (rows at final positions), by p % NBUF
      float linv[Cfg<NM>::NBUF][16][20];  // L11^-1 of panel p, by p % NBUF
      float stg[W][32][16 * Cfg<NM>::NTC + 1];  // per-wavefront interchange staging: 32 row slots
    } f;
    float rbuf[2][W][W][256];  // redistribution rounds: [I - h W][J - jc W] 16 x 16 tiles, swizzled
    float ld[W][4][16][20];    // loading A: four row-major FP32 tiles per wavefront
    struct {
      double x64[NM][RMAX];    // FP64 solution
      float cbuf[NM][16];      // right-hand side / solution blocks (columns >= nrhs are zero)
      float dscr[W][Cfg<NM>::NTC][16][20];  // diagonal blocks being inverted (16-byte aligned rows)
      float dinv[W][Cfg<NM>::NTC][16][33];  // their L^-1 (columns 0..15) and U^-1 (columns 16..31)
    } s;
  } u;
  // interchange tables of panel p, by p % NBUF
  signed char slot[Cfg<NM>::NBUF][NM];     // row -> interchange slot (-1: not affected)
  signed char srcslot[Cfg<NM>::NBUF][32];  // slot of the row whose values move into slot s
  int dlist[Cfg<NM>::NBUF][32];            // rows of the slots
  int dcount[Cfg<NM>::NBUF];
  int piv[Cfg<NM>::NBUF][16];              // absolute pivot rows
  int perm[NM];                // row i of the factored system = original row perm[i]
  int pinv[NM];
  unsigned long long cmax[2 * RMAX + 2];  // |r|, |x| column maxima, ||A||_inf (bit patterns)
  int sys;
  int info;
  int noswap[Cfg<NM>::NBUF];  // panel p kept every pivot on the diagonal: no interchanges to apply
  int spec;       // 1 while the diagonal-pivot attempt has not failed on this system
};

extern "C" __device__ __attribute__((const)) unsigned __ockl_wfred_max_u32(unsigned);
extern "C" __device__ __attribute__((const)) unsigned __ockl_wfred_min_u32(unsigned);

// Column c of the inverse of a 16 x 16 unit lower triangular L (row-major, leading dimension
// ldl, a multiple of 4, 16-byte aligned rows; the diagonal and upper part are not read), by
// forward substitution.  The rows are read four values per load, independently of the
// recurrence, which runs in registers; returns the column in x.  No scheduling barriers: row
// i + 1's sum can start while row i's finishes (it needs x[i] only at its last step).
__device__ __forceinline__ void unit_lower_inverse_column(const float* L, int ldl, int c, float (&x)[16]) {
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    float row[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (4 * q < i) {
        const float4v v = *reinterpret_cast<const float4v*>(L + i * ldl + 4 * q);
        row[4 * q] = v[0]; row[4 * q + 1] = v[1]; row[4 * q + 2] = v[2]; row[4 * q + 3] = v[3];
      }
    }
    float v = (i == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < i; ++k) v -= row[k] * x[k];
    x[i] = v;
  }
}

// Column c of the inverse of a 16 x 16 upper triangular U (row-major, ldu as above), by back
// substitution; returns the column in x.
__device__ __forceinline__ void upper_inverse_column(const float* U, int ldu, int c, float (&x)[16]) {
#pragma unroll
  for (int i = 15; i >= 0; --i) {
    float row[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (4 * q + 3 >= i) {
        const float4v v = *reinterpret_cast<const float4v*>(U + i * ldu + 4 * q);
        row[4 * q] = v[0]; row[4 * q + 1] = v[1]; row[4 * q + 2] = v[2]; row[4 * q + 3] = v[3];
      }
    }
    float v = (i == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = i + 1; k < 16; ++k) v -= row[k] * x[k];
    x[i] = v / row[i];
  }
}

// ---- panel factorisation (owner wavefront; panel column JC of its register tiles) -------------
// Factors the tile column p (rows j0 = 16p .. n_pad - 1) with partial pivoting (first row of
// maximal |a|; a NaN ranks below every number), writes the factored panel to pbuf[par] and back
// into the tiles, L11^-1 to linv[par], the pivots, the interchange tables and the permutation.
// Returns the 1-based column of the first exactly-zero pivot in this panel, or 0.
template <int NM, int JC>
__device__ __forceinline__ int factor_panel(f4 (&T)[Cfg<NM>::NT][Cfg<NM>::NTC], int p, int NTn, Shm<NM>& sh, int par, int tid) {
  constexpr int NT = Cfg<NM>::NT, Q = Cfg<NM>::Q;
  const int lane = tid & 63;
  const int g = lane >> 4, c16 = lane & 15;
  const int j0 = 16 * p;
  const int m = 16 * NTn - j0;  // rows of the panel
  float(*pb)[20] = sh.u.f.pbuf[par];
  // the whole tile column -> pbuf (absolute rows; rows above j0 hold U and pass through), so
  // that its registers are free for the panel while it is factored
#pragma unroll
  for (int I = 0; I < NT; ++I) {
#pragma unroll
    for (int r = 0; r < 4; ++r) pb[16 * I + 4 * g + r][c16] = T[I][JC][r];
  }
  wave_sync();
  float P[Q][16];
  int lpos[Q];
  const int qm = (m + 63) >> 6;  // row blocks of 64 that hold panel rows (uniform)
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int rr = lane + 64 * q;
    lpos[q] = rr < m ? rr : (1 << 20);
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      float4v v = {0.f, 0.f, 0.f, 0.f};
      if (rr < m) v = *reinterpret_cast<const float4v*>(&pb[j0 + rr][4 * c4]);
      P[q][4 * c4] = v[0]; P[q][4 * c4 + 1] = v[1]; P[q][4 * c4 + 2] = v[2]; P[q][4 * c4 + 3] = v[3];
    }
  }
  // First attempt (while no panel of this system has failed it): every pivot on the diagonal.
  // Row jj of the panel sits in lane jj, so a column step needs no search; each lane checks
  // that its candidate rows' |a| do not exceed |pivot| (a tie keeps the first row, the
  // diagonal), and a zero, infinite or NaN pivot fails the attempt.  If no lane saw a larger
  // entry, the diagonal was the partial-pivoting choice at every step and the arithmetic below
  // is that of the searching loop with those pivots, bit for bit; otherwise the panel is
  // reloaded and factored by the searching loop, and later panels of the system skip the attempt.
  bool fast = false;
  if (sh.spec) {
    unsigned bad = 0u;
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
      float prow[16];
#pragma unroll
      for (int c = jj; c < 16; ++c)
        prow[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(P[0][c]), jj));
      const unsigned pbits = __float_as_uint(prow[jj]) & 0x7FFFFFFFu;
      bad |= (pbits == 0u || pbits >= 0x7F800000u) ? 1u : 0u;
      const float rinv = __builtin_amdgcn_rcpf(prow[jj]);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (q >= qm) break;
        const int rr = lane + 64 * q;
        if (rr > jj && rr < m) {
          const float a = P[q][jj];
          bad |= ((__float_as_uint(a) & 0x7FFFFFFFu) > pbits) ? 1u : 0u;
          const float l = a * rinv;
          P[q][jj] = l;
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c > jj) P[q][c] -= l * prow[c];
        }
      }
    }
    fast = __ballot(bad != 0u) == 0ull;
    if (!fast) {
      // reload the panel for the searching loop
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int rr = lane + 64 * q;
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
          float4v v = {0.f, 0.f, 0.f, 0.f};
          if (rr < m) v = *reinterpret_cast<const float4v*>(&pb[j0 + rr][4 * c4]);
          P[q][4 * c4] = v[0]; P[q][4 * c4 + 1] = v[1]; P[q][4 * c4 + 2] = v[2]; P[q][4 * c4 + 3] = v[3];
        }
      }
      if (lane == 0) sh.spec = 0;
    }
  }
  if (lane == 0) sh.noswap[par] = fast ? 1 : 0;
  int info = 0;
  if (!fast) {
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) {
    // candidates: logical rows jj .. m-1, ranked by v = |a| bits + 2 (1 for NaN, 0 for none):
    // per lane the best of its rows by the 64-bit key (v, -row), so the first logical row of
    // maximal v wins; then the wavefront maximum of v and the lane holding it (a second
    // reduction only when several lanes hold the maximum)
    unsigned long long kb = 0ull;
    int bq = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (q >= qm) break;
      const bool cand = lpos[q] >= jj && lpos[q] < m;
      const float a = P[q][jj];
      unsigned hi = (__float_as_uint(a) & 0x7FFFFFFFu) + 2u;
      hi = (a != a) ? 1u : hi;
      hi = cand ? hi : 0u;
      const unsigned long long key = ((unsigned long long)hi << 32) | (unsigned)(~lpos[q]);
      const bool better = key > kb;
      kb = better ? key : kb;
      bq = better ? q : bq;
    }
    const unsigned vb = (unsigned)(kb >> 32);
    const int lb = (int)~(unsigned)kb;
    const unsigned vmax = __ockl_wfred_max_u32(vb);
    unsigned long long hold = __ballot(vb == vmax);
    int pr;
    if (__popcll(hold) == 1) {
      pr = __builtin_amdgcn_readlane(lb, __ffsll((long long)hold) - 1);
    } else {
      pr = (int)__ockl_wfred_min_u32(vb == vmax ? (unsigned)lb : 0xFFFFFFFFu);
      pr = __builtin_amdgcn_readfirstlane(pr);
      hold = __ballot(vb == vmax && lb == pr);
    }
    pr = __builtin_amdgcn_readfirstlane(pr);
    const int pl = __builtin_amdgcn_readfirstlane(__ffsll((long long)hold) - 1);
    const int pq = __builtin_amdgcn_readlane(bq, pl);
    // the pivot row from column jj on, read from its lane into scalar registers (a uniform
    // branch selects the register row that holds it)
    float prow[16];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (pq == q) {
#pragma unroll
        for (int c = jj; c < 16; ++c)
          prow[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(P[q][c]), pl));
      }
    }
    const float piv = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(prow[jj])));
    if (piv != 0.0f) {
      if (lane == 0) sh.piv[par][jj] = j0 + pr;
#pragma unroll
      for (int q = 0; q < Q; ++q) lpos[q] = (lpos[q] == jj) ? pr : ((lpos[q] == pr) ? jj : lpos[q]);
      const float rinv = __builtin_amdgcn_rcpf(piv);  // reciprocal, then scale (sgetf2)
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (q >= qm) break;
        if (lpos[q] > jj && lpos[q] < m) {
          const float l = P[q][jj] * rinv;
          P[q][jj] = l;
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c > jj) P[q][c] -= l * prow[c];
        }
      }
    } else {
      // exactly-zero pivot (dgetf2): no interchange, no scaling; the column below is zero
      if (lane == 0) sh.piv[par][jj] = j0 + jj;
      if (info == 0) info = j0 + jj + 1;
    }
  }
  }
  // rows to their final positions (pbuf), then back into the tiles
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int rr = lane + 64 * q;
    if (rr < m) {
#pragma unroll
      for (int c4 = 0; c4 < 4; ++c4)
        *reinterpret_cast<float4v*>(&pb[j0 + lpos[q]][4 * c4]) =
            float4v{P[q][4 * c4], P[q][4 * c4 + 1], P[q][4 * c4 + 2], P[q][4 * c4 + 3]};
    }
  }
  wave_sync();
  // L11^-1 (unit lower): lane c < 16 solves for column c, in place in linv (before the tiles
  // are reloaded, while the panel's registers are free)
  if (lane < 16) {
    float x[16];
    unit_lower_inverse_column(&pb[j0][0], 20, lane, x);
#pragma unroll
    for (int i = 0; i < 16; ++i) sh.u.f.linv[par][i][lane] = x[i];
  }
#pragma unroll
  for (int I = 0; I < NT; ++I) {
#pragma unroll
    for (int r = 0; r < 4; ++r) T[I][JC][r] = pb[16 * I + 4 * g + r][c16];
  }
  // no interchanges: the tables keep the last searched panel with this index (cleared when the
  // next one is built) and apply_swaps skips the panel
  if (fast) return 0;
  // interchange tables: slots 0..15 = rows j0..j0+15, then each distinct pivot row below
  const int pv = lane < 16 ? sh.piv[par][lane] : -1;
  bool first = lane < 16 && pv >= j0 + 16;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    const int o = __builtin_amdgcn_readlane(pv, k);
    if (k < lane && o == pv) first = false;
  }
  const unsigned long long fm = __ballot(first);
  const int prev = sh.dcount[par];
  if (lane < prev) sh.slot[par][sh.dlist[par][lane]] = -1;  // entries of the last panel built here
  wave_sync();
  if (lane < 16) {
    sh.slot[par][j0 + lane] = (signed char)lane;
    sh.dlist[par][lane] = j0 + lane;
  }
  if (first) {
    const int s = 16 + __popcll(fm & ((1ull << lane) - 1ull));
    sh.slot[par][pv] = (signed char)s;
    sh.dlist[par][s] = pv;
  }
  const int cnt = 16 + __popcll(fm);
  wave_sync();
  int src_row = 0;
  if (lane < cnt) {
    int pos = sh.dlist[par][lane];
#pragma unroll
    for (int jj = 15; jj >= 0; --jj) {
      const int pj = __builtin_amdgcn_readlane(pv, jj);
      if (pos == j0 + jj) pos = pj;
      else if (pos == pj) pos = j0 + jj;
    }
    src_row = pos;
    sh.srcslot[par][lane] = sh.slot[par][pos];
  }
  if (lane == 0) sh.dcount[par] = cnt;
  // the permutation after the panel's interchanges: position d takes the entry of its source
  const int moved = lane < cnt ? sh.perm[src_row] : 0;
  wave_sync();
  if (lane < cnt) sh.perm[sh.dlist[par][lane]] = moved;
  return info;
}

// ---- row interchanges of panel p on a wavefront's own tile columns --------------------------
template <int NM>
__device__ __forceinline__ void apply_swaps(f4 (&T)[Cfg<NM>::NT][Cfg<NM>::NTC], int p, int NTn, Shm<NM>& sh, int par, int skip_jc, int tid) {
  constexpr int NT = Cfg<NM>::NT, W = Cfg<NM>::W, NTC = Cfg<NM>::NTC;
  if (sh.noswap[par]) return;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c16 = lane & 15;
  float(*stg)[16 * NTC + 1] = sh.u.f.stg[wave];
  bool col_ok[NTC];
#pragma unroll
  for (int jc = 0; jc < NTC; ++jc) col_ok[jc] = (jc * W + wave < NTn) && (jc != skip_jc);
  // the slots of this lane's four rows in tile row I (one byte each; -1: not affected)
  auto slots = [&](int I) -> int {
    return (I >= p && I < NTn) ? *reinterpret_cast<const int*>(&sh.slot[par][16 * I + 4 * g]) : -1;
  };
  // the affected rows' current values to their slots
#pragma unroll
  for (int I = 0; I < NT; ++I) {
    const int s4 = slots(I);
    if (!__ballot(s4 != -1)) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int sl = (signed char)(s4 >> (8 * r));
      if (sl >= 0) {
#pragma unroll
        for (int jc = 0; jc < NTC; ++jc)
          if (col_ok[jc]) stg[sl][16 * jc + c16] = T[I][jc][r];
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int I = 0; I < NT; ++I) {
    const int s4 = slots(I);
    if (!__ballot(s4 != -1)) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int sl = (signed char)(s4 >> (8 * r));
      if (sl >= 0) {
        const int src = sh.srcslot[par][sl];
#pragma unroll
        for (int jc = 0; jc < NTC; ++jc)
          if (col_ok[jc]) T[I][jc][r] = stg[src][16 * jc + c16];
      }
    }
  }
  wave_sync();
}

// ---- U12 and the trailing update of one tile column (JC) after panel p ----------------------
template <int NM, int JC>
__device__ __forceinline__ void update_column(f4 (&T)[Cfg<NM>::NT][Cfg<NM>::NTC], int p, int NTn, Shm<NM>& sh, int par, int tid) {
  constexpr int NT = Cfg<NM>::NT, W = Cfg<NM>::W;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c16 = lane & 15;
  const int J = JC * W + wave;
  if (J <= p || J >= NTn) return;
  // A12 = tile (p, J); U12 = L11^-1 A12
  f4 a12 = T[0][JC];
#pragma unroll
  for (int I = 1; I < NT; ++I)
    if (I == p) a12 = T[I][JC];
  const float4v li = *reinterpret_cast<const float4v*>(&sh.u.f.linv[par][c16][4 * g]);
  f4 u = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) u = mma32(li[kk], a12[kk], u);
#pragma unroll
  for (int I = 0; I < NT; ++I)
    if (I == p) T[I][JC] = u;
  // A22 -= L21 U12
  const float(*pb)[20] = sh.u.f.pbuf[par];
#pragma unroll
  for (int I = 0; I < NT; ++I) {
    if (I > p && I < NTn) {
      const float4v la = *reinterpret_cast<const float4v*>(&pb[16 * I + c16][4 * g]);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) T[I][JC] = mma32(-la[kk], u[kk], T[I][JC]);
    }
    // at most four tile rows' operands in flight: bounds the registers the schedule holds
    if ((I & 3) == 3) __builtin_amdgcn_sched_barrier(0);
  }
}

// Every own tile column except `skip` (the one already updated).
template <int NM, int JC = 0>
__device__ __forceinline__ void update_rest(f4 (&T)[Cfg<NM>::NT][Cfg<NM>::NTC], int p, int NTn, Shm<NM>& sh, int par,
                                            int skip, int tid) {
  if constexpr (JC < Cfg<NM>::NTC) {
    if (JC != skip) update_column<NM, JC>(T, p, NTn, sh, par, tid);
    update_rest<NM, JC + 1>(T, p, NTn, sh, par, skip, tid);
  }
}

// The owner of panel p + 1 (local column jn): that column first, the panel, then the rest.
template <int NM, int JC = 0>
__device__ __forceinline__ int lookahead(f4 (&T)[Cfg<NM>::NT][Cfg<NM>::NTC], int p, int NTn, Shm<NM>& sh, int par,
                                         int jn, int tid) {
  if constexpr (JC < Cfg<NM>::NTC) {
    if (JC == jn) {
      update_column<NM, JC>(T, p, NTn, sh, par, tid);
      const int inf = factor_panel<NM, JC>(T, p + 1, NTn, sh, (p + 1) % Cfg<NM>::NBUF, opaque_tid(tid));
      if (!Cfg<NM>::DEFER) update_rest<NM>(T, p, NTn, sh, par, JC, opaque_tid(tid));
      return inf;
    }
    return lookahead<NM, JC + 1>(T, p, NTn, sh, par, jn, tid);
  } else {
    return 0;
  }
}

// ---- triangular solves on the row-block layout ----------------------------------------------
// T[J][ir]: tile (I = ir W + wave, J) as an MFMA A operand (lane l: row l % 16, columns
// 4(l/16) .. 4(l/16)+3).  Dinv[ir][0] = L_II^-1, Dinv[ir][1] = U_II^-1 (same layout).
// cbuf holds the right-hand sides in the permuted order; on return, the solution.
template <int NM>
__device__ __forceinline__ void solve_blocks(const f4 (&T)[Cfg<NM>::NT][Cfg<NM>::NTC], const f4 (&Dinv)[Cfg<NM>::NTC][2], int NTn, Shm<NM>& sh,
                             int tid) {
  constexpr int NT = Cfg<NM>::NT, W = Cfg<NM>::W, NTC = Cfg<NM>::NTC;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c16 = lane & 15;
  float(*cb)[16] = sh.u.s.cbuf;
  // forward: L y = c (unit lower)
  f4 acc[NTC];
#pragma unroll
  for (int ir = 0; ir < NTC; ++ir) acc[ir] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < NT; ++s) {
    if (s < NTn) {
      const int irs = s / W;
      if (wave == s % W) {
        f4 t;
#pragma unroll
        for (int r = 0; r < 4; ++r) t[r] = cb[16 * s + 4 * g + r][c16] - acc[irs][r];
        f4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) y = mma32(Dinv[irs][0][kk], t[kk], y);
#pragma unroll
        for (int r = 0; r < 4; ++r) cb[16 * s + 4 * g + r][c16] = y[r];
      }
      __syncthreads();
      f4 yv;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) yv[kk] = cb[16 * s + 4 * g + kk][c16];
#pragma unroll
      for (int ir = 0; ir < NTC; ++ir) {
        const int I = ir * W + wave;
        if (I > s && I < NTn) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) acc[ir] = mma32(T[s][ir][kk], yv[kk], acc[ir]);
        }
      }
    }
  }
  // backward: U x = y
#pragma unroll
  for (int ir = 0; ir < NTC; ++ir) acc[ir] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = NT - 1; s >= 0; --s) {
    if (s < NTn) {
      const int irs = s / W;
      if (wave == s % W) {
        f4 t;
#pragma unroll
        for (int r = 0; r < 4; ++r) t[r] = cb[16 * s + 4 * g + r][c16] - acc[irs][r];
        f4 x = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) x = mma32(Dinv[irs][1][kk], t[kk], x);
#pragma unroll
        for (int r = 0; r < 4; ++r) cb[16 * s + 4 * g + r][c16] = x[r];
      }
      __syncthreads();
      f4 xv;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) xv[kk] = cb[16 * s + 4 * g + kk][c16];
#pragma unroll
      for (int ir = 0; ir < NTC; ++ir) {
        const int I = ir * W + wave;
        if (I < s) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) acc[ir] = mma32(T[s][ir][kk], xv[kk], acc[ir]);
        }
      }
    }
  }
  __syncthreads();
}

// ---- FP64 residual r = b - A x on the FP64 matrix cores --------------------------------------
// Wavefront w takes tile rows I = w, w + W; per 16-column chunk one A-operand load of 4
// doubles per lane (row l % 16, columns 4(l/16) .. +3) and four MFMAs against x (B operand:
// x64[16c + 4(l/16) + kk][l % 16], zero beyond nrhs).  Writes (float) P r into cbuf, the
// column maxima of |r| and |x| into cmax[0..2 RMAX), and with `norm` ||A||_inf into
// cmax[2 RMAX].
template <int NM>
__device__ __forceinline__ void residual(const double* __restrict__ a, const double* __restrict__ b, int n, int nrhs, int NTn,
                         bool norm, Shm<NM>& sh, int tid) {
  constexpr int W = Cfg<NM>::W, NTC = Cfg<NM>::NTC;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c16 = lane & 15;
#pragma unroll 1
  for (int it = 0; it < NTC; ++it) {
    const int I = it * W + wave;
    if (I >= NTn) continue;
    const int row = 16 * I + c16;  // this lane's A-operand row
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    double asum = 0.0;
    const double* arow = a + (long)row * n;
    // four 16-column chunks per round: 16 doubles per lane in flight
#pragma unroll 1
    for (int c0 = 0; c0 < NTn; c0 += 4) {
      double av[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int col0 = 16 * (c0 + u) + 4 * g;
        if (row < n && col0 + 3 < n) {
          // four consecutive doubles of the row in one 32-byte load (8-byte aligned)
          const dvec4u v = *reinterpret_cast<const dvec4u*>(arow + col0);
          av[u][0] = v[0]; av[u][1] = v[1]; av[u][2] = v[2]; av[u][3] = v[3];
        } else {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) av[u][kk] = (row < n && col0 + kk < n) ? arow[col0 + kk] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (c0 + u < NTn) {
          const int col0 = 16 * (c0 + u) + 4 * g;
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const double xv = (c16 < nrhs) ? sh.u.s.x64[col0 + kk][c16 < RMAX ? c16 : 0] : 0.0;
            acc = mma64(av[u][kk], xv, acc);
            asum += fabs(av[u][kk]);
          }
        }
      }
    }
    // acc register i of lane l: row 16 I + 4 i + l/16, right-hand side l % 16
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rr = 16 * I + 4 * i + g;
      if (rr < n) {
        float out = 0.0f;
        if (c16 < nrhs) {
          const double rv = b[(long)rr * nrhs + c16] - acc[i];
          out = (float)rv;
          atomicMax(&sh.cmax[c16], (unsigned long long)__double_as_longlong(fabs(rv)));
          atomicMax(&sh.cmax[RMAX + c16], (unsigned long long)__double_as_longlong(fabs(sh.u.s.x64[rr][c16])));
        }
        sh.u.s.cbuf[sh.pinv[rr]][c16] = out;
      }
    }
    if (norm) {
      // row sums of |a|: the four lanes of a row (l % 16 equal) hold its column quarters
      asum += __shfl_xor(asum, 16, 64);
      asum += __shfl_xor(asum, 32, 64);
      if (g == 0 && row < n) atomicMax(&sh.cmax[2 * RMAX], (unsigned long long)__double_as_longlong(asum));
    }
  }
}

// ---- the kernel ---------------------------------------------------------------------------
// list[0 .. *list_count): the systems of this order class (order <= NM); systems in range take
// the path described at the top; results: X (FP64), iter[s] (steps, or -2 zero pivot in FP32,
// -3 no convergence), info[s] = 0, and failures appended to fallback_list.
template <int NM>
__global__ __launch_bounds__(Cfg<NM>::THREADS, Cfg<NM>::WAVES_PER_EU) void dsgesv_onchip_kernel(
    const int32_t* __restrict__ list, const int* __restrict__ list_count, const int32_t* __restrict__ nv,
    const int32_t* __restrict__ rv, const int64_t* __restrict__ a_off, const int64_t* __restrict__ b_off,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ X,
    int32_t* __restrict__ iter, int32_t* __restrict__ info, int* counter, int32_t* __restrict__ fallback_list,
    int* fallback_count, double tol, int max_iter) {
  constexpr int NT = Cfg<NM>::NT, W = Cfg<NM>::W, NTC = Cfg<NM>::NTC;
  __shared__ Shm<NM> sh;
  const int tid0 = threadIdx.x;
  const int count = *list_count;
  if (tid0 < Cfg<NM>::NBUF) sh.dcount[tid0] = 0;
  for (int i = tid0; i < Cfg<NM>::NBUF * NM; i += Cfg<NM>::THREADS) (&sh.slot[0][0])[i] = -1;
  for (;;) {
    const int tid = opaque_tid(tid0), lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    __syncthreads();
    if (tid == 0) sh.sys = atomicAdd(counter, 1);
    __syncthreads();
    const int idx = sh.sys;
    if (idx >= count) break;
    const int s = list[idx];
    const int n = nv[s], nrhs = rv[s];
    const int NTn = (n + 15) / 16;
    const double* a = A + a_off[s];
    const double* b = B + b_off[s];
    for (int i = tid; i < NM; i += Cfg<NM>::THREADS) sh.perm[i] = i;
    if (tid == 0) {
      sh.info = 0x7fffffff;  // smallest zero-pivot column over the panels (none: max)
      sh.spec = 1;
    }
    // ---- A (FP64, HBM) -> FP32 tiles; identity beyond n ----
    // Four tiles at a time: lane l reads row l / 4 of each, columns 4 (l % 4) .. + 3, in one
    // 32-byte load (the scalar form at the matrix's right edge), converts to FP32 and writes the
    // row-major tile to its wavefront's LDS area, from which the accumulator layout is read.
    f4 T[NT][NTC];
    {
      float(*ldt)[16][20] = sh.u.ld[wave];
      const int lr = lane >> 2, lc = 4 * (lane & 3);
#pragma unroll
      for (int jc = 0; jc < NTC; ++jc) {
        const int J = jc * W + wave;
#pragma unroll
        for (int I0 = 0; I0 < NT; I0 += 4) {
          double v[4][4];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int row = 16 * (I0 + t) + lr, col = 16 * J + lc;
            if (row < n && col + 3 < n) {
              const dvec4u x = *reinterpret_cast<const dvec4u*>(a + (long)row * n + col);
              v[t][0] = x[0]; v[t][1] = x[1]; v[t][2] = x[2]; v[t][3] = x[3];
            } else {
#pragma unroll
              for (int kk = 0; kk < 4; ++kk)
                v[t][kk] = (row < n && col + kk < n) ? a[(long)row * n + col + kk] : (row == col + kk ? 1.0 : 0.0);
            }
          }
#pragma unroll
          for (int t = 0; t < 4; ++t)
            *reinterpret_cast<float4v*>(&ldt[t][lr][lc]) =
                float4v{(float)v[t][0], (float)v[t][1], (float)v[t][2], (float)v[t][3]};
          wave_sync();
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int I = I0 + t;
            if (I < NTn && J < NTn) {
#pragma unroll
              for (int r = 0; r < 4; ++r) T[I][jc][r] = ldt[t][4 * g + r][c16];
            } else {
              T[I][jc] = f4{0.f, 0.f, 0.f, 0.f};
            }
          }
          wave_sync();
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    __syncthreads();
    // ---- FP32 LU: panel 0 by its owner, then one barrier per panel ----
    int my_info = 0;
    if (wave == 0) my_info = factor_panel<NM, 0>(T, 0, NTn, sh, 0, tid);
    for (int p = 0; p < NTn; ++p) {
      const int par = p % Cfg<NM>::NBUF;
      const int tp = opaque_tid(tid), wp = tp >> 6;
      __syncthreads();
      // the owner of panel p factored it during step p - 1 and left its other tile columns'
      // update with panel p - 1 for now (before panel p's interchanges, as in the other columns)
      if (Cfg<NM>::DEFER && p >= 1 && wp == p % W)
        update_rest<NM>(T, p - 1, NTn, sh, (p - 1) % Cfg<NM>::NBUF, p / W, opaque_tid(tp));
      apply_swaps<NM>(T, p, NTn, sh, par, (wp == p % W) ? p / W : -1, tp);
      if (p + 1 < NTn && wp == (p + 1) % W) {
        const int inf = lookahead<NM>(T, p, NTn, sh, par, (p + 1) / W, opaque_tid(tp));
        if (my_info == 0) my_info = inf;
      } else {
        update_rest<NM>(T, p, NTn, sh, par, -1, opaque_tid(tp));
      }
    }
    // first zero pivot over the panels (each owner reports its first)
    if (lane == 0 && my_info) atomicMin(&sh.info, my_info);
    __syncthreads();
    const int info32 = sh.info == 0x7fffffff ? 0 : sh.info;
    int it_out = -2;
    if (info32 == 0) {
      // per-lane quantities of the solve phase, derived afresh (not held across the LU)
      const int tid = opaque_tid(tid0), lane = tid & 63, wave = tid >> 6;
      const int g = lane >> 4, c16 = lane & 15;
      // ---- redistribute to row-block ownership (A-operand layout), in place ----
      // After this, T[J][ir] holds tile (ir W + wave, J) as an MFMA A operand (lane l: row l % 16,
      // columns 4(l/16) .. +3).  Round (jc, h) sends the tiles of this wavefront's column jc in
      // tile rows [h W, h W + W) and receives its row block h's tiles in tile columns
      // [jc W, jc W + W) into T[jc W + v][h]: slots that round (h, jc) sends, so (0, 1) and
      // (1, 0) run as one round and the receive never overwrites a tile not yet sent.
      static_assert(NTC == 2, "in-place redistribution is written for two tile columns per wavefront");
      auto send = [&](int jc, int h, int k) {
        const int J = jc * W + wave;
#pragma unroll
        for (int v = 0; v < W; ++v) {
          const int I = h * W + v;
          if (J < NTn && I < NTn) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = 4 * g + r;
              sh.u.rbuf[k][v][wave][row * 16 + (c16 ^ ((row & 3) << 2))] = (jc == 0 ? T[I][0][r] : T[I][1][r]);
            }
          }
        }
      };
      auto receive = [&](int jc, int h, int k) {
        const int Iw = h * W + wave;
#pragma unroll
        for (int v = 0; v < W; ++v) {
          const int Jr = jc * W + v;
          if (Iw < NTn && Jr < NTn) {
            const float4v t =
                *reinterpret_cast<const float4v*>(&sh.u.rbuf[k][wave][v][c16 * 16 + ((4 * g) ^ ((c16 & 3) << 2))]);
            if (h == 0) T[Jr][0] = f4{t[0], t[1], t[2], t[3]};
            else T[Jr][1] = f4{t[0], t[1], t[2], t[3]};
          }
        }
      };
      send(0, 0, 0);
      __syncthreads();
      receive(0, 0, 0);
      __syncthreads();
      send(0, 1, 0);
      send(1, 0, 1);
      __syncthreads();
      receive(0, 1, 0);
      receive(1, 0, 1);
      __syncthreads();
      send(1, 1, 0);
      __syncthreads();
      receive(1, 1, 0);
      __syncthreads();
      // ---- inverses of the diagonal blocks (L unit lower, U upper) of the own row blocks ----
      // Both row blocks at once: lanes 0..15 / 16..31 form column c of L^-1 / U^-1 of the first
      // block, lanes 32..47 / 48..63 those of the second; then back in the A-operand layout.
      f4 Dinv[NTC][2];
      {
        float(*ds)[16][20] = sh.u.s.dscr[wave];
        float(*di)[16][33] = sh.u.s.dinv[wave];
#pragma unroll
        for (int ir = 0; ir < NTC; ++ir) {
          const int I = ir * W + wave;
          f4 d = T[0][ir];
#pragma unroll
          for (int J = 1; J < NT; ++J)
            if (J == I) d = T[J][ir];
          if (I < NTn) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) ds[ir][c16][4 * g + kk] = d[kk];
          }
        }
        wave_sync();
        {
          const int ir = lane >> 5, h = lane & 31;
          if (ir * W + wave < NTn) {
            float x[16];
            if (h < 16) unit_lower_inverse_column(&ds[ir][0][0], 20, h, x);
            else upper_inverse_column(&ds[ir][0][0], 20, h - 16, x);
#pragma unroll
            for (int i = 0; i < 16; ++i) di[ir][i][h] = x[i];
          }
        }
        wave_sync();
#pragma unroll
        for (int ir = 0; ir < NTC; ++ir) {
          f4 li = {0.f, 0.f, 0.f, 0.f}, ui = {0.f, 0.f, 0.f, 0.f};
          if (ir * W + wave < NTn) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
              li[kk] = di[ir][c16][4 * g + kk];
              ui[kk] = di[ir][c16][16 + 4 * g + kk];
            }
          }
          Dinv[ir][0] = li;
          Dinv[ir][1] = ui;
        }
        wave_sync();
      }
      // ---- x = FP32 solve of P b ----
      for (int i = tid; i < NM; i += Cfg<NM>::THREADS) sh.pinv[sh.perm[i]] = i;
      // c = P b: RMAX right-hand sides per row, in rounds of THREADS / 2 rows
      for (int e = tid; e < 2 * NM; e += Cfg<NM>::THREADS) {
        const int i = e >> 1, c0 = (e & 1) * RMAX;
        const int src = sh.perm[i];
        float v[RMAX];
#pragma unroll
        for (int c = 0; c < RMAX; ++c)
          v[c] = (c0 == 0 && c < nrhs && src < n) ? (float)b[(long)src * nrhs + c] : 0.0f;
#pragma unroll
        for (int c = 0; c < RMAX; ++c) sh.u.s.cbuf[i][c0 + c] = v[c];
      }
      __syncthreads();
      solve_blocks<NM>(T, Dinv, NTn, sh, opaque_tid(tid));
      for (int e = tid; e < NM * RMAX; e += Cfg<NM>::THREADS) {
        const int i = e / RMAX, c = e % RMAX;
        sh.u.s.x64[i][c] = (double)sh.u.s.cbuf[i][c];
      }
      const double eps = tol > 0 ? tol : 0x1.0p-53;
      double anrm = 0.0, prev = 1e308;
      int stalls = 0;
      it_out = -3;
      for (int it = 0; it <= max_iter; ++it) {
        const int ti = opaque_tid(tid);
        if (ti < 2 * RMAX + 1) sh.cmax[ti] = 0ull;
        __syncthreads();
        residual<NM>(a, b, n, nrhs, NTn, it == 0, sh, ti);
        __syncthreads();
        if (it == 0) anrm = __longlong_as_double((long long)sh.cmax[2 * RMAX]);
        const double cte = anrm * eps * sqrt((double)n);
        bool ok = true;
        double worst = 0.0;
        for (int c = 0; c < nrhs; ++c) {
          const double rn = __longlong_as_double((long long)sh.cmax[c]);
          const double xn = __longlong_as_double((long long)sh.cmax[RMAX + c]);
          if (!(rn <= xn * cte)) ok = false;
          worst = fmax(worst, rn);
        }
        if (ok) { it_out = it; break; }
        if (it == max_iter) break;
        if (!(worst < 0.5 * prev)) {  // the residual must halve at least every other step
          if (++stalls >= 2) break;
        } else {
          stalls = 0;
        }
        prev = worst;
        solve_blocks<NM>(T, Dinv, NTn, sh, opaque_tid(ti));
        for (int e = ti; e < NM * RMAX; e += Cfg<NM>::THREADS) {
          const int i = e / RMAX, c = e % RMAX;
          sh.u.s.x64[i][c] += (double)sh.u.s.cbuf[i][c];
        }
      }
      if (it_out >= 0) {
        double* x = X + b_off[s];
        for (int e = tid; e < n * nrhs; e += Cfg<NM>::THREADS) x[e] = sh.u.s.x64[e / nrhs][e % nrhs];
      }
    }
    if (tid == 0) {
      iter[s] = it_out;
      info[s] = 0;
      if (it_out < 0) fallback_list[atomicAdd(fallback_count, 1)] = s;
    }
  }
}

// Stable partition of the batch (in `order`, or index order) by order class: class 0..3 =
// orders 1..32, 33..64, 65..128, 129..256; class 4 = 257..512.  Systems outside 1..512 or with
// nrhs outside 0..RMAX get info = -1, iter = 0 and no class.  One workgroup of 1024 threads:
// thread t takes a contiguous chunk, counts per class, a block scan gives the offsets.
// lists + cls * cap; counts[cls] = class sizes.
__global__ __launch_bounds__(1024) void partition_kernel(int count, const int32_t* __restrict__ nv,
                                                         const int32_t* __restrict__ rv,
                                                         const int32_t* __restrict__ order,
                                                         int32_t* __restrict__ info, int32_t* __restrict__ iter,
                                                         int32_t* __restrict__ lists, int cap,
                                                         int* __restrict__ counts) {
  constexpr int NC = 5;
  __shared__ int scan[NC][1024];
  const int t = threadIdx.x;
  const int chunk = (count + 1023) / 1024;
  const int i0 = min(count, t * chunk), i1 = min(count, i0 + chunk);
  auto cls_of = [&](int s) -> int {
    const int n = nv[s], r = rv[s];
    if (n < 1 || n > 512 || r < 0 || r > RMAX) return -1;
    return n <= 32 ? 0 : n <= 64 ? 1 : n <= 128 ? 2 : n <= 256 ? 3 : 4;
  };
  int cnt[NC] = {0, 0, 0, 0, 0};
  for (int i = i0; i < i1; ++i) {
    const int s = order ? order[i] : i;
    const int c = cls_of(s);
    if (c >= 0) cnt[c]++;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) scan[c][t] = cnt[c];
  __syncthreads();
  // inclusive scan per class (Hillis-Steele over 1024 entries)
  for (int off = 1; off < 1024; off <<= 1) {
    int v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = t >= off ? scan[c][t - off] : 0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c) scan[c][t] += v[c];
    __syncthreads();
  }
  int pos[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) pos[c] = scan[c][t] - cnt[c];
  for (int i = i0; i < i1; ++i) {
    const int s = order ? order[i] : i;
    const int c = cls_of(s);
    if (c < 0) {
      if (info) info[s] = -1;
      if (iter) iter[s] = 0;
    } else {
      lists[c * cap + pos[c]++] = s;
    }
  }
  if (t == 1023) {
#pragma unroll
    for (int c = 0; c < NC; ++c) counts[c] = scan[c][1023];
  }
}

}  // namespace mx

// Launches the on-chip mixed solver for the order classes 1..256 (lists written by
// partition_kernel; counters[0..3] their work counters, cleared by the caller).
bzg_status launch_dsgesv_onchip(bzg_context* ctx, const int32_t* lists, int cap, const int* counts,
                                const int32_t* n, const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                                const double* A, const double* B, double* X, int32_t* iter, int32_t* info,
                                int* counters, int32_t* fb_list, int* fb_count, double tol, int max_iter,
                                const int* wg_per_cu, int64_t count) {
  const int64_t cap_grid = std::max<int64_t>(count, 1);
#define BZG_MX(NMV, CLS)                                                                                     \
  hipLaunchKernelGGL((mx::dsgesv_onchip_kernel<NMV>),                                                         \
                     dim3((unsigned)std::min<int64_t>((int64_t)ctx->num_cus * wg_per_cu[CLS], cap_grid)),     \
                     dim3(mx::Cfg<NMV>::THREADS), 0, ctx->stream, lists + (size_t)(CLS) * cap, counts + (CLS), n, \
                     nrhs, a_off, b_off, A, B, X, iter, info, counters + (CLS), fb_list, fb_count, tol, max_iter); \
  BZG_LAUNCH_CHECK()
  BZG_MX(256, 3);
  BZG_MX(128, 2);
  BZG_MX(64, 1);
  BZG_MX(32, 0);
#undef BZG_MX
  return BZG_OK;
}

bzg_status launch_partition(bzg_context* ctx, int64_t count, const int32_t* n, const int32_t* nrhs,
                            const int32_t* order, int32_t* info, int32_t* iter, int32_t* lists, int cap,
                            int* counts) {
  hipLaunchKernelGGL(mx::partition_kernel, dim3(1), dim3(1024), 0, ctx->stream, (int)count, n, nrhs, order, info,
                     iter, lists, cap, counts);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}

}  // namespace bzg
