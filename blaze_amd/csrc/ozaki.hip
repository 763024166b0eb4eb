// FP64-accurate GEMM from low-precision residue slices on MFMA (Ozaki scheme II).
//
// No reference function: this is a drop-in for the FP64 C = A * B assignment
// (blaze/math/expressions/DMatDMatMultExpr.h:516-569) that computes the product through
// exact integer arithmetic on the BF16 (or INT8) matrix cores.
//
// Method (Ozaki, Uchino & Imamura, "Ozaki Scheme II: a GEMM-oriented emulation of floating-
// point matrix multiplication using an integer modular technique", 2025; here with BF16
// operands and FP32 accumulation):
//   1. scale every row i of A by 2^s_i and every column j of B by 2^t_j so that the rounded
//      integers A' = rint(A_i * 2^s_i), B' = rint(B_j * 2^t_j) satisfy |A'|, |B'| <= 2^P;
//   2. for each of N pairwise-coprime odd moduli m_l <= 511, form the symmetric residues
//      A'_l = A' mod m_l, B'_l = B' mod m_l (|r| <= 255: exact in BF16);
//   3. C'_l = A'_l B'_l mod m_l on MFMA v_mfma_f32_32x32x16_bf16.  Every product is exact in
//      FP32 (|r r'| <= 65025) and every 256-long K chunk sums to at most 256*65025 + 767 < 2^24,
//      so the FP32 accumulator is an exact integer; it is reduced mod m_l after each chunk;
//   4. CRT: C' = sum_l C'_l W_l mod M, W_l = (M/m_l) * ((M/m_l)^-1 mod m_l), M = prod m_l.  W_l
//      and M are split completely into 40-bit (39 for BF16) chunks so every partial sum is exact in FP64;
//      |C'| <= k 2^(2P) <= M/4 makes the symmetric reduction unique;
//   5. C = C' * 2^-(s_i + t_j), rounded once to FP64.
// The INT8 variant uses moduli <= 255 (|r| <= 127) on v_mfma_i32_32x32x32_i8 with exact INT32
// accumulation (no chunking for k <= 2^17).
// Error: |C - AB| <= 2^-P * (row max |A_i|)(col max |B_j|) * k  +  one FP64 rounding.
#include "common.hpp"

#include <mutex>

#include <cmath>
#include <type_traits>
#include <cstring>
#include <vector>

namespace bzg {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kMaxModuli = 24;

constexpr int kMaxChunks = 6;
// Residue planes (one per modulus) are spaced by their size plus 19 x 256 bytes: planes of a
// power-of-two size would put element (i, j) of every plane in the same DRAM channel and bank,
// and the kernels touch all planes at once.
constexpr long kPlanePadBytes = 19 * 256;
__host__ __device__ constexpr long plane_elems(long elems, int esz) { return elems + kPlanePadBytes / esz; }  // chunks of W_l and M: 24 moduli x 9 bits <= 216 bits, >= 39 bits each

struct CrtConsts {
  double m[kMaxModuli];
  double minv[kMaxModuli];
  double c30[kMaxModuli];  // 2^30 mod m
  double w[kMaxChunks][kMaxModuli];  // W_l = sum_c w[c][l] exactly (chunk c: cb bits at E-cb(c+1))
  double M[kMaxChunks];              // M = sum_c M[c] exactly
  double Minv;
  int n, nch;
};

// ------------------------------------------------------------------------------------------
// host: moduli and CRT constants (small big-integer arithmetic, 32-bit limbs)
// ------------------------------------------------------------------------------------------
struct Big {
  std::vector<uint32_t> d;  // little-endian limbs
  explicit Big(uint64_t v = 0) {
    d.push_back((uint32_t)v);
    d.push_back((uint32_t)(v >> 32));
    trim();
  }
  void trim() {
    while (d.size() > 1 && d.back() == 0) d.pop_back();
  }
  void mul_small(uint32_t s) {
    uint64_t c = 0;
    for (auto& x : d) {
      uint64_t v = (uint64_t)x * s + c;
      x = (uint32_t)v;
      c = v >> 32;
    }
    if (c) d.push_back((uint32_t)c);
  }
  uint32_t mod_small(uint32_t m) const {
    uint64_t r = 0;
    for (size_t i = d.size(); i-- > 0;) r = ((r << 32) | d[i]) % m;
    return (uint32_t)r;
  }
  int bits() const {
    int b = (int)(d.size() - 1) * 32;
    uint32_t t = d.back();
    while (t) {
      ++b;
      t >>= 1;
    }
    return b;
  }
  // floor(this / 2^sh) mod 2^width as a double (width <= 52)
  double chunk(int sh, int width) const {
    uint64_t v = 0;
    for (int b = width - 1; b >= 0; --b) {
      const int pos = sh + b;
      uint64_t bit = 0;
      if (pos >= 0) {
        const size_t limb = (size_t)pos / 32;
        if (limb < d.size()) bit = (d[limb] >> (pos % 32)) & 1u;
      }
      v = (v << 1) | bit;
    }
    return (double)v;
  }
};

static uint32_t modinv(uint32_t a, uint32_t m) {
  int64_t t = 0, nt = 1, r = m, nr = a % m;
  while (nr) {
    const int64_t q = r / nr;
    int64_t tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  if (t < 0) t += m;
  return (uint32_t)t;
}

static uint32_t gcd32(uint32_t a, uint32_t b) {
  while (b) { const uint32_t t = a % b; a = b; b = t; }
  return a;
}

// Pairwise-coprime odd moduli, largest first, not above `top`.
static std::vector<uint32_t> pick_moduli(int count, uint32_t top) {
  std::vector<uint32_t> ms;
  for (uint32_t m = top; m >= 3 && (int)ms.size() < count; m -= 2) {
    bool ok = true;
    for (uint32_t x : ms)
      if (gcd32(m, x) != 1) { ok = false; break; }
    if (ok) ms.push_back(m);
  }
  return ms;
}

static double log2_prod(const std::vector<uint32_t>& ms) {
  double s = 0;
  for (auto m : ms) s += std::log2((double)m);
  return s;
}

// Bits per operand after scaling: k * 2^(2P) <= M / 4.
static int operand_bits(const std::vector<uint32_t>& ms, int64_t k) {
  const double lm = log2_prod(ms) - 1e-9;
  const double lk = std::log2((double)std::max<int64_t>(k, 1));
  return (int)std::floor((lm - 2.0 - lk) / 2.0);
}

// Chunk width cb: a partial sum S_c = sum_l r_l w[c][l] with |r_l| <= rmax and up to 24 terms,
// and S_c - q M[c] with |q| <= 24 rmax, must stay below 2^53 units of the chunk: cb = 40 for
// INT8 residues (|r| <= 127), 39 for BF16 ones (|r| <= 255).
static void build_crt(const std::vector<uint32_t>& ms, CrtConsts& c, int cb) {
  const int n = (int)ms.size();
  Big M(1);
  for (auto m : ms) M.mul_small(m);
  const int E = M.bits();
  c.n = n;
  c.nch = (E + cb - 1) / cb;  // every bit of W_l < M: the partial sums below are exact
  for (int l = 0; l < n; ++l) {
    Big Ml(1);
    for (int q = 0; q < n; ++q)
      if (q != l) Ml.mul_small(ms[q]);
    const uint32_t inv = modinv(Ml.mod_small(ms[l]), ms[l]);
    Big W = Ml;
    W.mul_small(inv);  // W < M * m_l; reduce mod M below by using W mod M == W - M*floor
    // W = Ml * inv with inv < m_l, so W < M: no reduction needed.
    c.m[l] = (double)ms[l];
    c.minv[l] = 1.0 / (double)ms[l];
    c.c30[l] = (double)((1ull << 30) % ms[l]);
    for (int ch = 0; ch < c.nch; ++ch) c.w[ch][l] = std::ldexp(W.chunk(E - cb * (ch + 1), cb), E - cb * (ch + 1));
  }
  for (int ch = 0; ch < c.nch; ++ch) c.M[ch] = std::ldexp(M.chunk(E - cb * (ch + 1), cb), E - cb * (ch + 1));
  c.Minv = 1.0 / (c.M[0] + c.M[1]);
}

// ------------------------------------------------------------------------------------------
// device: scaling exponents
// ------------------------------------------------------------------------------------------
// |x| with NaN mapped to +Inf: the maxima of the fixed-count path are +Inf exactly for the rows
// (columns) that hold a NaN or an infinity, which nonfinite_fixup_kernel recomputes.
__device__ __forceinline__ double nonfinite_as_inf(double a) { return a != a ? __builtin_inf() : a; }

// Row maxima of |A| (m x k, lda), one wavefront per row; +Inf for a row with a NaN or an infinity.
__global__ __launch_bounds__(256) void row_max_kernel(int m, int k, const double* __restrict__ A, long lda,
                                                      double* __restrict__ rmax) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x % 64;
  if (wave >= m) return;
  const double* row = A + (long)wave * lda;
  double mx = 0.0;
  for (int j = lane; j < k; j += 64) mx = fmax(mx, nonfinite_as_inf(fabs(row[j])));
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) rmax[wave] = mx;
}

// Scaling exponents: s_i = P - 1 - ilogb(max_j |a_ij|) (0 for a zero row), so that
// |rint(a_ij 2^s_i)| <= 2^P; t_j likewise from the column maxima of B (FP64 bit patterns).
// A non-finite maximum (a row with a NaN or an infinity) also gets 0: its residues are never used.
__device__ __forceinline__ int row_exp(double mx, int P) {
  return (mx > 0.0 && mx <= 0x1.fffffffffffffp+1023) ? (P - 1 - ilogb(mx)) : 0;
}
__device__ __forceinline__ int col_exp(unsigned long long bits, int P) {
  return row_exp(__longlong_as_double((long long)bits), P);
}

// Column maxima of |B| (k x n, ldb) as FP64 bit patterns (non-negative doubles order like
// their bit patterns): a (n / 64) x (k / 256) grid of 64-column x 256-row blocks (y strides over
// K), combined with one atomicMax per column and block.  cmax must be zero on entry.  A column
// with a NaN or an infinity gets +Inf.

// C(i, j) = sum_q A(i, q) B(q, j) in FP64, in q order, for every entry in a row of A or a column
// of B whose maximum is +Inf (a NaN or an infinity in it): such entries are NaN or +-Inf, and
// IEEE arithmetic gives the same NaN / infinity in any summation order, so these entries match
// the plain FP64 GEMM.  Grid x: m + n workgroups; workgroup g < m owns row g, g >= m column g - m;
// a clean row or column returns at once.
__global__ __launch_bounds__(256) void nonfinite_fixup_kernel(int m, int n, int k, const double* __restrict__ A,
                                                              long lda, const double* __restrict__ B, long ldb,
                                                              const double* __restrict__ rmax,
                                                              const unsigned long long* __restrict__ cmax,
                                                              double* __restrict__ C, long ldc) {
  const int g = blockIdx.x;
  const bool is_row = g < m;
  const double mx = is_row ? rmax[g] : __longlong_as_double((long long)cmax[g - m]);
  if (!(mx > 0x1.fffffffffffffp+1023)) return;
  const int len = is_row ? n : m;
  for (int e = threadIdx.x; e < len; e += blockDim.x) {
    const int i = is_row ? g : e, j = is_row ? e : g - m;
    const double* a = A + (long)i * lda;
    double acc = 0.0;
    for (int q = 0; q < k; ++q) acc = fma(a[q], B[(long)q * ldb + j], acc);
    C[(long)i * ldc + j] = acc;
  }
}
__global__ __launch_bounds__(256) void col_max_kernel(int k, int n, const double* __restrict__ B,
                                                      long ldb, unsigned long long* __restrict__ cmax) {
  __shared__ double part[4][64];
  const int c = blockIdx.x * 64 + threadIdx.x % 64, ph = threadIdx.x / 64;
  double mx = 0.0;
  if (c < n)
    for (int r0 = blockIdx.y * 256; r0 < k; r0 += gridDim.y * 256) {
      const int r1 = min(k, r0 + 256);
      for (int i = r0 + ph; i < r1; i += 4) mx = fmax(mx, nonfinite_as_inf(fabs(B[(long)i * ldb + c])));
    }
  part[ph][threadIdx.x % 64] = mx;
  __syncthreads();
  if (ph == 0 && c < n) {
    mx = fmax(fmax(part[0][threadIdx.x], part[1][threadIdx.x]), fmax(part[2][threadIdx.x], part[3][threadIdx.x]));
    atomicMax(&cmax[c], (unsigned long long)__double_as_longlong(mx));
  }
}

// One pass over a row-major r x c matrix X (ld ldx) for the automatic moduli count: a
// workgroup covers 64 rows x 1024 columns (each wavefront 16 rows, lane l columns l + 64u).
// ROWMAX (A):  rowpart[x][i] = max |x_ij| over column range x, colpart[y][j] = sum of x_ij^2 over
//              the row blocks of workgroup row y;
// !ROWMAX (B): rowpart[x][i] = sum of x_ij^2 over column range x, colpart[y][j] = max |x_ij|.
// Every sum has a fixed order (rows in sequence per lane, a fixed butterfly per row, the four
// wavefronts in order); reduce_partials_kernel combines the partials.
template <bool ROWMAX>
__global__ __launch_bounds__(256) void stats_pass_kernel(int r, int c, const double* __restrict__ X, long ldx,
                                                         double* __restrict__ rowpart,
                                                         double* __restrict__ colpart) {
  __shared__ double comb[4][1024];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 1024;
  double acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = 0.0;
  for (int rb = blockIdx.y * 64 + wave * 16; rb < r; rb += gridDim.y * 64) {
    const int re = min(r, rb + 16);
    for (int i = rb; i < re; ++i) {
      const double* row = X + (long)i * ldx;
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = c0 + lane + 64 * u;
        v[u] = (j < c) ? row[j] : 0.0;
      }
      double red = 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        if (ROWMAX) {
          acc[u] = fma(v[u], v[u], acc[u]);
          red = fmax(red, fabs(v[u]));
        } else {
          acc[u] = fmax(acc[u], fabs(v[u]));
          red = fma(v[u], v[u], red);
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(red, off, 64);
        red = ROWMAX ? fmax(red, o) : red + o;
      }
      if (lane == 0) rowpart[(long)blockIdx.x * r + i] = red;
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) comb[wave][lane + 64 * u] = acc[u];
  __syncthreads();
  for (int jj = threadIdx.x; jj < 1024; jj += 256) {
    const int j = c0 + jj;
    if (j < c) {
      const double a = comb[0][jj], b = comb[1][jj], d = comb[2][jj], e = comb[3][jj];
      colpart[(long)blockIdx.y * c + j] = ROWMAX ? (a + b) + (d + e) : fmax(fmax(a, b), fmax(d, e));
    }
  }
}

// out[j] = sum (MAX: max) over y < ny of part[y * n + j], in order.
template <bool MAX>
__global__ __launch_bounds__(256) void reduce_partials_kernel(int ny, int n, const double* __restrict__ part,
                                                              double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double a = 0.0;
  int y = 0;
  for (; y + 8 <= ny; y += 8) {  // eight loads in flight, combined in order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = part[(long)(y + u) * n + j];
#pragma unroll
    for (int u = 0; u < 8; ++u) a = MAX ? fmax(a, v[u]) : a + v[u];
  }
  for (; y < ny; ++y) a = MAX ? fmax(a, part[(long)y * n + j]) : a + part[(long)y * n + j];
  out[j] = a;
}

// Accuracy statistics (one workgroup, fixed summation order): out[0] = sum_i D_i^2,
// out[1] = ||A||_F^2, out[2] = sum_j E_j^2, out[3] = ||B||_F^2, out[4] = sum_q ||A_:q||^2 ||B_q:||^2,
// where D_i = 2^(ilogb(rmax_i)+1) and E_j likewise for the column maxima of B; acn2 holds the
// column sums of squares of A, brn2 the row sums of squares of B (both of length k).
__global__ __launch_bounds__(1024) void fp64acc_stats_kernel(int m, int n, int k,
                                                            const double* __restrict__ rmax,
                                                            const unsigned long long* __restrict__ cmax,
                                                            const double* __restrict__ acn2,
                                                            const double* __restrict__ brn2, double* __restrict__ out) {
  __shared__ double red[5][1024];
  const int t = threadIdx.x;
  double sa = 0, fa = 0, sb = 0, fb = 0, cc = 0;
  for (int q = t; q < k; q += 1024) {
    cc = fma(acn2[q], brn2[q], cc);
    fa += acn2[q];
    fb += brn2[q];
  }
  for (int i = t; i < m; i += 1024) {
    const double mx = rmax[i];
    if (mx > 0.0) {
      const double d = ldexp(1.0, ilogb(mx) + 1);
      sa = fma(d, d, sa);
    }
  }
  for (int j = t; j < n; j += 1024) {
    const double mx = __longlong_as_double((long long)cmax[j]);
    if (mx > 0.0) {
      const double e = ldexp(1.0, ilogb(mx) + 1);
      sb = fma(e, e, sb);
    }
  }
  red[0][t] = sa; red[1][t] = fa; red[2][t] = sb; red[3][t] = fb; red[4][t] = cc;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int q = 0; q < 5; ++q) red[q][t] += red[q][t + w];
    __syncthreads();
  }
  if (t < 5) out[t] = red[t][0];
}

__global__ __launch_bounds__(256) void zero_u64_kernel(int n, unsigned long long* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Symmetric residue of the integer-valued double v modulo m (|result| <= (m-1)/2).
__device__ __forceinline__ double sym_mod(double v, double m, double minv) {
  double r = fma(-m, rint(v * minv), v);  // exact: the true value is a small integer
  const double h = 0.5 * (m - 1.0);
  r = (r > h) ? r - m : r;
  r = (r < -h) ? r + m : r;
  return r;
}

// Same for |v| up to 2^106: split v = vh 2^30 + vl exactly, reduce vh, then fold in vl.
// (sym_mod alone needs |v / m| < 2^50 for the quotient to be within one of the truth.)
__device__ __forceinline__ double sym_mod_wide(double v, double m, double minv, double c30) {
  const double vh = floor(v * 0x1.0p-30);
  const double vl = fma(-vh, 0x1.0p30, v);  // exact, 0 <= vl < 2^30
  const double rh = sym_mod(vh, m, minv);
  return sym_mod(fma(rh, c30, vl), m, minv);
}

template <int MODE>
struct SliceOf;
template <>
struct SliceOf<0> {  // BF16: 4 slices packed in 8 bytes
  using T = __bf16;
  using Pack = uint2;
  __device__ static uint32_t bits(double r) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)(float)r);
  }
  __device__ static Pack pack(const double* r) {
    return make_uint2(bits(r[0]) | (bits(r[1]) << 16), bits(r[2]) | (bits(r[3]) << 16));
  }
};
template <>
struct SliceOf<1> {  // INT8: 4 slices packed in 4 bytes
  using T = int8_t;
  using Pack = uint32_t;
  __device__ static uint32_t bits(double r) { return (uint32_t)(uint8_t)(int8_t)(int)r; }
  __device__ static Pack pack(const double* r) {
    return bits(r[0]) | (bits(r[1]) << 8) | (bits(r[2]) << 16) | (bits(r[3]) << 24);
  }
};

// RM = 0: |v| <= 2^50.  Then fl(v * fl(1/m)) is within |v/m| 2^-52 < 1/(2m) of v/m, and v/m (m
// odd) is at least 1/(2m) away from every half-integer, so rint returns the nearest quotient
// and v - m q already lies in [-(m-1)/2, (m-1)/2].  RM = 1: |v| <= 2^56, quotient within one
// (sym_mod's corrections).  RM = 2: wider (sym_mod_wide).
template <int RM>
__device__ __forceinline__ double residue(double v, const CrtConsts* __restrict__ cc, int l) {
  if constexpr (RM == 0) return fma(-cc->m[l], rint(v * cc->minv[l]), v);
  else if constexpr (RM == 1) return sym_mod(v, cc->m[l], cc->minv[l]);
  else return sym_mod_wide(v, cc->m[l], cc->minv[l], cc->c30[l]);
}

// 16 residues (one modulus) packed for a 16-byte (INT8) or two 16-byte (BF16) stores.
template <int MODE>
__device__ __forceinline__ void store16(typename SliceOf<MODE>::T* __restrict__ o, const double* r) {
  using S = SliceOf<MODE>;
  if constexpr (MODE == 1) {
    *reinterpret_cast<uint4*>(o) = make_uint4(S::pack(r), S::pack(r + 4), S::pack(r + 8), S::pack(r + 12));
  } else {
    const uint2 a = S::pack(r), b = S::pack(r + 4), c = S::pack(r + 8), d = S::pack(r + 12);
    reinterpret_cast<uint4*>(o)[0] = make_uint4(a.x, a.y, b.x, b.y);
    reinterpret_cast<uint4*>(o)[1] = make_uint4(c.x, c.y, d.x, d.y);
  }
}

// A residues: Ares[l][i][kk] (rows padded to mp, k padded to kp (a multiple of 64), zeros in
// the padding).  16 consecutive kk per thread: 16-byte loads of A where the row is aligned,
// one 16-byte (INT8) or two (BF16) stores per modulus, so a wavefront writes 1 KB (2 KB)
// contiguous per plane.  Grid: x covers kp in 4096-element chunks, y strides over rows.
template <int MODE, int RM>
__global__ __launch_bounds__(256) void residues_a_kernel(int m, int k, const double* __restrict__ A,
                                                         long lda, const double* __restrict__ rmax, int P,
                                                         int* __restrict__ s_out, int mp, int kp, int nmod,
                                                         const CrtConsts* __restrict__ cc,
                                                         typename SliceOf<MODE>::T* __restrict__ out) {
  const int k0 = (blockIdx.x * 256 + threadIdx.x) * 16;
  if (k0 >= kp) return;
  const long plane = plane_elems((long)mp * kp, sizeof(typename SliceOf<MODE>::T));
  for (int i = blockIdx.y; i < mp; i += gridDim.y) {
    double v[16];
    const double* src = A + (long)i * lda + k0;
    if (i < m && k0 + 16 <= k && ((uintptr_t)src & 15) == 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 t2 = reinterpret_cast<const double2*>(src)[u];
        v[2 * u] = t2.x;
        v[2 * u + 1] = t2.y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = (i < m && k0 + u < k) ? src[u] : 0.0;
    }
    const int si = i < m ? row_exp(rmax[i], P) : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0 && i < m) s_out[i] = si;  // for the CRT
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = rint(ldexp(v[u], si));
    typename SliceOf<MODE>::T* o = out + (long)i * kp + k0;
    for (int l = 0; l < nmod; ++l) {
      double r[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) r[u] = residue<RM>(v[u], cc, l);
      store16<MODE>(o + l * plane, r);
    }
  }
}

// B residues, transposed: Bres[l][j][kk] (n rows padded to np, k padded to kp).  A workgroup
// transposes a 128 (kk) x 32 (j) tile through LDS: 16-byte reads along j (256 bytes per row), then 16
// consecutive kk per thread and one 16-byte (INT8) or two (BF16) stores per modulus (128 bytes
// contiguous per j row).  LDS column c of row r is stored at c ^ (((r >> 4) & 3) << 3), which
// spreads both access patterns over the banks.  Grid x runs over the j blocks, so neighbouring
// workgroups read neighbouring pieces of the same rows of B (measured 1.25x faster than kk
// blocks first at 8192^2).
template <int MODE, int RM>
__global__ __launch_bounds__(256) void residues_b_kernel(int k, int n, const double* __restrict__ B,
                                                         long ldb, const unsigned long long* __restrict__ cmax,
                                                         int P, int* __restrict__ t_out, int np, int kp, int nmod,
                                                         const CrtConsts* __restrict__ cc,
                                                         typename SliceOf<MODE>::T* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) double tile[128][32];
  const int kb = blockIdx.y * 128, jb = blockIdx.x * 32;
  const int tid = threadIdx.x;
  {
    // 16-byte loads: lane pairs of columns (2 jp, 2 jp + 1), 16 rows per pass
    const int jp = tid & 15, j = jb + 2 * jp;
    const bool vec = ((((uintptr_t)B) | ((uintptr_t)ldb * 8)) & 15) == 0 && j + 1 < n;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = (tid >> 4) + 16 * i, kk = kb + r;
      double2 v = make_double2(0.0, 0.0);
      if (kk < k) {
        const double* src = B + (long)kk * ldb + j;
        if (vec) {
          v = *reinterpret_cast<const double2*>(src);
        } else {
          if (j < n) v.x = src[0];
          if (j + 1 < n) v.y = src[1];
        }
      }
      *reinterpret_cast<double2*>(&tile[r][(2 * jp) ^ (((r >> 4) & 3) << 3)]) = v;
    }
  }
  __syncthreads();
  const int jr = tid >> 3, g = tid & 7;
  const int j = jb + jr, kk0 = kb + 16 * g;
  if (j >= np || kk0 >= kp) return;
  const int tj = (j < n) ? col_exp(cmax[j], P) : 0;
  if (blockIdx.y == 0 && g == 0 && j < n) t_out[j] = tj;  // for the CRT
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = rint(ldexp(tile[16 * g + u][jr ^ ((g & 3) << 3)], tj));
  const long plane = plane_elems((long)np * kp, sizeof(typename SliceOf<MODE>::T));
  typename SliceOf<MODE>::T* o = out + (long)j * kp + kk0;
  for (int l = 0; l < nmod; ++l) {
    double r[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) r[u] = residue<RM>(v[u], cc, l);
    store16<MODE>(o + l * plane, r);
  }
}

// ------------------------------------------------------------------------------------------
// device: residue GEMM  C_l = A_l * B_l^T mod m_l   (256 x 256 tiles, 4 wavefronts)
// ------------------------------------------------------------------------------------------
constexpr int RT = 256;        // tile rows / cols
constexpr int RBYTES = 128;    // bytes of K per LDS stage and row
constexpr int RSTAGE = RT * RBYTES;  // bytes per operand per stage

// 16-byte chunk c of row r lives at chunk c ^ ((r >> 1) & 7).  The LDS serves a ds_read_b128
// 16 lanes at a time over 64 four-byte banks (256 bytes): 16 consecutive rows read at one
// logical chunk then land on 16 distinct 16-byte bank groups (row parity picks the half,
// (r >> 1) & 7 the group), and a row-major store of 2 rows x 8 chunks is conflict-free too.
__device__ __forceinline__ int swz(int r, int c) { return r * RBYTES + ((c ^ ((r >> 1) & 7)) << 4); }

// FN = 32-column MFMA blocks per wavefront: 4 -> 256 x 256 workgroup tile (INT8), 2 -> 256 x 128
// (BF16: its FP32 accumulators are reduced on the VALU every 256 K, which needs them in VGPRs).
// Output residues are symmetric: int8 for INT8 slices (m <= 255 -> |r| <= 127), int16 for BF16.
//
// Pipeline per K stage (128 bytes of K, 4 MFMA k-steps): the fragments of k-step ks+1 are read
// from LDS into a second register set while the 4 x FN MFMAs of k-step ks run; every k-step
// also writes a quarter of the staged registers (stage st+1, loaded one stage earlier) into the
// other LDS buffer (last read before the previous barrier) and issues the global loads of stage
// st+2 into the registers it has just written out, so each load has a whole stage to land
// (8192^3, 14 moduli: 6.60 ms, against 6.76 ms with all stores and loads on the last k-step and
// 7.05 ms with the loads issued at the top of the stage for the next one); one barrier closes
// the stage.
//
// MODE 2 (INT32 GEMM from balanced base-256 digits): Ares / Bres are single arrays whose rows
// hold the 4 digit planes side by side along K (A: d0 d1 d2 d3, B^T: d3 d2 d1 d0, kp each), and
// z = 3 - c runs digit class c (the pairs i + j = c) as one product over (c + 1) kp: A's segments
// 0..c against B's segments 3-c..3.  The class sums are written as raw INT32 planes.
template <int MODE>
using ResOut = typename std::conditional<MODE == 0, int16_t,
                                         typename std::conditional<MODE == 1, int8_t, int32_t>::type>::type;

template <int MODE, int FN>
__global__ __launch_bounds__(256, 1) void residue_gemm_kernel(int mp, int np, int kp,
                                                              const char* __restrict__ Ares,
                                                              const char* __restrict__ Bres,
                                                              ResOut<MODE>* __restrict__ Cres,
                                                              const CrtConsts* __restrict__ cc) {
  constexpr int ESZ = (MODE == 0) ? 2 : 1;          // bytes per element
  constexpr int KSTEP_B = 32;                        // bytes of K per MFMA (16 bf16 / 32 i8)
  constexpr int NKS = RBYTES / KSTEP_B;              // MFMA k-steps per stage
  constexpr int BN = 64 * FN;                        // workgroup tile columns
  constexpr int NBCH = BN * 8 / 256;                 // B chunks staged per thread
  extern __shared__ __attribute__((aligned(16))) char smem[];
  char* As = smem;                   // [2][RT rows][RBYTES]
  char* Bs = smem + 2 * RSTAGE;      // [2][BN rows][RBYTES]

  const int l = (MODE == 2) ? 3 - (int)blockIdx.z : (int)blockIdx.z;  // MODE 2: the longest class first
  const long a_plane = plane_elems((long)mp * kp * ESZ, 1), b_plane = plane_elems((long)np * kp * ESZ, 1);
  const char* Ag = (MODE == 2) ? Ares : Ares + l * a_plane;
  const char* Bg = (MODE == 2) ? Bres + (long)(3 - l) * kp : Bres + l * b_plane;
  float mf = 0.0f, minvf = 0.0f;
  if constexpr (MODE == 0) {
    mf = (float)cc->m[l];
    minvf = 1.0f / mf;
  }

  const int tiles_n = np / BN;
  const int tiles_m = mp / RT;
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if (nt >= 64 && nt % 8 == 0) bid = (bid % 8) * (nt / 8) + bid / 8;  // XCD-contiguous
  // walk tiles in column-major groups of 8 tile rows for L2 reuse (8192^3, 14 INT8 residue
  // products: 7.94 ms per call against 8.04 with groups of 4, 8.36 with 2, 8.10 with 16)
  const int G = 8;
  const int group = bid / (G * tiles_n);
  const int first_m = group * G;
  const int gsize = min(tiles_m - first_m, G);
  const int in_group = bid % (G * tiles_n);
  const int tm = first_m + in_group % gsize, tn = in_group / gsize;
  const long row0 = (long)tm * RT, col0 = (long)tn * BN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * 128, wn0 = (wave & 1) * (32 * FN);
  const long kbytes = (long)kp * ESZ;
  const long rowb = (MODE == 2) ? 4 * kbytes : kbytes;  // bytes per operand row

  using acc_t = typename std::conditional<MODE == 0, f16v, v16i>::type;
  acc_t acc[4][FN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

  // each thread stages 8 chunks of A and NBCH of B per stage: chunk q = tid + 256*e
  v4i ra[8], rb[NBCH];
  auto gload = [&](int st) {
    const long kb = (long)st * RBYTES;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int q = tid + 256 * e, r = q >> 3, c = q & 7;
      ra[e] = *reinterpret_cast<const v4i*>(Ag + (row0 + r) * rowb + kb + c * 16);
    }
#pragma unroll
    for (int e = 0; e < NBCH; ++e) {
      const int q = tid + 256 * e, r = q >> 3, c = q & 7;
      rb[e] = *reinterpret_cast<const v4i*>(Bg + (col0 + r) * rowb + kb + c * 16);
    }
  };
  // the stores and loads of the flattened staging chunks e in [lo, hi) (A: 0..7, B: 8..)
  auto sstore_part = [&](int buf, int lo, int hi) {
#pragma unroll
    for (int e = 0; e < 8 + NBCH; ++e) {
      if (e < lo || e >= hi) continue;
      if (e < 8) {
        const int q = tid + 256 * e, r = q >> 3, c = q & 7;
        *reinterpret_cast<v4i*>(As + buf * RSTAGE + swz(r, c)) = ra[e];
      } else {
        const int q = tid + 256 * (e - 8), r = q >> 3, c = q & 7;
        *reinterpret_cast<v4i*>(Bs + buf * RSTAGE + swz(r, c)) = rb[e - 8];
      }
    }
  };
  auto gload_part = [&](int st, int lo, int hi) {
    const long kb = (long)st * RBYTES;
#pragma unroll
    for (int e = 0; e < 8 + NBCH; ++e) {
      if (e < lo || e >= hi) continue;
      if (e < 8) {
        const int q = tid + 256 * e, r = q >> 3, c = q & 7;
        ra[e] = *reinterpret_cast<const v4i*>(Ag + (row0 + r) * rowb + kb + c * 16);
      } else {
        const int q = tid + 256 * (e - 8), r = q >> 3, c = q & 7;
        rb[e - 8] = *reinterpret_cast<const v4i*>(Bg + (col0 + r) * rowb + kb + c * 16);
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int q = tid + 256 * e, r = q >> 3, c = q & 7;
      *reinterpret_cast<v4i*>(As + buf * RSTAGE + swz(r, c)) = ra[e];
    }
#pragma unroll
    for (int e = 0; e < NBCH; ++e) {
      const int q = tid + 256 * e, r = q >> 3, c = q & 7;
      *reinterpret_cast<v4i*>(Bs + buf * RSTAGE + swz(r, c)) = rb[e];
    }
  };
  // MFMA operand fragments of k-step ks of LDS buffer buf
  v4i fa[2][4], fb[2][FN];
  auto lfrag = [&](int buf, int ks, v4i* a, v4i* b) {
    const char* as = As + buf * RSTAGE;
    const char* bs = Bs + buf * RSTAGE;
    const int c = ks * 2 + (lane >> 5);
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const v4i*>(as + swz(wm0 + i * 32 + (lane & 31), c));
#pragma unroll
    for (int j = 0; j < FN; ++j) b[j] = *reinterpret_cast<const v4i*>(bs + swz(wn0 + j * 32 + (lane & 31), c));
  };

  auto kstep = [&](int cur) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        if constexpr (MODE == 0) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, fa[cur][i]),
                                                               __builtin_bit_cast(bf8, fb[cur][j]),
                                                               acc[i][j], 0, 0, 0);
        } else {
          acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[cur][i], fb[cur][j], acc[i][j], 0, 0, 0);
        }
      }
  };
  auto reduce_acc = [&]() {
    if constexpr (MODE == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const float v = acc[i][j][r];
            acc[i][j][r] = __builtin_fmaf(-mf, __builtin_rintf(v * minvf), v);
          }
    }
  };

  const int nst = (int)((MODE == 2 ? (l + 1) * kbytes : kbytes) / RBYTES);
  // K elements per stage: 64 (bf16) / 128 (i8); bf16 reduces every 4 stages (256 K)
  // The global loads of stage st + 2 are issued right after the stores of stage st + 1 (on the
  // last k-step of stage st, the staging registers are free again), so they have a whole
  // stage and a barrier to land; near the end the index is clamped to the last stage (a
  // redundant, in-bounds reload keeps the loop free of branches).
  gload(0);
  sstore(0);
  gload(nst > 1 ? 1 : 0);
  __syncthreads();
  lfrag(0, 0, fa[0], fb[0]);
  for (int st = 0; st + 1 < nst; ++st) {
    const int buf = st & 1;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int cur = ks & 1;
      if (ks + 1 < NKS) lfrag(buf, ks + 1, fa[cur ^ 1], fb[cur ^ 1]);
      kstep(cur);
      constexpr int QS = (8 + NBCH) / NKS;
      sstore_part(buf ^ 1, ks * QS, (ks + 1) * QS);
      gload_part(min(st + 2, nst - 1), ks * QS, (ks + 1) * QS);
    }
    // instruction order of the stage: the first 4 + FN MFMAs of k-steps 0..NKS-2 are each
    // followed by one LDS fragment read of the following k-step; the next (8 + NBCH) / NKS
    // MFMAs of every k-step by one LDS store of the next stage and one global load of the
    // stage after it (the load reuses the register the store just read).
    constexpr int NM = 4 * FN;
    if constexpr (MODE != 0)  // (the BF16 tile measured faster with the compiler's own order)
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
      for (int t = 0; t < NM; ++t) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);               // MFMA
        if (ks + 1 < NKS && t < 4 + FN) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        if (t >= 4 + FN && t < 4 + FN + (8 + NBCH) / NKS) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
        if (t >= 4 + FN && t < 4 + FN + (8 + NBCH) / NKS) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
      }
    if ((st & 3) == 3) reduce_acc();
    __syncthreads();
    lfrag(buf ^ 1, 0, fa[0], fb[0]);
  }
#pragma unroll
  for (int ks = 0; ks < NKS; ++ks) {
    const int cur = ks & 1;
    if (ks + 1 < NKS) lfrag((nst - 1) & 1, ks + 1, fa[cur ^ 1], fb[cur ^ 1]);
    kstep(cur);
  }

  if constexpr (MODE == 2) {
    // class sums as they are (|sum| <= 4 k 2^14 < 2^31 for k < 2^15)
    int32_t* Cg = Cres + l * plane_elems((long)mp * np, sizeof(int32_t));
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const long gr = row0 + wm0 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const long gc = col0 + wn0 + j * 32 + (lane & 31);
          Cg[gr * np + gc] = acc[i][j][r];
        }
    return;
  } else {
  // epilogue: symmetric residue mod m
  const long cplane = plane_elems((long)mp * np, sizeof(ResOut<MODE>));
  ResOut<MODE>* Cg = Cres + l * cplane;
  const int mi = (int)cc->m[l];
  const int half = (mi - 1) / 2;
  const double mid = cc->m[l], minvd = cc->minv[l];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        int v;
        if constexpr (MODE == 0) {
          const float f = acc[i][j][r];
          v = (int)__builtin_fmaf(-mf, __builtin_rintf(f * minvf), f);
        } else {
          // |acc| < 2^31: the FP64 quotient is within 2^-20 of v / m, whose fractional part
          // is 0 or >= 1/m, so floor() is exact except at multiples of m (result m -> 0 below)
          const double d = (double)acc[i][j][r];
          v = (int)fma(-mid, floor(d * minvd), d);
        }
        v = (v > half) ? v - mi : v;
        v = (v < -half) ? v + mi : v;
        const long gr = row0 + wm0 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long gc = col0 + wn0 + j * 32 + (lane & 31);
        Cg[gr * np + gc] = (ResOut<MODE>)v;
      }
  }
}

// ------------------------------------------------------------------------------------------
// device: CRT recombination and scaling
// ------------------------------------------------------------------------------------------
// Columns per thread of crt_kernel: 8 keeps it at 111 VGPRs (4 wavefronts per SIMD); 16 needed
// 179 and measured 1.16x slower at 8192^2.
constexpr int kCrtCols = 8;
// CPT consecutive columns per thread (np is a multiple of 256): per residue plane one load of
// CPT residues (16 bytes for CPT = 16 int8); NCH = number of CRT chunks.  A workgroup covers
// 256 * CPT columns of a row; the results go through LDS (one padding double per 16) so the
// stores to C are 16 bytes per lane and contiguous across the wavefront.  Grid x covers
// np / (256 CPT), y strides over rows.
__device__ __forceinline__ int crt_pos(int g) { return g + (g >> 4); }

template <typename R, int NCH>
__global__ __launch_bounds__(256) void crt_kernel(int m, int n, int mp, int np, int nmod,
                                                  const R* __restrict__ Cres,
                                                  const int* __restrict__ s, const int* __restrict__ t,
                                                  const CrtConsts* __restrict__ cc, double* __restrict__ C,
                                                  long ldc) {
  constexpr int CPT = kCrtCols;
  constexpr int SPAN = 256 * CPT;
  constexpr int WPL = CPT * (int)sizeof(R) / 4;  // 32-bit words per plane and thread
  constexpr int EPW = 4 / (int)sizeof(R);        // residues per word
  __shared__ double stage[SPAN + SPAN / 16];
  const int jb = blockIdx.x * SPAN;
  const int j0 = jb + threadIdx.x * CPT;
  const bool active = j0 < n;
  const long plane = plane_elems((long)mp * np, sizeof(R));
  int tj[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) tj[u] = (j0 + u < n) ? t[j0 + u] : 0;
  for (int i = blockIdx.y; i < m; i += gridDim.y) {
    double out[CPT];
    if (active) {
      const R* src = Cres + (long)i * np + j0;
      // S[c] = sum_l r_l w[c][l]: every term a multiple of chunk c's unit, below 2^53 units: exact
      double S[NCH][CPT];
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
        for (int u = 0; u < CPT; ++u) S[ch][u] = 0.0;
      // planes in groups of 8: a group's loads are issued before its first multiply-add
      constexpr int GRP = 8;
      for (int l0 = 0; l0 < nmod; l0 += GRP) {
        int raw[GRP][WPL];
#pragma unroll
        for (int g = 0; g < GRP; ++g)
          if (l0 + g < nmod) {
            const R* q = src + (l0 + g) * plane;
            if constexpr (WPL == 4) {
              const int4 v = *reinterpret_cast<const int4*>(q);
              raw[g][0] = v.x; raw[g][1] = v.y; raw[g][2] = v.z; raw[g][3] = v.w;
            } else if constexpr (WPL == 2) {
              const int2 v = *reinterpret_cast<const int2*>(q);
              raw[g][0] = v.x; raw[g][1] = v.y;
            } else {
#pragma unroll
              for (int h = 0; h < WPL / 4; ++h) {
                const int4 v = reinterpret_cast<const int4*>(q)[h];
                raw[g][4 * h] = v.x; raw[g][4 * h + 1] = v.y; raw[g][4 * h + 2] = v.z; raw[g][4 * h + 3] = v.w;
              }
            }
          }
#pragma unroll
        for (int g = 0; g < GRP; ++g) {
          if (l0 + g < nmod) {
            double r[CPT];
#pragma unroll
            for (int u = 0; u < CPT; ++u)
              r[u] = sizeof(R) == 1 ? (double)(int8_t)(raw[g][u / EPW] >> (8 * (u % EPW)))
                                    : (double)(int16_t)(raw[g][u / EPW] >> (16 * (u % EPW)));
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
              const double wl = cc->w[ch][l0 + g];
#pragma unroll
              for (int u = 0; u < CPT; ++u) S[ch][u] = fma(r[u], wl, S[ch][u]);
            }
          }
        }
      }
      const int si = s[i];
#pragma unroll
      for (int u = 0; u < CPT; ++u) {
        // X = sum_c S[c] = C' + q M with |C'| <= M/4: q from the leading chunk, then the chunks
        // of X - q M (each exact) summed with error-free transformations, rounded once
        const double q = rint(S[0][u] * cc->Minv);
        double hi = fma(-q, cc->M[0], S[0][u]);  // exact
        double lo = 0.0;
#pragma unroll
        for (int ch = 1; ch < NCH; ++ch) {
          const double tt = fma(-q, cc->M[ch], S[ch][u]);  // exact
          const double sum = hi + tt;                      // TwoSum(hi, tt)
          const double bb = sum - hi;
          lo += (hi - (sum - bb)) + (tt - bb);
          hi = sum;
        }
        out[u] = ldexp(hi + lo, -(si + tj[u]));
      }
    }
    __syncthreads();  // the previous row's stores have read the stage
    if (active)
#pragma unroll
      for (int u = 0; u < CPT; ++u) stage[crt_pos(threadIdx.x * CPT + u)] = out[u];
    __syncthreads();
    double* dst = C + (long)i * ldc + jb;
#pragma unroll
    for (int v = 0; v < CPT / 2; ++v) {
      const int g = 2 * threadIdx.x + 512 * v;
      const double a = stage[crt_pos(g)], b = stage[crt_pos(g + 1)];
      if (jb + g + 1 < n && ((uintptr_t)(dst + g) & 15) == 0) {
        *reinterpret_cast<double2*>(dst + g) = make_double2(a, b);
      } else {
        if (jb + g < n) dst[g] = a;
        if (jb + g + 1 < n) dst[g + 1] = b;
      }
    }
  }
}

struct Fp64AccPlan {
  int mode;  // 0 bf16, 1 int8
  int nmod;
  int P;
  double est;  // estimated relative Frobenius error (automatic count), else 0
  CrtConsts host;
};

// Automatic moduli count: the smallest count whose operand precision P brings the estimated
// relative Frobenius error below kAutoTarget.  Rounding A_i to P bits below 2^(ilogb(max)+1)
// leaves independent errors of variance (2^-P D_i)^2 / 12 per entry, so
//   ||dC||_F^2 ~ 2^-2P / 12 (sum_i D_i^2 ||B||_F^2 + sum_j E_j^2 ||A||_F^2),
// and for operands with independent signs ||C||_F^2 ~ sum_q ||A_:q||^2 ||B_q:||^2 = CC, giving
//   est(P) = 2^-P sqrt((sum D_i^2 ||B||_F^2 + sum E_j^2 ||A||_F^2) / (12 CC)).
// (Uniform [-1, 1) entries: est = 0.71 2^-P; for wide exponent spreads CC is far below
// ||A||^2 ||B||^2 / k, which is what makes them need more moduli.)  Sign-coherent operands make
// ||C|| larger and the estimate conservative; a product that cancels to far below sqrt(CC) is
// not covered by any a-priori choice.
// The estimate is an expectation over the rounding errors; it is tight when many terms of
// comparable size make up each error (uniform entries: measured/estimated = 1.00).  When a
// few entries dominate their rows or columns (wide exponent spreads), a few roundings, each
// as large as 1/2 against an rms of 1/sqrt(12), make up the error and the realised value can
// exceed the estimate by ~2x (measured 1.84x at K = 16384, sigma = 10).  Such operands are
// recognised by tau = k sum_i D_i^2 / ||A||_F^2 (3 for uniform entries: every row's norm is
// spread over its k entries; ~k when one entry carries its row) and get a target kPeakyFactor
// times smaller.
constexpr double kAutoTarget = 8e-15;
constexpr double kPeakyTau = 12.0;
constexpr double kPeakyFactor = 2.5;

double auto_target(int64_t k, const double* st) {
  const double ta = st[1] > 0.0 ? st[0] * (double)k / st[1] : 0.0;
  const double tb = st[3] > 0.0 ? st[2] * (double)k / st[3] : 0.0;
  return (ta > kPeakyTau || tb > kPeakyTau) ? kAutoTarget / kPeakyFactor : kAutoTarget;
}

double auto_estimate(int P, int64_t, const double* st) {
  if (!(st[4] > 0.0)) return 0.0;  // a zero operand or no overlapping support: C = 0 exactly
  const double q = (st[0] * st[3] + st[2] * st[1]) / (12.0 * st[4]);
  if (!(q < 1e300)) return 1.0;  // overflowing statistics (|entries| ~ 1e150): most moduli
  return std::ldexp(std::sqrt(q), -P);
}

bzg_status plan_fp64acc(int64_t k, int32_t moduli, int mode, Fp64AccPlan& pl, const double* stats = nullptr) {
  const uint32_t top = mode == 0 ? 511u : 255u;
  int nm = moduli;
  pl.est = 0.0;
  if (nm <= 0 && stats) {
    const double target = auto_target(k, stats);
    for (nm = 4; nm < kMaxModuli; ++nm) {
      const int P = std::min(operand_bits(pick_moduli(nm, top), k), 104);
      if (P >= 1 && auto_estimate(P, k, stats) <= target) break;
    }
    pl.est = auto_estimate(std::min(operand_bits(pick_moduli(nm, top), k), 104), k, stats);
  } else if (nm <= 0) {
    // without statistics: smallest count giving P >= 50 bits per operand
    for (nm = 4; nm <= kMaxModuli; ++nm)
      if (operand_bits(pick_moduli(nm, top), k) >= 50) break;
  }
  BZG_REQUIRE(nm >= 2 && nm <= kMaxModuli, BZG_ERR_INVALID_ARGUMENT, "fp64acc: moduli must be 2..24");
  auto ms = pick_moduli(nm, top);
  BZG_REQUIRE((int)ms.size() == nm, BZG_ERR_UNSUPPORTED, "fp64acc: not enough coprime moduli");
  pl.mode = mode;
  pl.nmod = nm;
  pl.P = operand_bits(ms, k);
  BZG_REQUIRE(pl.P >= 1, BZG_ERR_UNSUPPORTED, "fp64acc: too few moduli for this k");
  pl.P = std::min(pl.P, 104);  // |A'| <= 2^104 keeps the residue split exact
  BZG_REQUIRE(mode == 0 || k <= (1LL << 17), BZG_ERR_UNSUPPORTED, "fp64acc int8: k > 2^17");
  // the CRT constants depend only on the moduli: built once per (slice type, count)
  static std::mutex mu;
  static CrtConsts cache[2][kMaxModuli + 1];
  static bool have[2][kMaxModuli + 1] = {};
  std::lock_guard<std::mutex> lock(mu);
  if (!have[mode][nm]) {
    build_crt(ms, cache[mode][nm], mode == 1 ? 40 : 39);
    have[mode][nm] = true;
  }
  pl.host = cache[mode][nm];
  return BZG_OK;
}

// ------------------------------------------------------------------------------------------
// device: INT32 GEMM from INT8 digits (wrap-around, exact modulo 2^32)
// ------------------------------------------------------------------------------------------
// Balanced base-256 digits: x = d0 + 2^8 d1 + 2^16 d2 + 2^24 d3 (mod 2^32), d_i in [-128, 127].
// Each step subtracts the digit, which leaves a multiple of 256 and makes the shift exact; the
// last digit only counts modulo 2^8.  The product then is
//   A B = sum_{i+j <= 3} 2^{8(i+j)} A_i B_j (mod 2^32)
// (the pairs with i + j >= 4 vanish modulo 2^32): 10 INT8 products in 4 classes c = i + j.
__device__ __forceinline__ void digits4(int x, uint32_t& b0, uint32_t& b1, uint32_t& b2, uint32_t& b3) {
  long v = x;
  const long d0 = (long)(int8_t)(uint8_t)(v & 0xff);
  v = (v - d0) >> 8;
  const long d1 = (long)(int8_t)(uint8_t)(v & 0xff);
  v = (v - d1) >> 8;
  const long d2 = (long)(int8_t)(uint8_t)(v & 0xff);
  v = (v - d2) >> 8;
  b0 = (uint32_t)(uint8_t)d0;
  b1 = (uint32_t)(uint8_t)d1;
  b2 = (uint32_t)(uint8_t)d2;
  b3 = (uint32_t)(v & 0xff);
}

// 16 values -> 4 digit segments of 16 bytes at o + seg(q) kp (seg(q) = q, or 3 - q with REV).
template <bool REV>
__device__ __forceinline__ void store_digits16(const int* v, char* __restrict__ o, long kp) {
  uint32_t w[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t) w[q][t] = 0u;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    uint32_t b0, b1, b2, b3;
    digits4(v[u], b0, b1, b2, b3);
    const int sh = 8 * (u & 3), t = u >> 2;
    w[0][t] |= b0 << sh;
    w[1][t] |= b1 << sh;
    w[2][t] |= b2 << sh;
    w[3][t] |= b3 << sh;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    *reinterpret_cast<uint4*>(o + (long)(REV ? 3 - q : q) * kp) = make_uint4(w[q][0], w[q][1], w[q][2], w[q][3]);
}

// Digit rows from a source whose rows run along K: row i of out (rp rows of 4 kp bytes) from
// src[i * ld + kk] (rows x k; zeros in the padding).  16 consecutive kk per thread; grid x covers
// kp in 4096-element chunks, y strides over the rows.
template <bool REV>
__global__ __launch_bounds__(256) void digits_rows_kernel(int rows, int k, const int* __restrict__ src, long ld,
                                                          int rp, int kp, char* __restrict__ out) {
  const int k0 = (blockIdx.x * 256 + threadIdx.x) * 16;
  if (k0 >= kp) return;
  for (int i = blockIdx.y; i < rp; i += gridDim.y) {
    int v[16];
    const int* s = src + (long)i * ld + k0;
    if (i < rows && k0 + 16 <= k && ((uintptr_t)s & 15) == 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int4 t4 = reinterpret_cast<const int4*>(s)[u];
        v[4 * u] = t4.x;
        v[4 * u + 1] = t4.y;
        v[4 * u + 2] = t4.z;
        v[4 * u + 3] = t4.w;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = (i < rows && k0 + u < k) ? s[u] : 0;
    }
    store_digits16<REV>(v, out + (long)i * 4 * kp + k0, kp);
  }
}

// Digit rows from a source whose columns run along K: row j of out from src[kk * ld + j]
// (k x rows).  A workgroup transposes a 128 (kk) x 64 (j) tile through LDS (rows padded to 65
// words); grid x = rp / 64, y = kp / 128.
template <bool REV>
__global__ __launch_bounds__(256) void digits_cols_kernel(int rows, int k, const int* __restrict__ src, long ld,
                                                          int rp, int kp, char* __restrict__ out) {
  __shared__ int tile[128][65];
  const int kb = blockIdx.y * 128, jb = blockIdx.x * 64;
  const int tid = threadIdx.x, c = tid & 63;
#pragma unroll 8
  for (int p = 0; p < 32; ++p) {
    const int r = (tid >> 6) + 4 * p, kk = kb + r, j = jb + c;
    tile[r][c] = (kk < k && j < rows) ? src[(long)kk * ld + j] : 0;
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int t = tid + 256 * e, jr = t >> 3, g = t & 7;
    const int j = jb + jr, kk0 = kb + 16 * g;
    if (j >= rp || kk0 >= kp) continue;
    int v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = tile[16 * g + u][jr];
    store_digits16<REV>(v, out + (long)j * 4 * kp + kk0, kp);
  }
}

// C = alpha (P0 + 2^8 P1 + 2^16 P2 + 2^24 P3) + beta C (mod 2^32) from the class planes P_c
// (mp x np, cplane apart), Horner from the top class.  beta == 0 never reads C.  4 consecutive
// columns per thread; grid x covers n in 1024-column chunks, y strides over rows.
__global__ __launch_bounds__(256) void digits_combine_kernel(int m, int n, int np, long cplane, uint32_t alpha,
                                                             const int* __restrict__ P, uint32_t beta,
                                                             int* __restrict__ C, long ldc) {
  const int j0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (j0 >= n) return;
  for (int i = blockIdx.y; i < m; i += gridDim.y) {
    const long o = (long)i * np + j0;
    const int4 p0 = *reinterpret_cast<const int4*>(P + o);
    const int4 p1 = *reinterpret_cast<const int4*>(P + cplane + o);
    const int4 p2 = *reinterpret_cast<const int4*>(P + 2 * cplane + o);
    const int4 p3 = *reinterpret_cast<const int4*>(P + 3 * cplane + o);
    const int a0[4] = {p0.x, p0.y, p0.z, p0.w}, a1[4] = {p1.x, p1.y, p1.z, p1.w};
    const int a2[4] = {p2.x, p2.y, p2.z, p2.w}, a3[4] = {p3.x, p3.y, p3.z, p3.w};
    int* crow = C + (long)i * ldc + j0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (j0 + u >= n) break;
      uint32_t s = (uint32_t)a3[u];
      s = (s << 8) + (uint32_t)a2[u];
      s = (s << 8) + (uint32_t)a1[u];
      s = (s << 8) + (uint32_t)a0[u];
      uint32_t r = alpha * s;
      if (beta != 0u) r += beta * (uint32_t)crow[u];
      crow[u] = (int)r;
    }
  }
}

}  // namespace

bzg_status dgemm_fp64acc_impl(bzg_context* ctx, int64_t m, int64_t n, int64_t k, const double* A,
                              int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                              int32_t moduli, int mode) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "fp64acc: null context");
  BZG_REQUIRE(m >= 0 && n >= 0 && k >= 0, BZG_ERR_INVALID_ARGUMENT, "fp64acc: negative size");
  BZG_REQUIRE(m < (1LL << 31) && n < (1LL << 31) && k < (1LL << 31), BZG_ERR_UNSUPPORTED,
              "fp64acc: dimension >= 2^31");
  BZG_REQUIRE(lda >= (k > 0 ? k : 1) && ldb >= (n > 0 ? n : 1) && ldc >= (n > 0 ? n : 1),
              BZG_ERR_INVALID_ARGUMENT, "fp64acc: leading dimension smaller than row length");
  if (m == 0 || n == 0) return BZG_OK;
  if (k == 0) return bzg_dgemm(ctx, m, n, 0, 1.0, A, lda, B, ldb, 0.0, C, ldc);
  BZG_REQUIRE(A && B && C, BZG_ERR_INVALID_ARGUMENT, "fp64acc: null pointer");
  BZG_REQUIRE(moduli >= 0, BZG_ERR_INVALID_ARGUMENT, "fp64acc: moduli must be 0 (automatic) or 2..24");
  ProfScope prof_total(ctx, "fp64acc_total");
  // ---- operand statistics: row maxima of A and column maxima of B (the scaling exponents);
  // for the automatic count also the column sums of squares of A and the row sums of squares of
  // B, in one pass over each operand ----
  const long ny = std::min<long>((k + 255) / 256, 16384);
  const long gya = std::min<long>((m + 63) / 64, 65535), gyb = std::min<long>((k + 63) / 64, 65535);
  const long nxa = (k + 1023) / 1024, nxb = (n + 1023) / 1024;
  const size_t stat_bytes = 64 + (size_t)m * 8 + (size_t)n * 8 + (size_t)k * 16 +
                            (size_t)(nxa * m + gya * k + nxb * k + gyb * n) * 8;
  void* sws = nullptr;
  bzg_status st = workspace(ctx, 8, stat_bytes, &sws);
  if (st) return st;
  double* sout = static_cast<double*>(sws);
  double* rmax = sout + 8;
  unsigned long long* cmaxB = reinterpret_cast<unsigned long long*>(rmax + m);
  double* acol = reinterpret_cast<double*>(cmaxB + n);   // column sums of squares of A (k)
  double* brn2 = acol + k;       // row sums of squares of B (k)
  double* pa_row = brn2 + k;     // partials of the fused passes
  double* pa_col = pa_row + nxa * m;
  double* pb_row = pa_col + gya * k;
  double* pb_col = pb_row + nxb * k;
  Fp64AccPlan pl;
  if (moduli == 0) {
    // automatic count: the statistics come back to the host (one stream synchronisation)
    hipLaunchKernelGGL(stats_pass_kernel<true>, dim3((unsigned)nxa, (unsigned)gya), dim3(256), 0, ctx->stream,
                       (int)m, (int)k, A, (long)lda, pa_row, pa_col);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(stats_pass_kernel<false>, dim3((unsigned)nxb, (unsigned)gyb), dim3(256), 0, ctx->stream,
                       (int)k, (int)n, B, (long)ldb, pb_row, pb_col);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_partials_kernel<true>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int)nxa, (int)m, pa_row, rmax);
    hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int)gya, (int)k, pa_col, acol);
    hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int)nxb, (int)k, pb_row, brn2);
    // column maxima of B as FP64 bit patterns (non-negative doubles order like their bits)
    hipLaunchKernelGGL(reduce_partials_kernel<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int)gyb, (int)n, pb_col, reinterpret_cast<double*>(cmaxB));
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(fp64acc_stats_kernel, dim3(1), dim3(1024), 0, ctx->stream, (int)m, (int)n, (int)k, rmax,
                       cmaxB, acol, brn2, sout);
    BZG_LAUNCH_CHECK();
    if (!ctx->host_stats) BZG_HIP_CHECK(hipHostMalloc((void**)&ctx->host_stats, 8 * sizeof(double)));
    double* hs = ctx->host_stats;  // pinned: the readback is one small DMA
    BZG_HIP_CHECK(hipMemcpyAsync(hs, sout, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    // A NaN or an infinity in either operand (or sums of squares beyond the FP64 range) makes a
    // statistic non-finite: the product is then the plain FP64 GEMM's, NaN and infinities
    // propagated as IEEE arithmetic does
    bool finite = true;
    for (int q = 0; q < 5; ++q) finite = finite && std::isfinite(hs[q]);
    if (!finite) {
      ctx->fp64acc_moduli = 0;
      ctx->fp64acc_bits = 0;
      ctx->fp64acc_est = 0.0;
      return bzg_dgemm(ctx, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, ldc);
    }
    st = plan_fp64acc(k, 0, mode, pl, hs);
  } else {
    hipLaunchKernelGGL(row_max_kernel, dim3((unsigned)((m * 64 + 255) / 256)), dim3(256), 0, ctx->stream, (int)m,
                       (int)k, A, (long)lda, rmax);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (int)n,
                       cmaxB);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(col_max_kernel, dim3((unsigned)((n + 63) / 64), (unsigned)ny), dim3(256), 0, ctx->stream,
                       (int)k, (int)n, B, (long)ldb, cmaxB);
    BZG_LAUNCH_CHECK();
    st = plan_fp64acc(k, moduli, mode, pl);
  }
  if (st) return st;
  ctx->fp64acc_moduli = pl.nmod;
  ctx->fp64acc_bits = pl.P;
  ctx->fp64acc_est = pl.est;
  const int esz = mode == 0 ? 2 : 1;
  const int kq = RBYTES / esz;  // K elements per stage
  const long mp = (m + RT - 1) / RT * RT, np = (n + RT - 1) / RT * RT;
  const long kp = (k + kq - 1) / kq * kq;
  const size_t a_bytes = (size_t)pl.nmod * plane_elems(mp * kp * esz, 1);
  const size_t b_bytes = (size_t)pl.nmod * plane_elems(np * kp * esz, 1);
  const size_t c_bytes = (size_t)pl.nmod * plane_elems(mp * np * (mode == 0 ? 2 : 1), 1);
  const size_t misc_bytes = 4096 + ((size_t)(mp + np) * sizeof(int) + 255) / 256 * 256;
  const size_t total = a_bytes + b_bytes + c_bytes + misc_bytes;
  void* ws = nullptr;
  st = workspace(ctx, 4, total, &ws);
  if (st) return st;
  char* base = static_cast<char*>(ws);
  CrtConsts* dcc = reinterpret_cast<CrtConsts*>(base);
  int* sA = reinterpret_cast<int*>(base + 4096);  // row exponents of A, written by residues_a
  int* tB = sA + mp;                              // column exponents of B, by residues_b
  char* Ares = base + misc_bytes;
  char* Bres = Ares + a_bytes;
  char* Cres = Bres + b_bytes;
  static_assert(sizeof(CrtConsts) <= 4096, "CRT constants block");
  // the constants block keeps the last plan's constants: upload only when they change
  if (ctx->fp64acc_cc_gen != ctx->ws[4].generation || ctx->fp64acc_cc_key != mode * 64 + pl.nmod) {
    BZG_HIP_CHECK(hipMemcpyAsync(dcc, &pl.host, sizeof(CrtConsts), hipMemcpyHostToDevice, ctx->stream));
    ctx->fp64acc_cc_gen = ctx->ws[4].generation;
    ctx->fp64acc_cc_key = mode * 64 + pl.nmod;
  }

  const int rm = pl.P > 56 ? 2 : (pl.P > 50 ? 1 : 0);
  const dim3 ga((unsigned)((kp + 4095) / 4096), (unsigned)std::min<long>(mp, 16384));
  const dim3 gb((unsigned)(np / 32), (unsigned)((kp + 127) / 128));
#define BZG_RES(MD, RMV, T)                                                                          \
  hipLaunchKernelGGL((residues_a_kernel<MD, RMV>), ga, dim3(256), 0, ctx->stream, (int)m, (int)k, A, (long)lda, \
                     rmax, pl.P, sA, (int)mp, (int)kp, pl.nmod, dcc, reinterpret_cast<T*>(Ares));     \
  hipLaunchKernelGGL((residues_b_kernel<MD, RMV>), gb, dim3(256), 0, ctx->stream, (int)k, (int)n, B, (long)ldb, \
                     cmaxB, pl.P, tB, (int)np, (int)kp, pl.nmod, dcc, reinterpret_cast<T*>(Bres))
  if (mode == 0) {
    if (rm == 0) { BZG_RES(0, 0, __bf16); } else if (rm == 1) { BZG_RES(0, 1, __bf16); } else { BZG_RES(0, 2, __bf16); }
  } else {
    if (rm == 0) { BZG_RES(1, 0, int8_t); } else if (rm == 1) { BZG_RES(1, 1, int8_t); } else { BZG_RES(1, 2, int8_t); }
  }
#undef BZG_RES
  BZG_LAUNCH_CHECK();
  const size_t lds = 4 * RSTAGE;
  if (!ctx->residue_lds_attr[mode]) {
    if (mode == 0)
      BZG_HIP_CHECK(hipFuncSetAttribute((const void*)residue_gemm_kernel<0, 2>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    else
      BZG_HIP_CHECK(hipFuncSetAttribute((const void*)residue_gemm_kernel<1, 4>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ctx->residue_lds_attr[mode] = true;
  }
  {
  ProfScope prof_gemm(ctx, "fp64acc_gemm");
  if (mode == 0)
    hipLaunchKernelGGL((residue_gemm_kernel<0, 2>), dim3((unsigned)((mp / RT) * (np / 128)), 1, pl.nmod),
                       dim3(256), lds, ctx->stream, (int)mp, (int)np, (int)kp, Ares, Bres,
                       reinterpret_cast<int16_t*>(Cres), dcc);
  else
    hipLaunchKernelGGL((residue_gemm_kernel<1, 4>), dim3((unsigned)((mp / RT) * (np / RT)), 1, pl.nmod),
                       dim3(256), lds, ctx->stream, (int)mp, (int)np, (int)kp, Ares, Bres,
                       reinterpret_cast<int8_t*>(Cres), dcc);
  BZG_LAUNCH_CHECK();
  }
  const dim3 gc((unsigned)((n + 256 * kCrtCols - 1) / (256 * kCrtCols)), (unsigned)std::min<long>(m, 16384));
  const int nch = pl.host.nch;
  BZG_REQUIRE(nch >= 2 && nch <= kMaxChunks, BZG_ERR_UNSUPPORTED, "fp64acc: CRT chunk count out of range");
#define BZG_CRT(R, NC)                                                                             \
  hipLaunchKernelGGL((crt_kernel<R, NC>), gc, dim3(256), 0, ctx->stream, (int)m, (int)n, (int)mp, (int)np, \
                     pl.nmod, reinterpret_cast<const R*>(Cres), sA, tB, dcc, C, (long)ldc)
#define BZG_CRT_ALL(R)                     \
  switch (nch) {                           \
    case 2: BZG_CRT(R, 2); break;          \
    case 3: BZG_CRT(R, 3); break;          \
    case 4: BZG_CRT(R, 4); break;          \
    case 5: BZG_CRT(R, 5); break;          \
    default: BZG_CRT(R, 6); break;         \
  }
  if (mode == 0) {
    BZG_CRT_ALL(int16_t)
  } else {
    BZG_CRT_ALL(int8_t)
  }
#undef BZG_CRT_ALL
#undef BZG_CRT
  BZG_LAUNCH_CHECK();
  if (moduli != 0) {
    // the fixed count has no statistics pass: rows / columns with a NaN or an infinity are
    // found on the device (maxima of +Inf) and recomputed in plain FP64
    hipLaunchKernelGGL(nonfinite_fixup_kernel, dim3((unsigned)(m + n)), dim3(256), 0, ctx->stream, (int)m,
                       (int)n, (int)k, A, (long)lda, B, (long)ldb, rmax, cmaxB, C, (long)ldc);
    BZG_LAUNCH_CHECK();
  }
  return BZG_OK;
}

namespace {
// One launch sequence: 1 <= k <= kIgemmDigitsMaxK (the class sums stay below 2^31).
bzg_status igemm_digits_chunk(bzg_context* ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k, int32_t alpha,
                              const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb, int32_t beta,
                              int32_t* C, int64_t ldc) {
  BZG_REQUIRE(k >= 1 && k <= kIgemmDigitsMaxK, BZG_ERR_UNSUPPORTED, "igemm (INT8 digits): chunk k out of range");
  const long mp = (m + RT - 1) / RT * RT, np = (n + RT - 1) / RT * RT;
  const long kp = (k + RBYTES - 1) / RBYTES * RBYTES;
  const size_t a_bytes = ((size_t)mp * 4 * kp + 255) / 256 * 256;
  const size_t b_bytes = ((size_t)np * 4 * kp + 255) / 256 * 256;
  const long cplane = plane_elems(mp * np, sizeof(int32_t));
  const size_t c_bytes = (size_t)4 * cplane * sizeof(int32_t);
  void* ws = nullptr;
  bzg_status st = workspace(ctx, 12, a_bytes + b_bytes + c_bytes, &ws);
  if (st) return st;
  char* Ad = static_cast<char*>(ws);
  char* Bd = Ad + a_bytes;
  int32_t* P = reinterpret_cast<int32_t*>(Bd + b_bytes);
  // digit rows of op(A) (m x k) and of op(B)^T (n x k)
  const dim3 grows_a((unsigned)((kp + 4095) / 4096), (unsigned)std::min<long>(mp, 16384));
  const dim3 grows_b((unsigned)((kp + 4095) / 4096), (unsigned)std::min<long>(np, 16384));
  if (!ta)
    hipLaunchKernelGGL(digits_rows_kernel<false>, grows_a, dim3(256), 0, ctx->stream, (int)m, (int)k, A, (long)lda,
                       (int)mp, (int)kp, Ad);
  else
    hipLaunchKernelGGL(digits_cols_kernel<false>, dim3((unsigned)(mp / 64), (unsigned)(kp / 128)), dim3(256), 0,
                       ctx->stream, (int)m, (int)k, A, (long)lda, (int)mp, (int)kp, Ad);
  BZG_LAUNCH_CHECK();
  if (!tb)
    hipLaunchKernelGGL(digits_cols_kernel<true>, dim3((unsigned)(np / 64), (unsigned)(kp / 128)), dim3(256), 0,
                       ctx->stream, (int)n, (int)k, B, (long)ldb, (int)np, (int)kp, Bd);
  else
    hipLaunchKernelGGL(digits_rows_kernel<true>, grows_b, dim3(256), 0, ctx->stream, (int)n, (int)k, B, (long)ldb,
                       (int)np, (int)kp, Bd);
  BZG_LAUNCH_CHECK();
  const size_t lds = 4 * RSTAGE;
  if (!ctx->residue_lds_attr[2]) {
    BZG_HIP_CHECK(hipFuncSetAttribute((const void*)residue_gemm_kernel<2, 4>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ctx->residue_lds_attr[2] = true;
  }
  {
    ProfScope prof_gemm(ctx, "igemm_digits_gemm");
    hipLaunchKernelGGL((residue_gemm_kernel<2, 4>), dim3((unsigned)((mp / RT) * (np / RT)), 1, 4), dim3(256), lds,
                       ctx->stream, (int)mp, (int)np, (int)kp, Ad, Bd, P, nullptr);
    BZG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(digits_combine_kernel, dim3((unsigned)((n + 1023) / 1024), (unsigned)std::min<long>(m, 16384)),
                     dim3(256), 0, ctx->stream, (int)m, (int)n, (int)np, cplane, (uint32_t)alpha, P,
                     (uint32_t)beta, C, (long)ldc);
  BZG_LAUNCH_CHECK();
  return BZG_OK;
}
}  // namespace

bzg_status igemm_digits_impl(bzg_context* ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k, int32_t alpha,
                             const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb, int32_t beta,
                             int32_t* C, int64_t ldc) {
  // arguments checked by the caller; m, n, k >= 1.  Longer K runs in chunks of at most
  // kIgemmDigitsChunk (a multiple of 128, so only the last chunk is padded), the later ones
  // accumulated with beta = 1: wrap-around sums are exact in any grouping.
  const int64_t nch = (k + kIgemmDigitsChunk - 1) / kIgemmDigitsChunk;
  const int64_t kc = ((k + nch - 1) / nch + 127) / 128 * 128;
  for (int64_t k0 = 0; k0 < k; k0 += kc) {
    const int64_t kk = std::min(kc, k - k0);
    const int32_t* Ak = ta ? A + k0 * lda : A + k0;
    const int32_t* Bk = tb ? B + k0 : B + k0 * ldb;
    const bzg_status st = igemm_digits_chunk(ctx, ta, tb, m, n, kk, alpha, Ak, lda, Bk, ldb,
                                             k0 == 0 ? beta : 1, C, ldc);
    if (st) return st;
  }
  return BZG_OK;
}

}  // namespace bzg

extern "C" {

bzg_status bzg_dgemm_fp64acc(bzg_context* ctx, int64_t m, int64_t n, int64_t k, const double* A,
                             int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                             int32_t moduli) {
  return bzg::dgemm_fp64acc_impl(ctx, m, n, k, A, lda, B, ldb, C, ldc, moduli, 0);
}

bzg_status bzg_dgemm_fp64acc_i8(bzg_context* ctx, int64_t m, int64_t n, int64_t k, const double* A,
                                int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                                int32_t moduli) {
  return bzg::dgemm_fp64acc_impl(ctx, m, n, k, A, lda, B, ldb, C, ldc, moduli, 1);
}

bzg_status bzg_fp64acc_last_plan(bzg_context* ctx, int32_t* moduli, int32_t* bits, double* est_rel_err) {
  BZG_REQUIRE(ctx != nullptr, BZG_ERR_INVALID_ARGUMENT, "fp64acc_last_plan: null context");
  if (moduli) *moduli = ctx->fp64acc_moduli;
  if (bits) *bits = ctx->fp64acc_bits;
  if (est_rel_err) *est_rel_err = ctx->fp64acc_est;
  return BZG_OK;
}

int32_t bzg_fp64acc_slices(int64_t k, int32_t moduli) {
  bzg::Fp64AccPlan pl;
  if (bzg::plan_fp64acc(k, moduli, 0, pl) != BZG_OK) return -1;
  return pl.nmod;
}

int32_t bzg_fp64acc_operand_bits(int64_t k, int32_t moduli, int32_t int8_slices) {
  bzg::Fp64AccPlan pl;
  if (bzg::plan_fp64acc(k, moduli, int8_slices ? 1 : 0, pl) != BZG_OK) return -1;
  return pl.P;
}

}  // extern "C"
