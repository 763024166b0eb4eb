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

#include <cmath>
#include <type_traits>
#include <cstring>
#include <vector>

namespace bzg {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kMaxModuli = 24;

constexpr int kMaxChunks = 6;  // chunks of W_l and M: 24 moduli x 9 bits <= 216 bits, >= 39 bits each

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
// Row statistics of A (m x k, lda), one wavefront per row: rmax_i = max_j |a_ij| and
// rn2_i = sum_j a_ij^2 (fixed summation order: lane-strided partials, then a butterfly).
__global__ __launch_bounds__(256) void row_stats_kernel(int m, int k, const double* __restrict__ A,
                                                        long lda, double* __restrict__ rmax,
                                                        double* __restrict__ rn2) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x % 64;
  if (wave >= m) return;
  const double* row = A + (long)wave * lda;
  double mx = 0.0, sq = 0.0;
  for (int j = lane; j < k; j += 64) {
    const double v = row[j];
    mx = fmax(mx, fabs(v));
    sq = fma(v, v, sq);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, off, 64));
    sq += __shfl_xor(sq, off, 64);
  }
  if (lane == 0) {
    rmax[wave] = mx;
    rn2[wave] = sq;
  }
}

// s_i = P - 1 - ilogb(rmax_i)  (0 for zero rows): |rint(a_ij 2^s_i)| <= 2^P.
__global__ __launch_bounds__(256) void row_scale_kernel(int m, const double* __restrict__ rmax, int P,
                                                        int* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double mx = rmax[i];
  s[i] = (mx > 0.0) ? (P - 1 - ilogb(mx)) : 0;
}

// Column maxima of |B| (k x n, ldb) as FP64 bit patterns (non-negative doubles order like
// their bit patterns): a (n / 64) x (k / 256) grid of 64-column x 256-row blocks (y strides over K), combined with
// one atomicMax per column and block.  cmax must be zero on entry.
// Also writes the block's partial sum of squares of each column to cn2[blockIdx.y * n + c]
// (fixed order within the block), for the accuracy estimate of the automatic moduli count.
__global__ __launch_bounds__(256) void col_max_kernel(int k, int n, const double* __restrict__ B,
                                                      long ldb, unsigned long long* __restrict__ cmax,
                                                      double* __restrict__ cn2) {
  __shared__ double part[4][64], psq[4][64];
  const int c = blockIdx.x * 64 + threadIdx.x % 64, ph = threadIdx.x / 64;
  double mx = 0.0, sq = 0.0;
  if (c < n)
    for (int r0 = blockIdx.y * 256; r0 < k; r0 += gridDim.y * 256) {
      const int r1 = min(k, r0 + 256);
      for (int i = r0 + ph; i < r1; i += 4) {
        const double v = B[(long)i * ldb + c];
        mx = fmax(mx, fabs(v));
        sq = fma(v, v, sq);
      }
    }
  part[ph][threadIdx.x % 64] = mx;
  psq[ph][threadIdx.x % 64] = sq;
  __syncthreads();
  if (ph == 0 && c < n) {
    mx = fmax(fmax(part[0][threadIdx.x], part[1][threadIdx.x]), fmax(part[2][threadIdx.x], part[3][threadIdx.x]));
    if (cmax) atomicMax(&cmax[c], (unsigned long long)__double_as_longlong(mx));
    cn2[(long)blockIdx.y * n + c] = (psq[0][threadIdx.x] + psq[1][threadIdx.x]) + (psq[2][threadIdx.x] + psq[3][threadIdx.x]);
  }
}

// out[j] = sum_y part[y * n + j] for j < n (fixed order): column sums of squares from the
// per-block partials of col_max_kernel.
__global__ __launch_bounds__(256) void sum_partials_kernel(int ny, int n, const double* __restrict__ part,
                                                           double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double a = 0.0;
  for (int y = 0; y < ny; ++y) a += part[(long)y * n + j];
  out[j] = a;
}

// Accuracy statistics (one workgroup, fixed summation order): out[0] = sum_i D_i^2,
// out[1] = ||A||_F^2, out[2] = sum_j E_j^2, out[3] = ||B||_F^2, out[4] = sum_q ||A_:q||^2 ||B_q:||^2,
// where D_i = 2^(ilogb(rmax_i)+1) and E_j likewise for the column maxima of B; bcn2 holds the
// column sums of squares of B, acn2 those of A (k columns), brn2 the row sums of squares of B.
__global__ __launch_bounds__(1024) void fp64acc_stats_kernel(int m, int n, int k,
                                                            const double* __restrict__ rmax,
                                                            const double* __restrict__ rn2,
                                                            const unsigned long long* __restrict__ cmax,
                                                            const double* __restrict__ bcn2,
                                                            const double* __restrict__ acn2,
                                                            const double* __restrict__ brn2, double* __restrict__ out) {
  __shared__ double red[5][1024];
  const int t = threadIdx.x;
  double sa = 0, fa = 0, sb = 0, fb = 0, cc = 0;
  for (int q = t; q < k; q += 1024) cc = fma(acn2[q], brn2[q], cc);
  for (int i = t; i < m; i += 1024) {
    const double mx = rmax[i];
    if (mx > 0.0) {
      const double d = ldexp(1.0, ilogb(mx) + 1);
      sa = fma(d, d, sa);
    }
    fa += rn2[i];
  }
  for (int j = t; j < n; j += 1024) {
    const double mx = __longlong_as_double((long long)cmax[j]);
    if (mx > 0.0) {
      const double e = ldexp(1.0, ilogb(mx) + 1);
      sb = fma(e, e, sb);
    }
    fb += bcn2[j];
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

// t_j = P - 1 - ilogb(max_i |b_ij|)  (0 for zero columns).
__global__ __launch_bounds__(256) void col_scale_kernel(int n, const unsigned long long* __restrict__ cmax,
                                                        int P, int* __restrict__ t) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double mx = __longlong_as_double((long long)cmax[j]);
  t[j] = (mx > 0.0) ? (P - 1 - ilogb(mx)) : 0;
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

__device__ __forceinline__ double residue(double v, int wide, const CrtConsts* __restrict__ cc, int l) {
  return wide ? sym_mod_wide(v, cc->m[l], cc->minv[l], cc->c30[l]) : sym_mod(v, cc->m[l], cc->minv[l]);
}

// A residues: Ares[l][i][kk] (rows padded to mp, k padded to kp (a multiple of 4), zeros in the
// padding).  Grid: x covers kp in 1024-element chunks (4 per thread), y strides over rows.
template <int MODE>
__global__ __launch_bounds__(256) void residues_a_kernel(int m, int k, const double* __restrict__ A,
                                                         long lda, const int* __restrict__ s, int mp,
                                                         int kp, int nmod, int wide,
                                                         const CrtConsts* __restrict__ cc,
                                                         typename SliceOf<MODE>::T* __restrict__ out) {
  using S = SliceOf<MODE>;
  const int k0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (k0 >= kp) return;
  const long plane = (long)mp * kp;
  for (int i = blockIdx.y; i < mp; i += gridDim.y) {
    double v[4];
    const int si = i < m ? s[i] : 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kk = k0 + u;
      v[u] = (i < m && kk < k) ? rint(ldexp(A[(long)i * lda + kk], si)) : 0.0;
    }
    typename S::T* o = out + (long)i * kp + k0;
    for (int l = 0; l < nmod; ++l) {
      double r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) r[u] = residue(v[u], wide, cc, l);
      *reinterpret_cast<typename S::Pack*>(o + l * plane) = S::pack(r);
    }
  }
}

// B residues, transposed: Bres[l][j][kk] (n rows padded to np, k padded to kp).  64 x 64 tiles:
// coalesced reads along j into LDS, then 4 consecutive kk per thread, packed stores along kk.
template <int MODE>
__global__ __launch_bounds__(256) void residues_b_kernel(int k, int n, const double* __restrict__ B,
                                                         long ldb, const int* __restrict__ t, int np,
                                                         int kp, int nmod, int wide,
                                                         const CrtConsts* __restrict__ cc,
                                                         typename SliceOf<MODE>::T* __restrict__ out) {
  using S = SliceOf<MODE>;
  __shared__ double tile[64][65];
  const int kb = blockIdx.x * 64, jb = blockIdx.y * 64;
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  const int tj = (jb + tx < n) ? t[jb + tx] : 0;
  for (int r = ty; r < 64; r += 4) {  // read B[kb+r][jb+tx]
    const int kk = kb + r, j = jb + tx;
    double v = 0.0;
    if (kk < k && j < n) v = rint(ldexp(B[(long)kk * ldb + j], tj));
    tile[r][tx] = v;
  }
  __syncthreads();
  const long plane = (long)np * kp;
  const int jj = threadIdx.x / 16, g = threadIdx.x % 16;  // 16 rows of j per pass, 4 kk each
  for (int jr = jj; jr < 64; jr += 16) {
    const int j = jb + jr, kk = kb + 4 * g;
    if (j >= np || kk >= kp) continue;
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = tile[4 * g + u][jr];
    typename S::T* o = out + (long)j * kp + kk;
    for (int l = 0; l < nmod; ++l) {
      double r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) r[u] = residue(v[u], wide, cc, l);
      *reinterpret_cast<typename S::Pack*>(o + l * plane) = S::pack(r);
    }
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
// Pipeline per K stage (128 bytes of K, 4 MFMA k-steps): the global loads of stage st+1 are
// issued at the top; the fragments of k-step ks+1 are read from LDS into a second register set
// while the 4 x FN MFMAs of k-step ks run; the staged registers of st+1 are written into the
// other LDS buffer during the last k-step (that buffer was last read before the previous
// barrier); one barrier closes the stage.
template <int MODE>
using ResOut = typename std::conditional<MODE == 0, int16_t, int8_t>::type;

template <int MODE, int FN>
__global__ __launch_bounds__(256, 1) void residue_gemm_kernel(int mp, int np, int kp,
                                                              const char* __restrict__ Ares,
                                                              const char* __restrict__ Bres,
                                                              ResOut<MODE>* __restrict__ Cres,
                                                              const CrtConsts* __restrict__ cc) {
  constexpr int ESZ = (MODE == 0) ? 2 : 1;          // bytes per element
  constexpr int KSTEP_B = 32;                        // bytes of K per MFMA (16 bf16 / 32 i8)
  constexpr int NKS = RBYTES / KSTEP_B;              // MFMA k-steps per stage
  constexpr int kStoreStep = NKS - 1;  // k-step that carries the next stage's LDS stores (the
                                       // last: the global loads get the whole stage to land)
  constexpr int BN = 64 * FN;                        // workgroup tile columns
  constexpr int NBCH = BN * 8 / 256;                 // B chunks staged per thread
  extern __shared__ __attribute__((aligned(16))) char smem[];
  char* As = smem;                   // [2][RT rows][RBYTES]
  char* Bs = smem + 2 * RSTAGE;      // [2][BN rows][RBYTES]

  const int l = blockIdx.z;
  const long a_plane = (long)mp * kp * ESZ, b_plane = (long)np * kp * ESZ;
  const char* Ag = Ares + l * a_plane;
  const char* Bg = Bres + l * b_plane;
  const float mf = (float)cc->m[l];
  const float minvf = 1.0f / mf;

  const int tiles_n = np / BN;
  const int tiles_m = mp / RT;
  const int nt = tiles_m * tiles_n;
  int bid = blockIdx.x;
  if (nt >= 64 && nt % 8 == 0) bid = (bid % 8) * (nt / 8) + bid / 8;  // XCD-contiguous
  // walk tiles in column-major groups of 4 tile rows for L2 reuse
  const int G = 4;
  const int group = bid / (G * tiles_n);
  const int first_m = group * G;
  const int gsize = min(tiles_m - first_m, G);
  const int in_group = bid % (G * tiles_n);
  const int tm = first_m + in_group % gsize, tn = in_group / gsize;
  const long row0 = (long)tm * RT, col0 = (long)tn * BN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * 128, wn0 = (wave & 1) * (32 * FN);
  const long kbytes = (long)kp * ESZ;

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
      ra[e] = *reinterpret_cast<const v4i*>(Ag + (row0 + r) * kbytes + kb + c * 16);
    }
#pragma unroll
    for (int e = 0; e < NBCH; ++e) {
      const int q = tid + 256 * e, r = q >> 3, c = q & 7;
      rb[e] = *reinterpret_cast<const v4i*>(Bg + (col0 + r) * kbytes + kb + c * 16);
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

  const int nst = (int)(kbytes / RBYTES);
  // K elements per stage: 64 (bf16) / 128 (i8); bf16 reduces every 4 stages (256 K)
  gload(0);
  sstore(0);
  __syncthreads();
  lfrag(0, 0, fa[0], fb[0]);
  for (int st = 0; st + 1 < nst; ++st) {
    const int buf = st & 1;
    gload(st + 1);
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int cur = ks & 1;
      if (ks + 1 < NKS) lfrag(buf, ks + 1, fa[cur ^ 1], fb[cur ^ 1]);
      kstep(cur);
      if (ks == kStoreStep) sstore(buf ^ 1);
    }
    // instruction order of the stage: each MFMA of k-step 0 is followed by one global load of
    // the next stage, and the first 4 + FN MFMAs of k-steps 0..NKS-2 by one LDS fragment read
    // of the following k-step; the LDS stores of the next stage ride on the last k-step.
    constexpr int NM = 4 * FN;
    if constexpr (MODE == 1)  // (the BF16 tile measured faster with the compiler's own order)
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
      for (int t = 0; t < NM; ++t) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);               // MFMA
        if (ks == 0 && t < 8 + NBCH) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
        if (ks + 1 < NKS && t < 4 + FN) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        if (ks == kStoreStep && t < 8 + NBCH) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
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

  // epilogue: symmetric residue mod m
  const long cplane = (long)mp * np;
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

// ------------------------------------------------------------------------------------------
// device: CRT recombination and scaling
// ------------------------------------------------------------------------------------------
template <typename R>
struct Quad;
template <>
struct Quad<int8_t> {
  using V = char4;
};
template <>
struct Quad<int16_t> {
  using V = short4;
};

// 4 consecutive columns per thread (np is a multiple of 256); grid x covers np / 1024, y strides
// over rows.
template <typename R>
__global__ __launch_bounds__(256) void crt_kernel(int m, int n, int mp, int np, int nmod,
                                                  const R* __restrict__ Cres,
                                                  const int* __restrict__ s, const int* __restrict__ t,
                                                  const CrtConsts* __restrict__ cc, double* __restrict__ C,
                                                  long ldc) {
  const int j0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (j0 >= n) return;
  const long plane = (long)mp * np;
  const int nch = cc->nch;
  int tj[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) tj[u] = (j0 + u < n) ? t[j0 + u] : 0;
  for (int i = blockIdx.y; i < m; i += gridDim.y) {
    const R* src = Cres + (long)i * np + j0;
    // S[c] = sum_l r_l w[c][l]: every term a multiple of chunk c's unit, below 2^53 units: exact
    double S[kMaxChunks][4];
#pragma unroll
    for (int ch = 0; ch < kMaxChunks; ++ch)
#pragma unroll
      for (int u = 0; u < 4; ++u) S[ch][u] = 0.0;
    for (int l = 0; l < nmod; ++l) {
      const auto q4 = *reinterpret_cast<const typename Quad<R>::V*>(src + l * plane);
      const double r[4] = {(double)q4.x, (double)q4.y, (double)q4.z, (double)q4.w};
#pragma unroll
      for (int ch = 0; ch < kMaxChunks; ++ch) {
        if (ch < nch) {
          const double wl = cc->w[ch][l];
#pragma unroll
          for (int u = 0; u < 4; ++u) S[ch][u] = fma(r[u], wl, S[ch][u]);
        }
      }
    }
    const int si = s[i];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (j0 + u >= n) break;
      // X = sum_c S[c] = C' + q M with |C'| <= M/4: q from the leading chunk, then the chunks of
      // X - q M (each exact) summed with error-free transformations, rounded once
      const double q = rint(S[0][u] * cc->Minv);
      double hi = fma(-q, cc->M[0], S[0][u]);  // exact
      double lo = 0.0;
#pragma unroll
      for (int ch = 1; ch < kMaxChunks; ++ch) {
        if (ch < nch) {
          const double t = fma(-q, cc->M[ch], S[ch][u]);  // exact
          const double sum = hi + t;                      // TwoSum(hi, t)
          const double bb = sum - hi;
          lo += (hi - (sum - bb)) + (t - bb);
          hi = sum;
        }
      }
      const double Cp = hi + lo;
      C[(long)i * ldc + j0 + u] = ldexp(Cp, -(si + tj[u]));
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
constexpr double kAutoTarget = 4e-15;

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
    for (nm = 4; nm < kMaxModuli; ++nm) {
      const int P = std::min(operand_bits(pick_moduli(nm, top), k), 104);
      if (P >= 1 && auto_estimate(P, k, stats) <= kAutoTarget) break;
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
  build_crt(ms, pl.host, mode == 1 ? 40 : 39);
  return BZG_OK;
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
  // ---- operand statistics (row / column maxima, sums of squares) ----
  const long ny = std::min<long>((k + 255) / 256, 16384);
  const long nya = std::min<long>((m + 255) / 256, 16384);
  const size_t stat_bytes = 64 + (size_t)m * 16 + (size_t)n * 8 + (size_t)ny * n * 8 + (size_t)nya * k * 8 +
                            (size_t)k * 16 + (size_t)(n + k) * 8;
  void* sws = nullptr;
  bzg_status st = workspace(ctx, 8, stat_bytes, &sws);
  if (st) return st;
  double* sout = static_cast<double*>(sws);
  double* rmax = sout + 8;
  double* rn2 = rmax + m;
  unsigned long long* cmaxB = reinterpret_cast<unsigned long long*>(rn2 + m);
  double* cn2 = reinterpret_cast<double*>(cmaxB + n);
  double* acn2 = cn2 + ny * n;     // partial column sums of squares of A
  double* bmax = acn2 + nya * k;   // row maxima of B (unused) and row sums of squares
  double* brn2 = bmax + k;
  double* bcn2 = brn2 + k;         // column sums of squares of B and of A
  double* acol = bcn2 + n;
  hipLaunchKernelGGL(row_stats_kernel, dim3((unsigned)((m * 64 + 255) / 256)), dim3(256), 0, ctx->stream, (int)m,
                     (int)k, A, (long)lda, rmax, rn2);
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (int)n,
                     cmaxB);
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(col_max_kernel, dim3((unsigned)((n + 63) / 64), (unsigned)ny), dim3(256), 0, ctx->stream,
                     (int)k, (int)n, B, (long)ldb, cmaxB, cn2);
  BZG_LAUNCH_CHECK();
  Fp64AccPlan pl;
  if (moduli == 0) {
    // automatic count: column norms of A and row norms of B complete the statistics, which
    // come back to the host (one stream synchronisation)
    hipLaunchKernelGGL(col_max_kernel, dim3((unsigned)((k + 63) / 64), (unsigned)nya), dim3(256), 0, ctx->stream,
                       (int)m, (int)k, A, (long)lda, (unsigned long long*)nullptr, acn2);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(row_stats_kernel, dim3((unsigned)((k * 64 + 255) / 256)), dim3(256), 0, ctx->stream, (int)k,
                       (int)n, B, (long)ldb, bmax, brn2);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (int)ny,
                       (int)n, cn2, bcn2);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, ctx->stream, (int)nya,
                       (int)k, acn2, acol);
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(fp64acc_stats_kernel, dim3(1), dim3(1024), 0, ctx->stream, (int)m, (int)n, (int)k, rmax, rn2,
                       cmaxB, bcn2, acol, brn2, sout);
    BZG_LAUNCH_CHECK();
    double hs[5];
    BZG_HIP_CHECK(hipMemcpyAsync(hs, sout, sizeof hs, hipMemcpyDeviceToHost, ctx->stream));
    BZG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    st = plan_fp64acc(k, 0, mode, pl, hs);
  } else {
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
  const size_t a_bytes = (size_t)pl.nmod * mp * kp * esz;
  const size_t b_bytes = (size_t)pl.nmod * np * kp * esz;
  const size_t c_bytes = (size_t)pl.nmod * mp * np * (mode == 0 ? sizeof(int16_t) : sizeof(int8_t));
  const size_t misc_bytes = 4096 + (size_t)(mp + np) * sizeof(int);
  const size_t total = a_bytes + b_bytes + c_bytes + misc_bytes;
  void* ws = nullptr;
  st = workspace(ctx, 4, total, &ws);
  if (st) return st;
  char* base = static_cast<char*>(ws);
  CrtConsts* dcc = reinterpret_cast<CrtConsts*>(base);
  int* sA = reinterpret_cast<int*>(base + 4096);
  int* tB = sA + mp;
  char* Ares = base + misc_bytes;
  char* Bres = Ares + a_bytes;
  char* Cres = Bres + b_bytes;
  static_assert(sizeof(CrtConsts) <= 4096, "CRT constants block");
  BZG_HIP_CHECK(hipMemcpyAsync(dcc, &pl.host, sizeof(CrtConsts), hipMemcpyHostToDevice, ctx->stream));

  hipLaunchKernelGGL(row_scale_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, ctx->stream, (int)m, rmax,
                     pl.P, sA);
  BZG_LAUNCH_CHECK();
  hipLaunchKernelGGL(col_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (int)n,
                     cmaxB, pl.P, tB);
  BZG_LAUNCH_CHECK();
  const int wide = pl.P > 56 ? 1 : 0;
  const dim3 ga((unsigned)((kp + 1023) / 1024), (unsigned)std::min<long>(mp, 16384));
  const dim3 gb((unsigned)(kp / 64), (unsigned)(np / 64));
  if (mode == 0) {
    hipLaunchKernelGGL(residues_a_kernel<0>, ga, dim3(256), 0, ctx->stream, (int)m, (int)k, A, (long)lda, sA,
                       (int)mp, (int)kp, pl.nmod, wide, dcc, reinterpret_cast<__bf16*>(Ares));
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(residues_b_kernel<0>, gb, dim3(256), 0, ctx->stream, (int)k, (int)n, B, (long)ldb, tB,
                       (int)np, (int)kp, pl.nmod, wide, dcc, reinterpret_cast<__bf16*>(Bres));
    BZG_LAUNCH_CHECK();
  } else {
    hipLaunchKernelGGL(residues_a_kernel<1>, ga, dim3(256), 0, ctx->stream, (int)m, (int)k, A, (long)lda, sA,
                       (int)mp, (int)kp, pl.nmod, wide, dcc, reinterpret_cast<int8_t*>(Ares));
    BZG_LAUNCH_CHECK();
    hipLaunchKernelGGL(residues_b_kernel<1>, gb, dim3(256), 0, ctx->stream, (int)k, (int)n, B, (long)ldb, tB,
                       (int)np, (int)kp, pl.nmod, wide, dcc, reinterpret_cast<int8_t*>(Bres));
    BZG_LAUNCH_CHECK();
  }
  const size_t lds = 4 * RSTAGE;
  static bool attr_set[2] = {false, false};
  if (!attr_set[mode]) {
    if (mode == 0)
      BZG_HIP_CHECK(hipFuncSetAttribute((const void*)residue_gemm_kernel<0, 2>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    else
      BZG_HIP_CHECK(hipFuncSetAttribute((const void*)residue_gemm_kernel<1, 4>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set[mode] = true;
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
  const dim3 gc((unsigned)((n + 1023) / 1024), (unsigned)std::min<long>(m, 16384));
  if (mode == 0)
    hipLaunchKernelGGL(crt_kernel<int16_t>, gc, dim3(256), 0, ctx->stream, (int)m, (int)n, (int)mp, (int)np,
                       pl.nmod, reinterpret_cast<const int16_t*>(Cres), sA, tB, dcc, C, (long)ldc);
  else
    hipLaunchKernelGGL(crt_kernel<int8_t>, gc, dim3(256), 0, ctx->stream, (int)m, (int)n, (int)mp, (int)np,
                       pl.nmod, reinterpret_cast<const int8_t*>(Cres), sA, tB, dcc, C, (long)ldc);
  BZG_LAUNCH_CHECK();
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
