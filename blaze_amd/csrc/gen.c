/*
 * Synthetic-input generators for the multiplication & solve workloads (host side, portable).
 *
 * Two random sources, both written from their published definitions:
 *  - MT19937-64 (Matsumoto & Nishimura 2000, 64-bit variant, init_genrand64): the sequential
 *    stream for the 512^3 GEMM inputs (A first, then B, row-major order);
 *  - SplitMix64 (Steele, Lea & Flood 2014) on a counter: every value is a pure function of
 *    (seed, index), so large inputs are generated in parallel and any entry can be recomputed.
 * Uniforms on [-1,1) are u = (x >> 11) * 2^-53 * 2 - 1; normals use Box-Muller on two
 * counter uniforms.  None of std::*_distribution is used (those are implementation-defined).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- MT19937-64 ------------------------------------------------------------------------ */
#define MT_NN 312
#define MT_MM 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

typedef struct {
  uint64_t mt[MT_NN];
  int mti;
} mt64_t;

static void mt64_seed(mt64_t* s, uint64_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < MT_NN; i++)
    s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
  s->mti = MT_NN;
}

static uint64_t mt64_next(mt64_t* s) {
  static const uint64_t mag01[2] = {0ULL, MT_MATRIX_A};
  uint64_t x;
  if (s->mti >= MT_NN) {
    int i;
    for (i = 0; i < MT_NN - MT_MM; i++) {
      x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
      s->mt[i] = s->mt[i + MT_MM] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    for (; i < MT_NN - 1; i++) {
      x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
      s->mt[i] = s->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    x = (s->mt[MT_NN - 1] & MT_UM) | (s->mt[0] & MT_LM);
    s->mt[MT_NN - 1] = s->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    s->mti = 0;
  }
  x = s->mt[s->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= (x >> 43);
  return x;
}

static inline double u53_pm1(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53 * 2.0 - 1.0; }
static inline double u53_01(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

/* Raw MT19937-64 outputs (for checking the generator against its published test vector). */
void bzg_gen_mt64_raw(uint64_t seed, int64_t n, uint64_t* out) {
  mt64_t s;
  mt64_seed(&s, seed);
  for (int64_t i = 0; i < n; i++) out[i] = mt64_next(&s);
}

/* n1 uniforms on [-1,1) into a, then n2 into b, from one MT19937-64 stream. */
void bzg_gen_uniform_mt(uint64_t seed, int64_t n1, double* a, int64_t n2, double* b) {
  mt64_t s;
  mt64_seed(&s, seed);
  for (int64_t i = 0; i < n1; i++) a[i] = u53_pm1(mt64_next(&s));
  for (int64_t i = 0; i < n2; i++) b[i] = u53_pm1(mt64_next(&s));
}

/* ---- SplitMix64 counter generator -------------------------------------------------------- */
static inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t ctr_bits(uint64_t seed, uint64_t stream, uint64_t i) {
  return splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) + i * 0x9E3779B97F4A7C15ULL);
}
static inline double ctr_pm1(uint64_t seed, uint64_t stream, uint64_t i) {
  return u53_pm1(ctr_bits(seed, stream, i));
}
static inline double ctr_normal(uint64_t seed, uint64_t stream, uint64_t i) {
  /* Box-Muller: u1 in (0,1], u2 in [0,1). */
  double u1 = 1.0 - u53_01(ctr_bits(seed, stream, 2 * i));
  double u2 = u53_01(ctr_bits(seed, stream, 2 * i + 1));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void bzg_gen_uniform_ctr(uint64_t seed, uint64_t stream, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = ctr_pm1(seed, stream, (uint64_t)i);
}

void bzg_gen_uniform_ctr_f32(uint64_t seed, uint64_t stream, int64_t n, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = (float)ctr_pm1(seed, stream, (uint64_t)i);
}

/* Integer-valued entries uniform in [lo, hi] (exact in every floating type used here). */
void bzg_gen_int_ctr(uint64_t seed, uint64_t stream, int64_t n, int32_t lo, int32_t hi, int32_t* out) {
  const uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = (int32_t)(lo + (int64_t)(ctr_bits(seed, stream, (uint64_t)i) % span));
}

/* Graded-exponent entries u * 2^e, u uniform[-1,1), e = round(N(0, sigma)). */
void bzg_gen_graded(uint64_t seed, uint64_t stream, int64_t n, double sigma, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double u = ctr_pm1(seed, stream, (uint64_t)i);
    double e = sigma > 0 ? nearbyint(sigma * ctr_normal(seed, stream + 0x5151, (uint64_t)i)) : 0.0;
    out[i] = ldexp(u, (int)e);
  }
}

/* ---- 7-point 3-D Laplacian on an nx^3 grid (lexicographic i,j,k; k fastest) ------------- */
int64_t bzg_gen_laplace3d_nnz(int64_t nx) {
  const int64_t n = nx * nx * nx;
  return 7 * n - 6 * nx * nx;
}

void bzg_gen_laplace3d(int64_t nx, int32_t* rowptr, int32_t* colidx, double* val) {
  const int64_t n = nx * nx * nx;
  /* Row r = (i*nx + j)*nx + k; neighbours sorted by column: -nx^2, -nx, -1, 0, +1, +nx, +nx^2 */
  rowptr[0] = 0;
  for (int64_t r = 0; r < n; r++) {
    const int64_t i = r / (nx * nx), j = (r / nx) % nx, k = r % nx;
    rowptr[r + 1] = rowptr[r] + 1 + (i > 0) + (j > 0) + (k > 0) + (k < nx - 1) + (j < nx - 1) +
                    (i < nx - 1);
  }
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    const int64_t i = r / (nx * nx), j = (r / nx) % nx, k = r % nx;
    int64_t p = rowptr[r];
    if (i > 0) { colidx[p] = (int32_t)(r - nx * nx); val[p++] = -1.0; }
    if (j > 0) { colidx[p] = (int32_t)(r - nx); val[p++] = -1.0; }
    if (k > 0) { colidx[p] = (int32_t)(r - 1); val[p++] = -1.0; }
    colidx[p] = (int32_t)r; val[p++] = 6.0;
    if (k < nx - 1) { colidx[p] = (int32_t)(r + 1); val[p++] = -1.0; }
    if (j < nx - 1) { colidx[p] = (int32_t)(r + nx); val[p++] = -1.0; }
    if (i < nx - 1) { colidx[p] = (int32_t)(r + nx * nx); val[p++] = -1.0; }
  }
}

/* ---- power-law (truncated Zipf) random rows ---------------------------------------------- */
/* Row lengths k in [1, kmax] with P(k) ~ k^-s, by inverse CDF on a counter uniform. */
void bzg_gen_zipf_rowptr(uint64_t seed, int64_t n, double s, int32_t kmax, int32_t* rowptr) {
  double* cdf = (double*)malloc(sizeof(double) * (size_t)kmax);
  double acc = 0;
  for (int k = 1; k <= kmax; k++) { acc += pow((double)k, -s); cdf[k - 1] = acc; }
  for (int k = 0; k < kmax; k++) cdf[k] /= acc;
  int32_t* len = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    double u = u53_01(ctr_bits(seed, 101, (uint64_t)r));
    int lo = 0, hi = kmax - 1;
    while (lo < hi) { int mid = (lo + hi) / 2; if (cdf[mid] > u) hi = mid; else lo = mid + 1; }
    len[r] = lo + 1;
  }
  rowptr[0] = 0;
  for (int64_t r = 0; r < n; r++) rowptr[r + 1] = rowptr[r] + len[r];
  free(len);
  free(cdf);
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Distinct uniform columns per row (sorted) and N(0,1) values. */
void bzg_gen_zipf_fill(uint64_t seed, int64_t n, int64_t ncols, const int32_t* rowptr,
                       int32_t* colidx, double* val) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < n; r++) {
    const int32_t b = rowptr[r], e = rowptr[r + 1];
    const int32_t k = e - b;
    uint64_t draw = 0;
    int32_t have = 0;
    while (have < k) {
      int32_t c = (int32_t)(ctr_bits(seed, 202 + (uint64_t)r * 7919ULL, draw++) % (uint64_t)ncols);
      int dup = 0;
      for (int32_t t = 0; t < have; t++)
        if (colidx[b + t] == c) { dup = 1; break; }
      if (!dup) colidx[b + have++] = c;
    }
    qsort(colidx + b, (size_t)k, sizeof(int32_t), cmp_i32);
    for (int32_t t = 0; t < k; t++) val[b + t] = ctr_normal(seed, 303, (uint64_t)(b + t));
  }
}

/* ---- dense solve systems ------------------------------------------------------------------ */
/* Haar-distributed random orthogonal factors by Stewart's method (G. W. Stewart, "The efficient
 * generation of random orthogonal matrices with an application to condition estimators",
 * SIAM J. Numer. Anal. 17(3), 1980): Q = H_0 H_1 ... H_{n-2} D, where H_k is the Householder
 * reflector of a fresh N(0, I) vector x_k of length n - k (acting on coordinates k..n-1), taking
 * x_k to -sign(x_k[0]) ||x_k|| e_0, and D = diag(d) with d_k = -sign(x_k[0]) (d_{n-1} = the sign
 * of one more N(0,1) draw).  This is the distribution of the Q factor of the Householder QR of
 * an n x n Gaussian matrix with the signs of R's diagonal made positive (fresh Gaussian vectors
 * stand in for the transformed columns by the rotational invariance of N(0, I)).
 * The reflector vectors v_k (x_k with sign(x_k[0]) ||x_k|| added to its first entry) are stored
 * one after another in `vs` (n(n+1)/2 - 1 values), their factors 2 / v_k^T v_k in `f`. */
static void haar_reflectors(uint64_t seed, uint64_t stream, int n, double* vs, double* f, double* d) {
  int64_t pos = 0;
  for (int k = 0; k + 1 < n; k++) {
    const int len = n - k;
    double* v = vs + pos;
    double nrm2 = 0.0;
    for (int i = 0; i < len; i++) {
      v[i] = ctr_normal(seed, stream, (uint64_t)pos + (uint64_t)i);
      nrm2 += v[i] * v[i];
    }
    const double sg = v[0] < 0.0 ? -1.0 : 1.0;
    v[0] += sg * sqrt(nrm2);          /* H x = -sg ||x|| e_0 */
    double vtv = 0.0;
    for (int i = 0; i < len; i++) vtv += v[i] * v[i];
    f[k] = vtv > 0.0 ? 2.0 / vtv : 0.0;
    d[k] = -sg;                       /* R_kk = -sg ||x||: its sign is made positive */
    pos += len;
  }
  if (n >= 1) d[n - 1] = ctr_normal(seed, stream, (uint64_t)pos) < 0.0 ? -1.0 : 1.0;
}

/* M <- Q M (side 0) or M <- M Q^T (side 1) for Q = H_0 ... H_{n-2} D, M n x n row-major. */
static void haar_apply(int n, const double* vs, const double* f, const double* d, double* M, int side,
                       double* tmp) {
  int64_t end = 0;  /* one past the last reflector vector */
  for (int k = 0; k + 1 < n; k++) end += n - k;
  if (side == 0) {
    /* Q M = H_0 (H_1 (... H_{n-2} (D M))) */
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) M[(int64_t)i * n + j] *= d[i];
    int64_t pos = end;
    for (int k = n - 2; k >= 0; k--) {
      const int len = n - k;
      pos -= len;
      const double* v = vs + pos;
      for (int j = 0; j < n; j++) tmp[j] = 0.0;
      for (int i = 0; i < len; i++) {
        const double vi = v[i];
        const double* row = M + (int64_t)(k + i) * n;
        for (int j = 0; j < n; j++) tmp[j] += vi * row[j];
      }
      for (int i = 0; i < len; i++) {
        const double c = f[k] * v[i];
        double* row = M + (int64_t)(k + i) * n;
        for (int j = 0; j < n; j++) row[j] -= c * tmp[j];
      }
    }
  } else {
    /* M Q^T = ((M D) H_{n-2}) ... H_0  (the reflectors are symmetric) */
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) M[(int64_t)i * n + j] *= d[j];
    int64_t pos = end;
    for (int k = n - 2; k >= 0; k--) {
      const int len = n - k;
      pos -= len;
      const double* v = vs + pos;
      for (int i = 0; i < n; i++) {
        double* row = M + (int64_t)i * n + k;
        double s = 0.0;
        for (int j = 0; j < len; j++) s += row[j] * v[j];
        s *= f[k];
        for (int j = 0; j < len; j++) row[j] -= s * v[j];
      }
    }
  }
}

/* Q itself (n x n row-major, n <= 512), for the generator's tests. */
void bzg_gen_haar_orthogonal(uint64_t seed, uint64_t stream, int32_t n, double* Q) {
  if (n <= 0 || n > 512) return;
  double* vs = (double*)malloc(sizeof(double) * ((size_t)n * (n + 1) / 2 + 1));
  double* f = (double*)malloc(sizeof(double) * (size_t)n);
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
  if (vs && f && d && tmp) {
    haar_reflectors(seed, stream, n, vs, f, d);
    memset(Q, 0, sizeof(double) * (size_t)n * (size_t)n);
    for (int i = 0; i < n; i++) Q[(int64_t)i * n + i] = 1.0;
    haar_apply(n, vs, f, d, Q, 0, tmp);
  }
  free(vs); free(f); free(d); free(tmp);
}

/* kind 0: A = R + n*I, R uniform[-1,1) (diagonally dominant).
 * kind 1: A = Q1 diag(logspace(0, -8, n)) Q2^T with Q1, Q2 independent Haar-random orthogonal
 *         matrices (Stewart's method above: the Q factors of Gaussian matrices): singular
 *         values exactly 10^(-8 t/(n-1)), so cond_2(A) = 1e8.
 * B uniform[-1,1), n x nrhs row-major.  Systems are independent (stream = system index). */
void bzg_gen_solve_batch(uint64_t seed, int64_t count, const int32_t* n, const int32_t* nrhs,
                         const int8_t* kind, const int64_t* a_off, const int64_t* b_off,
                         double* A, double* B) {
#pragma omp parallel
  {
    double* vs = (double*)malloc(sizeof(double) * (512 * 513 / 2 + 1));
    double* f = (double*)malloc(sizeof(double) * 512);
    double* d = (double*)malloc(sizeof(double) * 512);
    double* tmp = (double*)malloc(sizeof(double) * 512);
#pragma omp for schedule(dynamic, 1)
    for (int64_t s = 0; s < count; s++) {
      const int ns = n[s];
      double* M = A + a_off[s];
      const uint64_t st = 1000003ULL * (uint64_t)s + 17;
      if (kind[s] == 0) {
        for (int64_t i = 0; i < (int64_t)ns * ns; i++) M[i] = ctr_pm1(seed, st, (uint64_t)i);
        for (int i = 0; i < ns; i++) M[(int64_t)i * ns + i] += (double)ns;
      } else if (ns > 0 && ns <= 512) {
        memset(M, 0, sizeof(double) * (size_t)ns * (size_t)ns);
        for (int i = 0; i < ns; i++)
          M[(int64_t)i * ns + i] = ns > 1 ? pow(10.0, -8.0 * (double)i / (double)(ns - 1)) : 1.0;
        haar_reflectors(seed, st + 2, ns, vs, f, d); /* Q2 */
        haar_apply(ns, vs, f, d, M, 1, tmp);
        haar_reflectors(seed, st + 1, ns, vs, f, d); /* Q1 */
        haar_apply(ns, vs, f, d, M, 0, tmp);
      }
      double* R = B + b_off[s];
      for (int64_t i = 0; i < (int64_t)ns * nrhs[s]; i++) R[i] = ctr_pm1(seed, st + 7, (uint64_t)i);
    }
    free(vs);
    free(f);
    free(d);
    free(tmp);
  }
}
