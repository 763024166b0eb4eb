/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference (Blaze 3.9-dev) algorithms on the multiplication &
 * solve path.  Only the parity tests, the smoke check and the CPU-baseline leg of the
 * benchmark load this file's library; the product path never does.  Each function names the
 * reference file:line it restates (paths relative to the Blaze source tree).
 *
 * Pinning (what checks each function independently of this code; tests/test_oracle.py):
 *   orc_gemm_*    the reference's known answers (ClassTest2.cpp:1244-1273,
 *                 AliasingTest.cpp:1803-1850); golden fixtures from the compiled reference
 *                 kernels (int32 bit-exact, double within 2 gamma_k |A||B|); exact integer
 *                 arithmetic; permutation products (a transposed operand fails).
 *   orc_gemm_dd_entries  exact rational sums of dyadic products (fractions), to 2^-100.
 *   orc_gemv_*    compiled reference DMatDVecMultExpr (oracle/_ref), exact integers.
 *   orc_csrmv_*   golden fixtures from the compiled reference SMatDVecMultExpr, bit-for-bit
 *                 (same sequential order, no contraction); the dense product of the same
 *                 integer matrix; empty rows.
 *   orc_getrf/getrs/gesv   LAPACK dgetrf factors and pivots (scipy.linalg.lu_factor fixture),
 *                 exact rational solves (fractions), P*A == L*U, info of a zero column.
 *   orc_solve_closed   golden fixtures from the reference's solve1x1..solve6x6.
 *   orc_dsgesv    FP64 solve agreement, zero refinement steps when the FP32 solve is exact,
 *                 fallback (-3) at kappa = 1e13 where FP32 refinement cannot converge.
 *   orc_backward_error   a hand-computed case.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- dense GEMM ---------------------------------------------------------------------------
 * C = A * B, row-major.  Blaze's small kernel (DMatDMatMultExpr.h:864-1887) computes every
 * element as a sequential sum over k, C(i,j) = A(i,0)*B(0,j) + ... in increasing k, with the
 * multiply-add fused when the target has FMA (blaze/math/simd/FMA.h:97-108).  `fused` selects
 * fma() or a rounded product followed by a rounded add. */
void orc_gemm_d(int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                int64_t ldb, double* C, int64_t ldc, int fused) {
  for (int64_t i = 0; i < m; i++)
    for (int64_t j = 0; j < n; j++) {
      double s = 0.0;
      for (int64_t p = 0; p < k; p++) {
        const double a = A[i * lda + p], b = B[p * ldb + j];
        if (fused) {
          s = fma(a, b, s);
        } else {
          volatile double prod = a * b;
          s = s + prod;
        }
      }
      C[i * ldc + j] = s;
    }
}

void orc_gemm_s(int64_t m, int64_t n, int64_t k, const float* A, int64_t lda, const float* B,
                int64_t ldb, float* C, int64_t ldc) {
  for (int64_t i = 0; i < m; i++)
    for (int64_t j = 0; j < n; j++) {
      float s = 0.0f;
      for (int64_t p = 0; p < k; p++) s = fmaf(A[i * lda + p], B[p * ldb + j], s);
      C[i * ldc + j] = s;
    }
}

/* Integer GEMM with wrap-around (mod 2^32) arithmetic: equal to Blaze's int kernels
 * (MMM.h:262-327, DMatDMatMultExpr.h:864-1887) whenever no partial sum overflows. */
void orc_gemm_i32(int64_t m, int64_t n, int64_t k, const int32_t* A, int64_t lda,
                  const int32_t* B, int64_t ldb, int32_t* C, int64_t ldc) {
  for (int64_t i = 0; i < m; i++)
    for (int64_t j = 0; j < n; j++) {
      uint32_t s = 0;
      for (int64_t p = 0; p < k; p++) s += (uint32_t)A[i * lda + p] * (uint32_t)B[p * ldb + j];
      C[i * ldc + j] = (int32_t)s;
    }
}

/* ---- double-double reference entries -------------------------------------------------------
 * C(i,j) of A*B to ~106 bits with error-free transformations: TwoSum (Knuth) and TwoProduct
 * via fma (Ogita, Rump & Oishi 2005, "Accurate sum and dot product", Alg. 3.1/Dot2).
 * Returns hi in out_hi and lo in out_lo for each requested (row, col). */
void orc_gemm_dd_entries(int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                         int64_t count, const int64_t* rows, const int64_t* cols, double* out_hi,
                         double* out_lo) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < count; t++) {
    const int64_t i = rows[t], j = cols[t];
    double s = 0.0, c = 0.0;
    for (int64_t p = 0; p < k; p++) {
      const double a = A[i * lda + p], b = B[p * ldb + j];
      const double h = a * b;
      const double e = fma(a, b, -h);  /* exact: a*b = h + e */
      const double s2 = s + h;         /* TwoSum(s, h) */
      const double bb = s2 - s;
      const double err = (s - (s2 - bb)) + (h - bb);
      s = s2;
      c += err + e;
    }
    const double hi = s + c;
    out_hi[t] = hi;
    out_lo[t] = c - (hi - s);
  }
}

/* ---- dense GEMV ----------------------------------------------------------------------------
 * y = A * x, row-major; every y(i) a sequential sum over j (DMatDVecMultExpr.h:528-850 keeps
 * one accumulator per row; the SIMD lanes split j, so Blaze's own order is lane-strided). */
void orc_gemv_d(int64_t m, int64_t n, const double* A, int64_t lda, const double* x, double* y) {
  for (int64_t i = 0; i < m; i++) {
    double s = 0.0;
    for (int64_t j = 0; j < n; j++) s = fma(A[i * lda + j], x[j], s);
    y[i] = s;
  }
}
void orc_gemv_s(int64_t m, int64_t n, const float* A, int64_t lda, const float* x, float* y) {
  for (int64_t i = 0; i < m; i++) {
    float s = 0.0f;
    for (int64_t j = 0; j < n; j++) s = fmaf(A[i * lda + j], x[j], s);
    y[i] = s;
  }
}
void orc_gemv_i32(int64_t m, int64_t n, const int32_t* A, int64_t lda, const int32_t* x, int32_t* y) {
  for (int64_t i = 0; i < m; i++) {
    uint32_t s = 0;
    for (int64_t j = 0; j < n; j++) s += (uint32_t)A[i * lda + j] * (uint32_t)x[j];
    y[i] = (int32_t)s;
  }
}

/* ---- CSR SpMV --------------------------------------------------------------------------------
 * y(i) = row(A,i) * x  (SMatDVecMultExpr.h:204-207) with the inner product of
 * SVecDVecInnerExpr.h:127-135:  sp = v0*x[c0]; for the rest sp += v_k*x[c_k], storage order,
 * each product rounded before the add (the reference built without FP contraction).
 * An empty row gives MultType{} = 0. */
void orc_csrmv_d(int64_t n_rows, const int32_t* rowptr, const int32_t* col, const double* val,
                 const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_rows; i++) {
    const int32_t b = rowptr[i], e = rowptr[i + 1];
    double sp = 0.0;
    if (b != e) {
      volatile double p0 = val[b] * x[col[b]];
      sp = p0;
      for (int32_t t = b + 1; t < e; t++) {
        volatile double p = val[t] * x[col[t]];
        sp = sp + p;
      }
    }
    y[i] = sp;
  }
}
void orc_csrmv_s(int64_t n_rows, const int32_t* rowptr, const int32_t* col, const float* val,
                 const float* x, float* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_rows; i++) {
    const int32_t b = rowptr[i], e = rowptr[i + 1];
    float sp = 0.0f;
    if (b != e) {
      volatile float p0 = val[b] * x[col[b]];
      sp = p0;
      for (int32_t t = b + 1; t < e; t++) {
        volatile float p = val[t] * x[col[t]];
        sp = sp + p;
      }
    }
    y[i] = sp;
  }
}
/* y = trans(S) * x for a row-major CSR S (n_rows x n_cols), y of n_cols entries: the
 * reference's TSMatDVecMultExpr assign (blaze/math/expressions/TSMatDVecMultExpr.h:329-383):
 * y is reset to +0, then for every row j of S in order and every stored (i, v) of it,
 * y[i] += v * x[j] (each product rounded before it is added). */
void orc_tcsrmv_d(int64_t n_rows, int64_t n_cols, const int32_t* rowptr, const int32_t* col,
                  const double* val, const double* x, double* y) {
  for (int64_t i = 0; i < n_cols; i++) y[i] = 0.0;
  for (int64_t j = 0; j < n_rows; j++)
    for (int32_t t = rowptr[j]; t < rowptr[j + 1]; t++) {
      volatile double p = val[t] * x[j];
      y[col[t]] = y[col[t]] + p;
    }
}
void orc_tcsrmv_s(int64_t n_rows, int64_t n_cols, const int32_t* rowptr, const int32_t* col,
                  const float* val, const float* x, float* y) {
  for (int64_t i = 0; i < n_cols; i++) y[i] = 0.0f;
  for (int64_t j = 0; j < n_rows; j++)
    for (int32_t t = rowptr[j]; t < rowptr[j + 1]; t++) {
      volatile float p = val[t] * x[j];
      y[col[t]] = y[col[t]] + p;
    }
}
void orc_csrmv_i32(int64_t n_rows, const int32_t* rowptr, const int32_t* col, const int32_t* val,
                   const int32_t* x, int32_t* y) {
  for (int64_t i = 0; i < n_rows; i++) {
    uint32_t sp = 0;
    for (int32_t t = rowptr[i]; t < rowptr[i + 1]; t++) sp += (uint32_t)val[t] * (uint32_t)x[col[t]];
    y[i] = (int32_t)sp;
  }
}

/* ---- LU with partial pivoting (LAPACK dgetf2 semantics) -------------------------------------
 * solve() for N >= 7 copies A into column-major storage and calls dgesv_ (LSE.h:5229-5257,
 * lapack/gesv.h:136-175, clapack/gesv.h:207) = dgetrf + dgetrs.  Restated unblocked: for
 * column j pick the FIRST row p >= j of maximal |a(p,j)| (idamax), swap rows j and p, divide
 * the column below the diagonal by the pivot and apply the rank-1 update.  A is n x n
 * row-major here (the same matrix LAPACK sees), ipiv is 1-based, return value = info
 * (0, or the 1-based column of the first exactly-zero pivot; factorisation continues). */
int orc_getrf(int n, double* A, int lda, int32_t* ipiv) {
  int info = 0;
  for (int j = 0; j < n; j++) {
    int p = j;
    double best = fabs(A[(int64_t)j * lda + j]);
    for (int i = j + 1; i < n; i++) {
      const double v = fabs(A[(int64_t)i * lda + j]);
      if (v > best) { best = v; p = i; }
    }
    ipiv[j] = p + 1;
    if (A[(int64_t)p * lda + j] != 0.0) {
      if (p != j)
        for (int c = 0; c < n; c++) {
          double t = A[(int64_t)j * lda + c];
          A[(int64_t)j * lda + c] = A[(int64_t)p * lda + c];
          A[(int64_t)p * lda + c] = t;
        }
      const double piv = A[(int64_t)j * lda + j];
      const double r = 1.0 / piv;
      for (int i = j + 1; i < n; i++) A[(int64_t)i * lda + j] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int i = j + 1; i < n; i++) {
      const double l = A[(int64_t)i * lda + j];
      for (int c = j + 1; c < n; c++) A[(int64_t)i * lda + c] -= l * A[(int64_t)j * lda + c];
    }
  }
  return info;
}

/* Single-precision twin (LAPACK sgetf2) used by the mixed-precision solver. */
int orc_sgetrf(int n, float* A, int lda, int32_t* ipiv) {
  int info = 0;
  for (int j = 0; j < n; j++) {
    int p = j;
    float best = fabsf(A[(int64_t)j * lda + j]);
    for (int i = j + 1; i < n; i++) {
      const float v = fabsf(A[(int64_t)i * lda + j]);
      if (v > best) { best = v; p = i; }
    }
    ipiv[j] = p + 1;
    if (A[(int64_t)p * lda + j] != 0.0f) {
      if (p != j)
        for (int c = 0; c < n; c++) {
          float t = A[(int64_t)j * lda + c];
          A[(int64_t)j * lda + c] = A[(int64_t)p * lda + c];
          A[(int64_t)p * lda + c] = t;
        }
      const float r = 1.0f / A[(int64_t)j * lda + j];
      for (int i = j + 1; i < n; i++) A[(int64_t)i * lda + j] *= r;
    } else if (info == 0) {
      info = j + 1;
    }
    for (int i = j + 1; i < n; i++) {
      const float l = A[(int64_t)i * lda + j];
      for (int c = j + 1; c < n; c++) A[(int64_t)i * lda + c] -= l * A[(int64_t)j * lda + c];
    }
  }
  return info;
}

/* dgetrs ('N'): apply the row interchanges to B (n x nrhs row-major), then L (unit) and U. */
void orc_getrs(int n, int nrhs, const double* LU, int lda, const int32_t* ipiv, double* B, int ldb) {
  for (int i = 0; i < n; i++) {
    const int p = ipiv[i] - 1;
    if (p != i)
      for (int c = 0; c < nrhs; c++) {
        double t = B[(int64_t)i * ldb + c];
        B[(int64_t)i * ldb + c] = B[(int64_t)p * ldb + c];
        B[(int64_t)p * ldb + c] = t;
      }
  }
  for (int c = 0; c < nrhs; c++) {
    for (int i = 0; i < n; i++) {
      double s = B[(int64_t)i * ldb + c];
      for (int k = 0; k < i; k++) s -= LU[(int64_t)i * lda + k] * B[(int64_t)k * ldb + c];
      B[(int64_t)i * ldb + c] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
      double s = B[(int64_t)i * ldb + c];
      for (int k = i + 1; k < n; k++) s -= LU[(int64_t)i * lda + k] * B[(int64_t)k * ldb + c];
      B[(int64_t)i * ldb + c] = s / LU[(int64_t)i * lda + i];
    }
  }
}

/* dgesv: A overwritten by its LU factors, B by the solution; returns info. */
int orc_gesv(int n, int nrhs, double* A, int lda, int32_t* ipiv, double* B, int ldb) {
  const int info = orc_getrf(n, A, lda, ipiv);
  if (info == 0) orc_getrs(n, nrhs, A, lda, ipiv, B, ldb);
  return info;
}

/* ---- closed-form solves for N <= 3 (LSE.h:172-200, 305-337, 1141-1185) -----------------------
 * N = 1: x = b / a;  N = 2: Cramer with D = a00 a11 - a01 a10;  N = 3: the cofactor form of
 * solve3x3.  Return 1 when the determinant (pivot) is exactly zero (Blaze throws
 * "Solving LSE with singular system matrix failed"), else 0.  Larger N use orc_gesv. */
int orc_solve_closed(int n, const double* A, const double* b, double* x) {
#define a_(i, j) A[(i) * n + (j)]
  if (n == 1) {
    if (a_(0, 0) == 0.0) return 1;
    x[0] = b[0] / a_(0, 0);
    return 0;
  }
  if (n == 2) {
    const double D = a_(0, 0) * a_(1, 1) - a_(0, 1) * a_(1, 0);
    if (D == 0.0) return 1;
    const double invD = 1.0 / D;
    x[0] = (a_(1, 1) * b[0] - a_(0, 1) * b[1]) * invD;
    x[1] = (a_(0, 0) * b[1] - a_(1, 0) * b[0]) * invD;
    return 0;
  }
  if (n == 3) {
    const double t1 = a_(0, 0) * a_(1, 1) - a_(1, 0) * a_(0, 1);
    const double t2 = a_(1, 0) * a_(0, 2) - a_(0, 0) * a_(1, 2);
    const double t3 = a_(0, 1) * a_(1, 2) - a_(1, 1) * a_(0, 2);
    const double D = t1 * a_(2, 2) + t2 * a_(2, 1) + t3 * a_(2, 0);
    if (D == 0.0) return 1;
    const double invD = 1.0 / D;
    const double t4 = a_(1, 1) * b[0] - a_(0, 1) * b[1];
    const double t5 = a_(0, 2) * b[1] - a_(1, 2) * b[0];
    const double t6 = a_(0, 0) * b[1] - a_(1, 0) * b[0];
    x[0] = (t4 * a_(2, 2) + t5 * a_(2, 1) + t3 * b[2]) * invD;
    x[1] = (t6 * a_(2, 2) + t2 * b[2] - t5 * a_(2, 0)) * invD;
    x[2] = (t1 * b[2] - t4 * a_(2, 0) - t6 * a_(2, 1)) * invD;
    return 0;
  }
#undef a_
  return -1;
}

/* ---- mixed-precision solve (LAPACK dsgesv, Langou et al. 2006) -------------------------------
 * No reference function: the drop-in for solve() at LSE.h:6343 with FP32 factors.
 * Per RHS set: convert A to FP32 and factor (sgetrf); x = FP32 solve of b; loop: r = b - A x in
 * FP64, stop when every column has ||r||_inf <= ||x||_inf * ||A||_inf * eps * sqrt(n), else
 * solve A d = r with the FP32 factors (r rounded to FP32) and x += d.  After max_iter steps,
 * or when the FP32 factorisation hits a zero pivot, fall back to the FP64 LU.
 * Returns iterations (>= 0) or -2 (FP32 zero pivot -> FP64) / -3 (no convergence -> FP64);
 * *info = FP64 dgesv info on fallback, else 0. */
int orc_dsgesv(int n, int nrhs, const double* A, const double* B, double* X, double eps,
               int max_iter, int* info) {
  *info = 0;
  if (eps <= 0) eps = ldexp(1.0, -53);
  float* Af = (float*)malloc(sizeof(float) * (size_t)n * n);
  int32_t* ip = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  double* R = (double*)malloc(sizeof(double) * (size_t)n * nrhs);
  int ret = 0;
  double anrm = 0;
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < n; j++) s += fabs(A[(int64_t)i * n + j]);
    if (s > anrm) anrm = s;
  }
  const double cte = anrm * eps * sqrt((double)n);
  for (int64_t i = 0; i < (int64_t)n * n; i++) Af[i] = (float)A[i];
  if (orc_sgetrf(n, Af, n, ip) != 0) { ret = -2; goto fallback; }
  /* x = solve(b) in FP32 */
  {
    float* t = (float*)malloc(sizeof(float) * (size_t)n * nrhs);
    for (int64_t i = 0; i < (int64_t)n * nrhs; i++) t[i] = (float)B[i];
    /* FP32 getrs */
    for (int i = 0; i < n; i++) {
      const int p = ip[i] - 1;
      if (p != i) for (int c = 0; c < nrhs; c++) { float q = t[i * nrhs + c]; t[i * nrhs + c] = t[p * nrhs + c]; t[p * nrhs + c] = q; }
    }
    for (int c = 0; c < nrhs; c++) {
      for (int i = 0; i < n; i++) { float s = t[i * nrhs + c]; for (int k = 0; k < i; k++) s -= Af[(int64_t)i * n + k] * t[k * nrhs + c]; t[i * nrhs + c] = s; }
      for (int i = n - 1; i >= 0; i--) { float s = t[i * nrhs + c]; for (int k = i + 1; k < n; k++) s -= Af[(int64_t)i * n + k] * t[k * nrhs + c]; t[i * nrhs + c] = s / Af[(int64_t)i * n + i]; }
    }
    for (int64_t i = 0; i < (int64_t)n * nrhs; i++) X[i] = (double)t[i];
    for (int it = 0; it <= max_iter; it++) {
      /* r = b - A x */
      int ok = 1;
      for (int c = 0; c < nrhs; c++) {
        double xn = 0, rn = 0;
        for (int i = 0; i < n; i++) {
          double s = B[(int64_t)i * nrhs + c];
          for (int k = 0; k < n; k++) s -= A[(int64_t)i * n + k] * X[(int64_t)k * nrhs + c];
          R[(int64_t)i * nrhs + c] = s;
          if (fabs(s) > rn) rn = fabs(s);
          if (fabs(X[(int64_t)i * nrhs + c]) > xn) xn = fabs(X[(int64_t)i * nrhs + c]);
        }
        if (!(rn <= xn * cte)) ok = 0;
      }
      if (ok) { ret = it; free(t); goto done; }
      if (it == max_iter) break;
      for (int64_t i = 0; i < (int64_t)n * nrhs; i++) t[i] = (float)R[i];
      for (int i = 0; i < n; i++) {
        const int p = ip[i] - 1;
        if (p != i) for (int c = 0; c < nrhs; c++) { float q = t[i * nrhs + c]; t[i * nrhs + c] = t[p * nrhs + c]; t[p * nrhs + c] = q; }
      }
      for (int c = 0; c < nrhs; c++) {
        for (int i = 0; i < n; i++) { float s = t[i * nrhs + c]; for (int k = 0; k < i; k++) s -= Af[(int64_t)i * n + k] * t[k * nrhs + c]; t[i * nrhs + c] = s; }
        for (int i = n - 1; i >= 0; i--) { float s = t[i * nrhs + c]; for (int k = i + 1; k < n; k++) s -= Af[(int64_t)i * n + k] * t[k * nrhs + c]; t[i * nrhs + c] = s / Af[(int64_t)i * n + i]; }
      }
      for (int64_t i = 0; i < (int64_t)n * nrhs; i++) X[i] += (double)t[i];
    }
    free(t);
    ret = -3;
  }
fallback : {
  double* Ad = (double*)malloc(sizeof(double) * (size_t)n * n);
  memcpy(Ad, A, sizeof(double) * (size_t)n * n);
  memcpy(X, B, sizeof(double) * (size_t)n * nrhs);
  *info = orc_gesv(n, nrhs, Ad, n, ip, X, nrhs);
  free(Ad);
}
done:
  free(Af);
  free(ip);
  free(R);
  return ret;
}

/* Normwise backward error of X for A X = B, per column maximum:
 * max_c ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf + ||b_c||_inf), residual in long double. */
double orc_backward_error(int n, int nrhs, const double* A, const double* B, const double* X) {
  double anrm = 0, worst = 0;
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int j = 0; j < n; j++) s += fabs(A[(int64_t)i * n + j]);
    if (s > anrm) anrm = s;
  }
  for (int c = 0; c < nrhs; c++) {
    double rn = 0, xn = 0, bn = 0;
    for (int i = 0; i < n; i++) {
      long double s = B[(int64_t)i * nrhs + c];
      for (int k = 0; k < n; k++) s -= (long double)A[(int64_t)i * n + k] * X[(int64_t)k * nrhs + c];
      if (fabsl(s) > rn) rn = (double)fabsl(s);
      if (fabs(X[(int64_t)i * nrhs + c]) > xn) xn = fabs(X[(int64_t)i * nrhs + c]);
      if (fabs(B[(int64_t)i * nrhs + c]) > bn) bn = fabs(B[(int64_t)i * nrhs + c]);
    }
    const double den = anrm * xn + bn;
    const double be = den > 0 ? rn / den : rn;
    if (be > worst) worst = be;
  }
  return worst;
}
