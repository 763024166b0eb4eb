// CPU ORACLE (compiled reference) -- test infrastructure only.
//
// Thin extern "C" driver around the reference's own header-only kernels: it copies plain
// arrays into blaze::DynamicMatrix / DynamicVector / CompressedMatrix and evaluates
// `C = A * B`, `y = A * x` and the closed-form small solves exactly as a Blaze user would.
// Built by oracle/Makefile from the headers under the reference tree; the binary goes to
// oracle/_ref/ and is never part of the product.  LAPACK-backed paths (solve() for N >= 7,
// lu()) need an external LAPACK and are not built here.
#include <blaze/Math.h>
#include <chrono>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>

using blaze::rowMajor;
using blaze::columnVector;

namespace {
template <typename T>
void to_blaze(blaze::DynamicMatrix<T, rowMajor>& M, int64_t m, int64_t n, const T* src) {
  M.resize(m, n, false);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) M(i, j) = src[i * n + j];
}
template <typename T>
void from_blaze(const blaze::DynamicMatrix<T, rowMajor>& M, T* dst) {
  for (size_t i = 0; i < M.rows(); ++i)
    for (size_t j = 0; j < M.columns(); ++j) dst[i * M.columns() + j] = M(i, j);
}
template <typename T>
void build_csr(blaze::CompressedMatrix<T, rowMajor>& S, int64_t n_rows, int64_t n_cols,
               const int32_t* rowptr, const int32_t* col, const T* val) {
  S.resize(n_rows, n_cols, false);
  S.reserve(rowptr[n_rows]);
  for (int64_t i = 0; i < n_rows; ++i) {
    for (int32_t t = rowptr[i]; t < rowptr[i + 1]; ++t) S.append(i, col[t], val[t], false);
    S.finalize(i);
  }
}
template <typename T>
void gemm(int64_t m, int64_t n, int64_t k, const T* A, const T* B, T* C) {
  blaze::DynamicMatrix<T, rowMajor> a, b, c;
  to_blaze(a, m, k, A);
  to_blaze(b, k, n, B);
  c = a * b;
  from_blaze(c, C);
}
template <typename T>
void gemv(int64_t m, int64_t n, const T* A, const T* x, T* y) {
  blaze::DynamicMatrix<T, rowMajor> a;
  to_blaze(a, m, n, A);
  blaze::DynamicVector<T, columnVector> xv(n), yv;
  for (int64_t j = 0; j < n; ++j) xv[j] = x[j];
  yv = a * xv;
  for (int64_t i = 0; i < m; ++i) y[i] = yv[i];
}
template <typename T>
void csrmv(int64_t n_rows, int64_t n_cols, const int32_t* rowptr, const int32_t* col, const T* val,
           const T* x, T* y) {
  blaze::CompressedMatrix<T, rowMajor> S;
  build_csr(S, n_rows, n_cols, rowptr, col, val);
  blaze::DynamicVector<T, columnVector> xv(n_cols), yv;
  for (int64_t j = 0; j < n_cols; ++j) xv[j] = x[j];
  yv = S * xv;
  for (int64_t i = 0; i < n_rows; ++i) y[i] = yv[i];
}
// y = trans(A) * x for a row-major A: the reference's TDMatDVecMultExpr
template <typename T>
void tgemv(int64_t m, int64_t n, const T* A, const T* x, T* y) {
  blaze::DynamicMatrix<T, rowMajor> a;
  to_blaze(a, m, n, A);
  blaze::DynamicVector<T, columnVector> xv(m), yv;
  for (int64_t i = 0; i < m; ++i) xv[i] = x[i];
  yv = trans(a) * xv;
  for (int64_t j = 0; j < n; ++j) y[j] = yv[j];
}
// y = trans(S) * x for a row-major S: the reference's TSMatDVecMultExpr (a column-major operand)
template <typename T>
void tcsrmv(int64_t n_rows, int64_t n_cols, const int32_t* rowptr, const int32_t* col, const T* val,
            const T* x, T* y) {
  blaze::CompressedMatrix<T, rowMajor> S;
  build_csr(S, n_rows, n_cols, rowptr, col, val);
  blaze::DynamicVector<T, columnVector> xv(n_rows), yv;
  for (int64_t i = 0; i < n_rows; ++i) xv[i] = x[i];
  yv = trans(S) * xv;
  for (int64_t j = 0; j < n_cols; ++j) y[j] = yv[j];
}
// C = op(A) * op(B) with op 0 = none, 1 = trans(), 2 = ctrans(); A stored m x k (op 0) or
// k x m, B stored k x n (op 0) or n x k.
template <typename T>
void gemm_op(int opa, int opb, int64_t m, int64_t n, int64_t k, const T* A, const T* B, T* C) {
  blaze::DynamicMatrix<T, rowMajor> a, b, c;
  to_blaze(a, opa ? k : m, opa ? m : k, A);
  to_blaze(b, opb ? n : k, opb ? k : n, B);
  switch (opa * 3 + opb) {
    case 0: c = a * b; break;
    case 1: c = a * trans(b); break;
    case 2: c = a * ctrans(b); break;
    case 3: c = trans(a) * b; break;
    case 4: c = trans(a) * trans(b); break;
    case 5: c = trans(a) * ctrans(b); break;
    case 6: c = ctrans(a) * b; break;
    case 7: c = ctrans(a) * trans(b); break;
    default: c = ctrans(a) * ctrans(b); break;
  }
  from_blaze(c, C);
}
// y = op(A) * x, A stored m x n
template <typename T>
void gemv_op(int op, int64_t m, int64_t n, const T* A, const T* x, T* y) {
  blaze::DynamicMatrix<T, rowMajor> a;
  to_blaze(a, m, n, A);
  const int64_t xl = op ? m : n;
  blaze::DynamicVector<T, columnVector> xv(xl), yv;
  for (int64_t j = 0; j < xl; ++j) xv[j] = x[j];
  if (op == 0) yv = a * xv;
  else if (op == 1) yv = trans(a) * xv;
  else yv = ctrans(a) * xv;
  for (size_t i = 0; i < yv.size(); ++i) y[i] = yv[i];
}
using zd = std::complex<double>;
using zf = std::complex<float>;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

extern "C" {

void ref_set_threads(int t) { blaze::setNumThreads(t); }
int ref_get_threads() { return (int)blaze::getNumThreads(); }

void ref_dgemm(int64_t m, int64_t n, int64_t k, const double* A, const double* B, double* C) { gemm(m, n, k, A, B, C); }
void ref_sgemm(int64_t m, int64_t n, int64_t k, const float* A, const float* B, float* C) { gemm(m, n, k, A, B, C); }
void ref_igemm(int64_t m, int64_t n, int64_t k, const int32_t* A, const int32_t* B, int32_t* C) { gemm(m, n, k, A, B, C); }

void ref_dgemv(int64_t m, int64_t n, const double* A, const double* x, double* y) { gemv(m, n, A, x, y); }
void ref_sgemv(int64_t m, int64_t n, const float* A, const float* x, float* y) { gemv(m, n, A, x, y); }
void ref_igemv(int64_t m, int64_t n, const int32_t* A, const int32_t* x, int32_t* y) { gemv(m, n, A, x, y); }

void ref_zgemm_op(int opa, int opb, int64_t m, int64_t n, int64_t k, const zd* A, const zd* B, zd* C) { gemm_op(opa, opb, m, n, k, A, B, C); }
void ref_cgemm_op(int opa, int opb, int64_t m, int64_t n, int64_t k, const zf* A, const zf* B, zf* C) { gemm_op(opa, opb, m, n, k, A, B, C); }
void ref_zgemv_op(int op, int64_t m, int64_t n, const zd* A, const zd* x, zd* y) { gemv_op(op, m, n, A, x, y); }
void ref_cgemv_op(int op, int64_t m, int64_t n, const zf* A, const zf* x, zf* y) { gemv_op(op, m, n, A, x, y); }

void ref_dcsrmv(int64_t r, int64_t c, const int32_t* rp, const int32_t* ci, const double* v, const double* x, double* y) { csrmv(r, c, rp, ci, v, x, y); }
void ref_scsrmv(int64_t r, int64_t c, const int32_t* rp, const int32_t* ci, const float* v, const float* x, float* y) { csrmv(r, c, rp, ci, v, x, y); }
void ref_icsrmv(int64_t r, int64_t c, const int32_t* rp, const int32_t* ci, const int32_t* v, const int32_t* x, int32_t* y) { csrmv(r, c, rp, ci, v, x, y); }
void ref_dtgemv(int64_t m, int64_t n, const double* A, const double* x, double* y) { tgemv(m, n, A, x, y); }
void ref_stgemv(int64_t m, int64_t n, const float* A, const float* x, float* y) { tgemv(m, n, A, x, y); }
void ref_itgemv(int64_t m, int64_t n, const int32_t* A, const int32_t* x, int32_t* y) { tgemv(m, n, A, x, y); }
void ref_dtcsrmv(int64_t r, int64_t c, const int32_t* rp, const int32_t* ci, const double* v, const double* x, double* y) { tcsrmv(r, c, rp, ci, v, x, y); }
void ref_stcsrmv(int64_t r, int64_t c, const int32_t* rp, const int32_t* ci, const float* v, const float* x, float* y) { tcsrmv(r, c, rp, ci, v, x, y); }

// Closed-form solves of LSE.h for 1 <= n <= 6.  Returns 0, or 1 when Blaze throws
// std::runtime_error("Solving LSE with singular system matrix failed").
int ref_dsolve_small(int n, const double* A, const double* b, double* x) {
  blaze::DynamicMatrix<double, rowMajor> a;
  to_blaze(a, n, n, A);
  blaze::DynamicVector<double, columnVector> bv(n), xv(n);
  for (int i = 0; i < n; ++i) bv[i] = b[i];
  try {
    switch (n) {
      case 1: blaze::solve1x1(a, xv, bv); break;
      case 2: blaze::solve2x2(a, xv, bv); break;
      case 3: blaze::solve3x3(a, xv, bv); break;
      case 4: blaze::solve4x4(a, xv, bv); break;
      case 5: blaze::solve5x5(a, xv, bv); break;
      case 6: blaze::solve6x6(a, xv, bv); break;
      default: return -1;
    }
  } catch (const std::runtime_error&) {
    return 1;
  }
  for (int i = 0; i < n; ++i) x[i] = xv[i];
  return 0;
}

// blazemark protocol (Config.h:58-59, src/blaze/DMatDMatMult.cpp:81-87): one warm-up
// assignment, then `reps` timed assignments; returns the minimum seconds per assignment.
double ref_time_dgemm(int64_t m, int64_t n, int64_t k, const double* A, const double* B, int reps) {
  blaze::DynamicMatrix<double, rowMajor> a, b, c(m, n);
  to_blaze(a, m, k, A);
  to_blaze(b, k, n, B);
  c = a * b;
  double best = 1e300;
  for (int r = 0; r < reps; ++r) {
    const double t0 = now();
    c = a * b;
    const double t = now() - t0;
    if (t < best) best = t;
  }
  return best;
}

// best of `reps` timed C = A * B for complex<double> (after one untimed evaluation)
double ref_time_zgemm(int64_t m, int64_t n, int64_t k, const zd* A, const zd* B, int reps) {
  blaze::DynamicMatrix<zd, rowMajor> a, b, c(m, n);
  to_blaze(a, m, k, A);
  to_blaze(b, k, n, B);
  c = a * b;
  double best = 1e300;
  for (int r = 0; r < reps; ++r) {
    const double t0 = now();
    c = a * b;
    const double t = now() - t0;
    if (t < best) best = t;
  }
  return best;
}

double ref_time_dcsrmv(int64_t nr, int64_t nc, const int32_t* rp, const int32_t* ci, const double* v,
                       const double* x, int reps) {
  blaze::CompressedMatrix<double, rowMajor> S;
  build_csr(S, nr, nc, rp, ci, v);
  blaze::DynamicVector<double, columnVector> xv(nc), yv(nr);
  for (int64_t j = 0; j < nc; ++j) xv[j] = x[j];
  yv = S * xv;
  double best = 1e300;
  for (int r = 0; r < reps; ++r) {
    const double t0 = now();
    yv = S * xv;
    const double t = now() - t0;
    if (t < best) best = t;
  }
  return best;
}

double ref_time_dgemv(int64_t m, int64_t n, const double* A, const double* x, int reps) {
  blaze::DynamicMatrix<double, rowMajor> a;
  to_blaze(a, m, n, A);
  blaze::DynamicVector<double, columnVector> xv(n), yv(m);
  for (int64_t j = 0; j < n; ++j) xv[j] = x[j];
  yv = a * xv;
  double best = 1e300;
  for (int r = 0; r < reps; ++r) {
    const double t0 = now();
    yv = a * xv;
    const double t = now() - t0;
    if (t < best) best = t;
  }
  return best;
}

}  // extern "C"
