// Host stand-in for the C ABI functions the BLAS/LAPACK seam calls (test infrastructure only):
// "device" memory is host memory, each call computes its documented row-major result with
// plain loops.  Linked with blaze_amd/csrc/seam.cpp into a test library, it lets the CPU
// suite check the seam's mapping of CBLAS/LAPACK arguments (storage orders, transposes,
// leading dimensions, strides, info codes) onto the ABI without a GPU.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "blaze_gpu.h"

struct bzg_context { int dummy; };

template <typename R>
static void cx_gemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const R* al, const R* A, int64_t lda,
                    const R* B, int64_t ldb, const R* be, R* C, int64_t ldc) {
  using Z = std::complex<R>;
  auto a = [&](int64_t i, int64_t p) {
    const R* e = ta ? A + 2 * (p * lda + i) : A + 2 * (i * lda + p);
    return Z(e[0], ta == 2 ? -e[1] : e[1]);
  };
  auto b = [&](int64_t p, int64_t j) {
    const R* e = tb ? B + 2 * (j * ldb + p) : B + 2 * (p * ldb + j);
    return Z(e[0], tb == 2 ? -e[1] : e[1]);
  };
  const Z alpha(al[0], al[1]), beta(be[0], be[1]);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) {
      Z s(0, 0);
      for (int64_t p = 0; p < k; ++p) s += a(i, p) * b(p, j);
      R* c = C + 2 * (i * ldc + j);
      const Z v = alpha * s + (beta != Z(0, 0) ? beta * Z(c[0], c[1]) : Z(0, 0));
      c[0] = v.real();
      c[1] = v.imag();
    }
}
template <typename R>
static void cx_gemv(int t, int64_t m, int64_t n, const R* al, const R* A, int64_t lda, const R* x, const R* be, R* y) {
  using Z = std::complex<R>;
  const Z alpha(al[0], al[1]), beta(be[0], be[1]);
  const int64_t L = t ? n : m, K = t ? m : n;
  for (int64_t o = 0; o < L; ++o) {
    Z s(0, 0);
    for (int64_t q = 0; q < K; ++q) {
      const R* e = t ? A + 2 * (q * lda + o) : A + 2 * (o * lda + q);
      s += Z(e[0], t == 2 ? -e[1] : e[1]) * Z(x[2 * q], x[2 * q + 1]);
    }
    const Z v = alpha * s + (beta != Z(0, 0) ? beta * Z(y[2 * o], y[2 * o + 1]) : Z(0, 0));
    y[2 * o] = v.real();
    y[2 * o + 1] = v.imag();
  }
}
extern "C" {
const char* bzg_status_string(bzg_status) { return "stub"; }
const char* bzg_last_error(void) { return ""; }
bzg_status bzg_create(bzg_context** out) { *out = new bzg_context(); return BZG_OK; }
bzg_status bzg_destroy(bzg_context* c) { delete c; return BZG_OK; }
bzg_status bzg_malloc(void** p, size_t b) { *p = std::malloc(b ? b : 1); return BZG_OK; }
bzg_status bzg_free(void* p) { std::free(p); return BZG_OK; }
bzg_status bzg_memcpy(bzg_context*, void* d, const void* s, size_t b, int) { std::memcpy(d, s, b); return BZG_OK; }

bzg_status bzg_dgemm_ex(bzg_context*, int32_t ta, int32_t tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                        int64_t ldc) {
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) {
      double s = 0;
      for (int64_t p = 0; p < k; ++p) s += (ta ? A[p * lda + i] : A[i * lda + p]) * (tb ? B[j * ldb + p] : B[p * ldb + j]);
      C[i * ldc + j] = alpha * s + (beta != 0.0 ? beta * C[i * ldc + j] : 0.0);
    }
  return BZG_OK;
}
bzg_status bzg_dgemv(bzg_context*, int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                     const double* x, double beta, double* y) {
  for (int64_t i = 0; i < m; ++i) {
    double s = 0;
    for (int64_t j = 0; j < n; ++j) s += A[i * lda + j] * x[j];
    y[i] = alpha * s + (beta != 0.0 ? beta * y[i] : 0.0);
  }
  return BZG_OK;
}
bzg_status bzg_dgemv_t(bzg_context*, int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                       const double* x, double beta, double* y) {
  for (int64_t j = 0; j < n; ++j) {
    double s = 0;
    for (int64_t i = 0; i < m; ++i) s += A[i * lda + j] * x[i];
    y[j] = alpha * s + (beta != 0.0 ? beta * y[j] : 0.0);
  }
  return BZG_OK;
}
bzg_status bzg_zgemm_ex(bzg_context*, int32_t ta, int32_t tb, int64_t m, int64_t n, int64_t k, const double* al,
                        const double* A, int64_t lda, const double* B, int64_t ldb, const double* be, double* C,
                        int64_t ldc) {
  cx_gemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  return BZG_OK;
}
bzg_status bzg_cgemm_ex(bzg_context*, int32_t ta, int32_t tb, int64_t m, int64_t n, int64_t k, const float* al,
                        const float* A, int64_t lda, const float* B, int64_t ldb, const float* be, float* C,
                        int64_t ldc) {
  cx_gemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  return BZG_OK;
}
bzg_status bzg_zgemv_ex(bzg_context*, int32_t t, int64_t m, int64_t n, const double* al, const double* A,
                        int64_t lda, const double* x, const double* be, double* y) {
  cx_gemv(t, m, n, al, A, lda, x, be, y);
  return BZG_OK;
}
bzg_status bzg_cgemv_ex(bzg_context*, int32_t t, int64_t m, int64_t n, const float* al, const float* A,
                        int64_t lda, const float* x, const float* be, float* y) {
  cx_gemv(t, m, n, al, A, lda, x, be, y);
  return BZG_OK;
}
bzg_status bzg_dtranspose(bzg_context*, int64_t m, int64_t n, const double* A, int64_t lda, double* B, int64_t ldb) {
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) B[j * ldb + i] = A[i * lda + j];
  return BZG_OK;
}
bzg_status bzg_dgetrf(bzg_context*, int64_t n, double* A, int64_t lda, int32_t* ipiv, int32_t* info) {
  *info = 0;
  for (int64_t j = 0; j < n; ++j) {
    int64_t p = j;
    for (int64_t i = j + 1; i < n; ++i)
      if (std::fabs(A[i * lda + j]) > std::fabs(A[p * lda + j])) p = i;
    ipiv[j] = (int32_t)(p + 1);
    if (A[p * lda + j] != 0.0) {
      for (int64_t c = 0; c < n; ++c) std::swap(A[j * lda + c], A[p * lda + c]);
      for (int64_t i = j + 1; i < n; ++i) A[i * lda + j] /= A[j * lda + j];
    } else if (*info == 0) {
      *info = (int32_t)(j + 1);
    }
    for (int64_t i = j + 1; i < n; ++i)
      for (int64_t c = j + 1; c < n; ++c) A[i * lda + c] -= A[i * lda + j] * A[j * lda + c];
  }
  return BZG_OK;
}
bzg_status bzg_dgetrs(bzg_context*, int64_t n, int64_t r, const double* A, int64_t lda, const int32_t* ipiv,
                      double* B, int64_t ldb) {
  for (int64_t i = 0; i < n; ++i)
    for (int64_t c = 0; c < r; ++c) std::swap(B[i * ldb + c], B[(ipiv[i] - 1) * ldb + c]);
  for (int64_t c = 0; c < r; ++c) {
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = 0; k < i; ++k) B[i * ldb + c] -= A[i * lda + k] * B[k * ldb + c];
    for (int64_t i = n - 1; i >= 0; --i) {
      for (int64_t k = i + 1; k < n; ++k) B[i * ldb + c] -= A[i * lda + k] * B[k * ldb + c];
      B[i * ldb + c] /= A[i * lda + i];
    }
  }
  return BZG_OK;
}
}
