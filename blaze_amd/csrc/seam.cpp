// The reference's BLAS/LAPACK seam served by the device library: host-memory cblas_dgemm,
// cblas_dgemv, their complex forms cblas_zgemm / cblas_cgemm / cblas_zgemv / cblas_cgemv, and
// dgesv_, dgetrf_ and dgetrs_ with the standard (LP64, 32-bit int) prototypes.
//
// The reference reaches external BLAS/LAPACK through these calls only:
//   cblas_dgemm  blaze/math/blas/cblas/gemm.h:156-162 (DMatDMatMultExpr's BLAS kernel)
//   cblas_dgemv  blaze/math/blas/cblas/gemv.h:153     (DMatDVecMultExpr's BLAS kernel)
//   cblas_zgemm / cblas_cgemm  blaze/math/blas/cblas/gemm.h:197-283 (complex elements)
//   cblas_zgemv / cblas_cgemv  blaze/math/blas/cblas/gemv.h:186-245
//   dgesv_       blaze/math/lapack/clapack/gesv.h:200-208 (solve() for N >= 7)
//   dgetrf_      blaze/math/lapack/clapack/getrf.h:190-197 (lu(), det(), invert())
//   dgetrs_      blaze/math/lapack/clapack/getrs.h:207-220
// Linking this library ahead of the system BLAS/LAPACK routes them to the GPU: operands are
// copied to HBM, the product or solve runs there (bzg_dgemm_ex, bzg_dgemv[_t], bzg_dgetrf,
// bzg_dgetrs), results are copied back before the call returns.  Column-major data is the
// row-major transpose: a column-major X (ld) is handed to the device as the row-major X^T
// (same ld), and LAPACK's column-major A is transposed on the device so that the factors and
// pivots are LAPACK's for A.  Unsupported forms (dgetrf_ with m != n, dgetrs_ with a
// transpose) return info = -1 / -2 instead of computing anything on the CPU.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "blaze_gpu.h"

namespace {

enum { kRowMajor = 101, kColMajor = 102 };
enum { kNoTrans = 111, kTrans = 112, kConjTrans = 113 };

struct SeamContext {
  bzg_context* ctx = nullptr;
  void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t bytes[4] = {0, 0, 0, 0};
  SeamContext() {
    if (bzg_create(&ctx) != BZG_OK) ctx = nullptr;
  }
  ~SeamContext() {
    for (int i = 0; i < 4; ++i)
      if (buf[i]) bzg_free(buf[i]);
    if (ctx) bzg_destroy(ctx);
  }
  // device buffer `slot` of at least `n` bytes (grown on demand, kept for later calls)
  void* get(int slot, size_t n) {
    if (n == 0) n = 8;
    if (bytes[slot] < n) {
      if (buf[slot]) bzg_free(buf[slot]);
      buf[slot] = nullptr;
      bytes[slot] = 0;
      if (bzg_malloc(&buf[slot], n) != BZG_OK) return nullptr;
      bytes[slot] = n;
    }
    return buf[slot];
  }
};

SeamContext& seam() {
  static thread_local SeamContext s;
  return s;
}

// the BLAS seam has no status channel: a device failure is reported and aborts, as a failing
// reference BLAS (xerbla) does
void fail(const char* what, bzg_status st) {
  std::fprintf(stderr, "blaze seam: %s failed: %s (%s)\n", what, bzg_status_string(st), bzg_last_error());
  std::abort();
}
#define SEAM_CHECK(expr, what)             \
  do {                                     \
    const bzg_status st_ = (expr);         \
    if (st_ != BZG_OK) fail(what, st_);    \
  } while (0)

// elements spanned by `rows` rows of `cols` entries at leading dimension ld
size_t extent(int64_t rows, int64_t cols, int64_t ld) {
  return rows > 0 && cols > 0 ? (size_t)((rows - 1) * ld + cols) : 0;
}

// a strided BLAS vector (n entries, increment inc; a negative inc starts at the end)
void gather(int n, const double* x, int inc, std::vector<double>& out) {
  out.resize((size_t)n);
  const int64_t step = inc < 0 ? -inc : inc;
  for (int i = 0; i < n; ++i) out[(size_t)i] = x[(inc < 0 ? (int64_t)(n - 1 - i) : (int64_t)i) * step];
}
void scatter(int n, const std::vector<double>& v, double* y, int inc) {
  const int64_t step = inc < 0 ? -inc : inc;
  for (int i = 0; i < n; ++i) y[(inc < 0 ? (int64_t)(n - 1 - i) : (int64_t)i) * step] = v[(size_t)i];
}

// complex vectors (n entries of 2 reals), optionally conjugated on the way
template <typename R>
void cgather(int n, const R* x, int inc, bool conj, std::vector<R>& out) {
  out.resize(2 * (size_t)n);
  const int64_t step = inc < 0 ? -inc : inc;
  for (int i = 0; i < n; ++i) {
    const int64_t e = (inc < 0 ? (int64_t)(n - 1 - i) : (int64_t)i) * step;
    out[2 * (size_t)i] = x[2 * e];
    out[2 * (size_t)i + 1] = conj ? -x[2 * e + 1] : x[2 * e + 1];
  }
}
template <typename R>
void cscatter(int n, const std::vector<R>& v, R* y, int inc, bool conj) {
  const int64_t step = inc < 0 ? -inc : inc;
  for (int i = 0; i < n; ++i) {
    const int64_t e = (inc < 0 ? (int64_t)(n - 1 - i) : (int64_t)i) * step;
    y[2 * e] = v[2 * (size_t)i];
    y[2 * e + 1] = conj ? -v[2 * (size_t)i + 1] : v[2 * (size_t)i + 1];
  }
}

// CBLAS transpose codes to the ABI's op codes (0 N, 1 T, 2 C); -1: not a CBLAS code
int op_code(int t) { return t == kNoTrans ? 0 : t == kTrans ? 1 : t == kConjTrans ? 2 : -1; }

template <typename R>
struct CxAbi;
template <>
struct CxAbi<double> {
  static bzg_status gemm(bzg_context* c, int ta, int tb, int64_t m, int64_t n, int64_t k, const double* al,
                         const double* A, int64_t lda, const double* B, int64_t ldb, const double* be, double* C,
                         int64_t ldc) {
    return bzg_zgemm_ex(c, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  }
  static bzg_status gemv(bzg_context* c, int t, int64_t m, int64_t n, const double* al, const double* A,
                         int64_t lda, const double* x, const double* be, double* y) {
    return bzg_zgemv_ex(c, t, m, n, al, A, lda, x, be, y);
  }
};
template <>
struct CxAbi<float> {
  static bzg_status gemm(bzg_context* c, int ta, int tb, int64_t m, int64_t n, int64_t k, const float* al,
                         const float* A, int64_t lda, const float* B, int64_t ldb, const float* be, float* C,
                         int64_t ldc) {
    return bzg_cgemm_ex(c, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
  }
  static bzg_status gemv(bzg_context* c, int t, int64_t m, int64_t n, const float* al, const float* A,
                         int64_t lda, const float* x, const float* be, float* y) {
    return bzg_cgemv_ex(c, t, m, n, al, A, lda, x, be, y);
  }
};

// Complex CBLAS gemm: the same row-major mapping as cblas_dgemm.  A column-major operand X is
// the row-major X^T, and (op(A) op(B))^T = op(B)^T op(A)^T keeps each operand's op code: for
// op = C, op(B)^T = conj(B) = (B^T)^H, the conjugate transpose of the stored row-major matrix.
template <typename R>
void cblas_cx_gemm(const char* name, int order, int transa, int transb, int m, int n, int k, const void* alpha,
                   const void* A, int lda, const void* B, int ldb, const void* beta, void* C, int ldc) {
  if (m <= 0 || n <= 0) return;
  SeamContext& s = seam();
  if (!s.ctx) fail(name, BZG_ERR_NO_DEVICE);
  const int ta = op_code(transa), tb = op_code(transb);
  if (ta < 0 || tb < 0) fail(name, BZG_ERR_INVALID_ARGUMENT);
  const bool row = order == kRowMajor;
  const int rm = row ? m : n, rn = row ? n : m;
  const R* X = static_cast<const R*>(row ? A : B);
  const R* Y = static_cast<const R*>(row ? B : A);
  const int tx = row ? ta : tb, ty = row ? tb : ta;
  const int ldx = row ? lda : ldb, ldy = row ? ldb : lda;
  const size_t ex = k > 0 ? (tx ? extent(k, rm, ldx) : extent(rm, k, ldx)) : 0;
  const size_t ey = k > 0 ? (ty ? extent(rn, k, ldy) : extent(k, rn, ldy)) : 0;
  const size_t ec = extent(rm, rn, ldc);
  const size_t es = 2 * sizeof(R);
  R* dX = static_cast<R*>(s.get(0, ex * es));
  R* dY = static_cast<R*>(s.get(1, ey * es));
  R* dC = static_cast<R*>(s.get(2, ec * es));
  if (!dX || !dY || !dC) fail(name, BZG_ERR_OUT_OF_MEMORY);
  const R* al = static_cast<const R*>(alpha);
  const R* be = static_cast<const R*>(beta);
  if (ex) SEAM_CHECK(bzg_memcpy(s.ctx, dX, X, ex * es, 0), name);
  if (ey) SEAM_CHECK(bzg_memcpy(s.ctx, dY, Y, ey * es, 0), name);
  if (be[0] != R(0) || be[1] != R(0) || ldc > rn) SEAM_CHECK(bzg_memcpy(s.ctx, dC, C, ec * es, 0), name);
  SEAM_CHECK(CxAbi<R>::gemm(s.ctx, tx, ty, rm, rn, k, al, dX, ldx, dY, ldy, be, dC, ldc), name);
  SEAM_CHECK(bzg_memcpy(s.ctx, C, dC, ec * es, 1), name);
}

// Complex CBLAS gemv.  Row-major: the stored A with the given op.  Column-major: the stored
// row-major matrix S = A^T, so A x = S^T x (op T), A^T x = S x (op N), and A^H x = conj(S) x,
// computed as conj(conj(alpha) S conj(x) + conj(beta) conj(y)) (conjugations on the host copies).
template <typename R>
void cblas_cx_gemv(const char* name, int order, int trans, int m, int n, const void* alpha, const void* A, int lda,
                   const void* x, int incx, const void* beta, void* y, int incy) {
  if (m <= 0 || n <= 0) return;
  SeamContext& s = seam();
  if (!s.ctx) fail(name, BZG_ERR_NO_DEVICE);
  const int t = op_code(trans);
  if (t < 0) fail(name, BZG_ERR_INVALID_ARGUMENT);
  const bool row = order == kRowMajor;
  const int sm = row ? m : n, sn = row ? n : m;
  const bool conj_all = !row && t == 2;
  const int dev_op = row ? t : (t == 0 ? 1 : 0);
  const int lx = t ? m : n, ly = t ? n : m;
  const R* al = static_cast<const R*>(alpha);
  const R* be = static_cast<const R*>(beta);
  const R alc[2] = {al[0], conj_all ? -al[1] : al[1]}, bec[2] = {be[0], conj_all ? -be[1] : be[1]};
  const bool use_y = be[0] != R(0) || be[1] != R(0);
  std::vector<R> hx, hy;
  cgather(lx, static_cast<const R*>(x), incx, conj_all, hx);
  if (use_y) cgather(ly, static_cast<const R*>(y), incy, conj_all, hy);
  else hy.assign(2 * (size_t)ly, R(0));
  const size_t es = 2 * sizeof(R), ea = extent(sm, sn, lda);
  R* dA = static_cast<R*>(s.get(0, ea * es));
  R* dx = static_cast<R*>(s.get(1, (size_t)lx * es));
  R* dy = static_cast<R*>(s.get(2, (size_t)ly * es));
  if (!dA || !dx || !dy) fail(name, BZG_ERR_OUT_OF_MEMORY);
  SEAM_CHECK(bzg_memcpy(s.ctx, dA, A, ea * es, 0), name);
  SEAM_CHECK(bzg_memcpy(s.ctx, dx, hx.data(), (size_t)lx * es, 0), name);
  if (use_y) SEAM_CHECK(bzg_memcpy(s.ctx, dy, hy.data(), (size_t)ly * es, 0), name);
  SEAM_CHECK(CxAbi<R>::gemv(s.ctx, dev_op, sm, sn, alc, dA, lda, dx, bec, dy), name);
  SEAM_CHECK(bzg_memcpy(s.ctx, hy.data(), dy, (size_t)ly * es, 1), name);
  cscatter(ly, hy, static_cast<R*>(y), incy, conj_all);
}

}  // namespace

extern "C" {

void cblas_zgemm(int order, int transa, int transb, int m, int n, int k, const void* alpha, const void* A, int lda,
                 const void* B, int ldb, const void* beta, void* C, int ldc) {
  cblas_cx_gemm<double>("cblas_zgemm", order, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
void cblas_cgemm(int order, int transa, int transb, int m, int n, int k, const void* alpha, const void* A, int lda,
                 const void* B, int ldb, const void* beta, void* C, int ldc) {
  cblas_cx_gemm<float>("cblas_cgemm", order, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
void cblas_zgemv(int order, int trans, int m, int n, const void* alpha, const void* A, int lda, const void* x,
                 int incx, const void* beta, void* y, int incy) {
  cblas_cx_gemv<double>("cblas_zgemv", order, trans, m, n, alpha, A, lda, x, incx, beta, y, incy);
}
void cblas_cgemv(int order, int trans, int m, int n, const void* alpha, const void* A, int lda, const void* x,
                 int incx, const void* beta, void* y, int incy) {
  cblas_cx_gemv<float>("cblas_cgemv", order, trans, m, n, alpha, A, lda, x, incx, beta, y, incy);
}

// C = alpha op(A) op(B) + beta C (CBLAS).  Row-major: the device call as given.  Column-major:
// C^T = alpha op(B)^T op(A)^T + beta C^T on the row-major transposes, i.e. the same call with
// the operands (and m, n) exchanged.
void cblas_dgemm(int order, int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
                 const double* B, int ldb, double beta, double* C, int ldc) {
  if (m <= 0 || n <= 0) return;
  SeamContext& s = seam();
  if (!s.ctx) fail("cblas_dgemm: context", BZG_ERR_NO_DEVICE);
  const int ta = transa != kNoTrans, tb = transb != kNoTrans;
  // row-major view: C is rm x rn, first operand X (trans tx), second Y (trans ty)
  const bool row = order == kRowMajor;
  const int rm = row ? m : n, rn = row ? n : m;
  const double* X = row ? A : B;
  const double* Y = row ? B : A;
  const int tx = row ? ta : tb, ty = row ? tb : ta;
  const int ldx = row ? lda : ldb, ldy = row ? ldb : lda;
  // stored shapes: X is rm x k (tx: k x rm), Y is k x rn (ty: rn x k)
  const size_t ex = k > 0 ? (tx ? extent(k, rm, ldx) : extent(rm, k, ldx)) : 0;
  const size_t ey = k > 0 ? (ty ? extent(rn, k, ldy) : extent(k, rn, ldy)) : 0;
  const size_t ec = extent(rm, rn, ldc);
  double* dX = static_cast<double*>(s.get(0, ex * sizeof(double)));
  double* dY = static_cast<double*>(s.get(1, ey * sizeof(double)));
  double* dC = static_cast<double*>(s.get(2, ec * sizeof(double)));
  if (!dX || !dY || !dC) fail("cblas_dgemm: device memory", BZG_ERR_OUT_OF_MEMORY);
  if (ex) SEAM_CHECK(bzg_memcpy(s.ctx, dX, X, ex * sizeof(double), 0), "cblas_dgemm: upload");
  if (ey) SEAM_CHECK(bzg_memcpy(s.ctx, dY, Y, ey * sizeof(double), 0), "cblas_dgemm: upload");
  // C is read only for beta != 0, but the download covers the padding between rows too: with
  // padding, C goes up (and its padding comes back unchanged) either way
  if (beta != 0.0 || ldc > rn) SEAM_CHECK(bzg_memcpy(s.ctx, dC, C, ec * sizeof(double), 0), "cblas_dgemm: upload");
  SEAM_CHECK(bzg_dgemm_ex(s.ctx, tx, ty, rm, rn, k, alpha, dX, ldx, dY, ldy, beta, dC, ldc), "cblas_dgemm");
  SEAM_CHECK(bzg_memcpy(s.ctx, C, dC, ec * sizeof(double), 1), "cblas_dgemm: download");
}

// y = alpha op(A) x + beta y (CBLAS).  A column-major m x n is the row-major n x m A^T, so its
// product without a transpose is the row-major transposed product and vice versa.
void cblas_dgemv(int order, int trans, int m, int n, double alpha, const double* A, int lda, const double* x,
                 int incx, double beta, double* y, int incy) {
  if (m <= 0 || n <= 0) return;
  SeamContext& s = seam();
  if (!s.ctx) fail("cblas_dgemv: context", BZG_ERR_NO_DEVICE);
  const bool t = trans != kNoTrans;
  const bool row = order == kRowMajor;
  const int sm = row ? m : n, sn = row ? n : m;  // stored row-major shape
  const bool dev_t = row ? t : !t;               // bzg_dgemv_t on the stored matrix
  const int lx = t ? m : n, ly = t ? n : m;      // vector lengths
  std::vector<double> hx, hy;
  gather(lx, x, incx, hx);
  if (beta != 0.0) gather(ly, y, incy, hy);
  else hy.assign((size_t)ly, 0.0);
  const size_t ea = extent(sm, sn, lda);
  double* dA = static_cast<double*>(s.get(0, ea * sizeof(double)));
  double* dx = static_cast<double*>(s.get(1, (size_t)lx * sizeof(double)));
  double* dy = static_cast<double*>(s.get(2, (size_t)ly * sizeof(double)));
  if (!dA || !dx || !dy) fail("cblas_dgemv: device memory", BZG_ERR_OUT_OF_MEMORY);
  SEAM_CHECK(bzg_memcpy(s.ctx, dA, A, ea * sizeof(double), 0), "cblas_dgemv: upload");
  SEAM_CHECK(bzg_memcpy(s.ctx, dx, hx.data(), (size_t)lx * sizeof(double), 0), "cblas_dgemv: upload");
  if (beta != 0.0) SEAM_CHECK(bzg_memcpy(s.ctx, dy, hy.data(), (size_t)ly * sizeof(double), 0), "cblas_dgemv: upload");
  if (dev_t)
    SEAM_CHECK(bzg_dgemv_t(s.ctx, sm, sn, alpha, dA, lda, dx, beta, dy), "cblas_dgemv");
  else
    SEAM_CHECK(bzg_dgemv(s.ctx, sm, sn, alpha, dA, lda, dx, beta, dy), "cblas_dgemv");
  SEAM_CHECK(bzg_memcpy(s.ctx, hy.data(), dy, (size_t)ly * sizeof(double), 1), "cblas_dgemv: download");
  scatter(ly, hy, y, incy);
}

// LAPACK dgetrf on the column-major n x n A (square only): the device factors the row-major
// copy of A (a transpose of the uploaded storage) and the factors are transposed back.
void dgetrf_(const int* m, const int* n, double* A, const int* lda, int* ipiv, int* info) {
  *info = 0;
  if (*m < 0) { *info = -1; return; }
  if (*n < 0 || *n != *m) { *info = -2; return; }  // rectangular factorisations are not provided
  if (*lda < (*m > 1 ? *m : 1)) { *info = -4; return; }
  const int N = *n, ld = *lda;
  if (N == 0) return;
  SeamContext& s = seam();
  if (!s.ctx) fail("dgetrf_: context", BZG_ERR_NO_DEVICE);
  const size_t ea = extent(N, N, ld);
  double* dS = static_cast<double*>(s.get(0, ea * sizeof(double)));          // storage (= A^T)
  double* dA = static_cast<double*>(s.get(1, (size_t)N * N * sizeof(double)));  // row-major A
  int32_t* dp = static_cast<int32_t*>(s.get(2, (size_t)N * sizeof(int32_t) + 64));
  if (!dS || !dA || !dp) fail("dgetrf_: device memory", BZG_ERR_OUT_OF_MEMORY);
  int32_t* dinfo = dp + N + (16 - N % 16) % 16;
  SEAM_CHECK(bzg_memcpy(s.ctx, dS, A, ea * sizeof(double), 0), "dgetrf_: upload");
  SEAM_CHECK(bzg_dtranspose(s.ctx, N, N, dS, ld, dA, N), "dgetrf_: transpose");
  SEAM_CHECK(bzg_dgetrf(s.ctx, N, dA, N, dp, dinfo), "dgetrf_");
  SEAM_CHECK(bzg_dtranspose(s.ctx, N, N, dA, N, dS, ld), "dgetrf_: transpose");
  SEAM_CHECK(bzg_memcpy(s.ctx, A, dS, ea * sizeof(double), 1), "dgetrf_: download");
  std::vector<int32_t> hp((size_t)N);
  int32_t hinfo = 0;
  SEAM_CHECK(bzg_memcpy(s.ctx, hp.data(), dp, (size_t)N * sizeof(int32_t), 1), "dgetrf_: download");
  SEAM_CHECK(bzg_memcpy(s.ctx, &hinfo, dinfo, sizeof(int32_t), 1), "dgetrf_: download");
  for (int i = 0; i < N; ++i) ipiv[i] = hp[(size_t)i];
  *info = hinfo;
}

// LAPACK dgetrs ('N') from dgetrf_'s column-major factors: B (column-major n x nrhs) is the
// row-major B^T; it is transposed to the row-major B on the device, solved, and transposed back.
void dgetrs_(const char* trans, const int* n, const int* nrhs, const double* A, const int* lda, const int* ipiv,
             double* B, const int* ldb, int* info, size_t /*trans_len*/) {
  *info = 0;
  if (!(*trans == 'N' || *trans == 'n')) { *info = -1; return; }  // transposed solves are not provided
  if (*n < 0) { *info = -2; return; }
  if (*nrhs < 0) { *info = -3; return; }
  if (*lda < (*n > 1 ? *n : 1)) { *info = -5; return; }
  if (*ldb < (*n > 1 ? *n : 1)) { *info = -8; return; }
  const int N = *n, R = *nrhs;
  if (N == 0 || R == 0) return;
  SeamContext& s = seam();
  if (!s.ctx) fail("dgetrs_: context", BZG_ERR_NO_DEVICE);
  const size_t ea = extent(N, N, *lda), eb = extent(R, N, *ldb);
  double* dS = static_cast<double*>(s.get(0, (ea + (size_t)N * N) * sizeof(double)));
  double* dA = dS + ea;
  double* dT = static_cast<double*>(s.get(1, (eb + (size_t)N * R) * sizeof(double)));
  double* dB = dT + eb;
  int32_t* dp = static_cast<int32_t*>(s.get(2, (size_t)N * sizeof(int32_t)));
  if (!dS || !dT || !dp) fail("dgetrs_: device memory", BZG_ERR_OUT_OF_MEMORY);
  std::vector<int32_t> hp(ipiv, ipiv + N);
  SEAM_CHECK(bzg_memcpy(s.ctx, dS, A, ea * sizeof(double), 0), "dgetrs_: upload");
  SEAM_CHECK(bzg_memcpy(s.ctx, dT, B, eb * sizeof(double), 0), "dgetrs_: upload");
  SEAM_CHECK(bzg_memcpy(s.ctx, dp, hp.data(), (size_t)N * sizeof(int32_t), 0), "dgetrs_: upload");
  SEAM_CHECK(bzg_dtranspose(s.ctx, N, N, dS, *lda, dA, N), "dgetrs_: transpose");
  SEAM_CHECK(bzg_dtranspose(s.ctx, R, N, dT, *ldb, dB, R), "dgetrs_: transpose");
  SEAM_CHECK(bzg_dgetrs(s.ctx, N, R, dA, N, dp, dB, R), "dgetrs_");
  SEAM_CHECK(bzg_dtranspose(s.ctx, N, R, dB, R, dT, *ldb), "dgetrs_: transpose");
  SEAM_CHECK(bzg_memcpy(s.ctx, B, dT, eb * sizeof(double), 1), "dgetrs_: download");
}

// LAPACK dgesv = dgetrf_ + dgetrs_ ('N'); B is left unchanged when info > 0.
void dgesv_(const int* n, const int* nrhs, double* A, const int* lda, int* ipiv, double* B, const int* ldb,
            int* info) {
  *info = 0;
  if (*n < 0) { *info = -1; return; }
  if (*nrhs < 0) { *info = -2; return; }
  if (*lda < (*n > 1 ? *n : 1)) { *info = -4; return; }
  if (*ldb < (*n > 1 ? *n : 1)) { *info = -7; return; }
  dgetrf_(n, n, A, lda, ipiv, info);
  if (*info != 0) return;
  dgetrs_("N", n, nrhs, A, lda, ipiv, B, ldb, info, 1);
}

}  // extern "C"
