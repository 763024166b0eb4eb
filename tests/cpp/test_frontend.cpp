// C++ front-end test: the reference's expression syntax on blaze::gpu containers.
//
// Mirrors the reference's own tests: the DynamicMatrix multiplication-assignment known answer
// (blazetest/src/mathtest/matrices/dynamicmatrix/ClassTest2.cpp:1244-1273), aliasing
// (operations/dmatdmatmult/AliasingTest.cpp:1803-1850: A = A * B equals the non-aliased
// result), CompressedMatrix products (matrices/compressedmatrix/ClassTest1.cpp), the LSE test
// (operations/lse/DenseTest.h:140-200: A*x == b for N = 1..12, relaxed equality 1e-8) and the
// exception types/messages of operator* and solve().  Every check prints PASS/FAIL; exit code
// is the number of failures.
#include <blaze/gpu/Math.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

using namespace blaze::gpu;

static int failures = 0;
static void check(bool ok, const std::string& name) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", name.c_str());
  std::fflush(stdout);
  if (!ok) ++failures;
}
template <typename E>
static bool throws(const std::function<void()>& f, const std::string& msg) {
  try {
    f();
  } catch (const E& e) {
    return msg.empty() || std::string(e.what()) == msg;
  } catch (...) {
    return false;
  }
  return false;
}
static bool relaxed(double a, double b, double eps = 1e-8) {  // blaze::equal (Equal.h:214-224)
  return std::fabs(a - b) <= eps * std::max(1.0, std::fabs(b));
}
template <typename T>
static std::vector<T> host_gemm(const std::vector<T>& A, const std::vector<T>& B, size_t m, size_t n, size_t k) {
  std::vector<T> C(m * n, T(0));
  for (size_t i = 0; i < m; ++i)
    for (size_t p = 0; p < k; ++p)
      for (size_t j = 0; j < n; ++j) C[i * n + j] += A[i * k + p] * B[p * n + j];
  return C;
}

int main() {
  std::mt19937_64 rng(128753984);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::uniform_int_distribution<int> I(-10, 10);

  // ClassTest2.cpp:1244-1273
  {
    DynamicMatrix<int> mat1{{0, 2, 0, 0}, {1, 3, 0, 4}, {0, 0, 0, 5}};
    DynamicMatrix<int> mat2{{1, 0, 2}, {0, 3, 0}, {4, 0, 5}};
    mat2 *= mat1;
    std::vector<int> got = mat2.toVector(), want{0, 2, 0, 10, 3, 9, 0, 12, 0, 8, 0, 25};
    check(mat2.rows() == 3 && mat2.columns() == 4 && got == want && mat2.nonZeros() == 7,
          "DynamicMatrix<int> multiplication assignment known answer");
  }
  // AliasingTest: A = A * B equals the non-aliased product
  {
    DynamicMatrix<int> A{{-1, 0, -2, 0}, {0, 2, -3, 1}, {0, 1, 2, 2}};
    DynamicMatrix<int> B{{1, 0, -3}, {0, -1, 0}, {0, 2, 1}, {2, 1, -2}};
    DynamicMatrix<int> C{{1, 0, 2}, {0, 3, -1}, {-1, 0, 2}};
    DynamicMatrix<int> ref;
    ref = A * B;
    DynamicMatrix<int> A2(A);
    A2 = A2 * B;  // aliased left operand
    check(A2.toVector() == ref.toVector(), "aliased C = C * B (int)");
    DynamicMatrix<int> ref2;
    ref2 = C * C;
    C = C * C;
    check(C.toVector() == ref2.toVector(), "aliased C = C * C (int)");
  }
  // random double / float products against a host triple loop
  for (size_t m : {1, 7, 37, 64, 130})
    for (size_t k : {0, 3, 63, 200})
      for (size_t n : {1, 16, 65}) {
        std::vector<double> a(m * k), b(k * n);
        for (auto& v : a) v = U(rng);
        for (auto& v : b) v = U(rng);
        DynamicMatrix<double> A(m, k, a.data()), B(k, n, b.data()), C(m, n, 5.0);
        C = A * B;
        auto want = host_gemm(a, b, m, n, k);
        auto got = C.toVector();
        bool ok = C.rows() == m && C.columns() == n;
        for (size_t q = 0; q < m * n && ok; ++q) ok = relaxed(got[q], want[q]);
        check(ok, "C = A * B double " + std::to_string(m) + "x" + std::to_string(n) + "x" + std::to_string(k));
        std::vector<float> af(a.begin(), a.end()), bf(b.begin(), b.end());
        DynamicMatrix<float> Af(m, k, af.data()), Bf(k, n, bf.data()), Cf;
        Cf = Af * Bf;
        auto gf = Cf.toVector();
        auto wf = host_gemm(af, bf, m, n, k);
        ok = true;
        for (size_t q = 0; q < m * n && ok; ++q) ok = std::fabs(gf[q] - wf[q]) <= 1e-5f * (1 + std::fabs(wf[q]));
        check(ok, "C = A * B float " + std::to_string(m) + "x" + std::to_string(n) + "x" + std::to_string(k));
      }
  // compound and scaled forms
  {
    const size_t m = 40, n = 30, k = 20;
    std::vector<double> a(m * k), b(k * n), c(m * n);
    for (auto& v : a) v = U(rng);
    for (auto& v : b) v = U(rng);
    for (auto& v : c) v = U(rng);
    DynamicMatrix<double> A(m, k, a.data()), B(k, n, b.data()), C(m, n, c.data()), D(m, n, c.data()), E;
    C += A * B;
    D -= A * B;
    E = 2.0 * (A * B);
    auto p = host_gemm(a, b, m, n, k);
    auto gc = C.toVector(), gd = D.toVector(), ge = E.toVector();
    bool ok = true;
    for (size_t q = 0; q < m * n; ++q)
      ok = ok && relaxed(gc[q], c[q] + p[q]) && relaxed(gd[q], c[q] - p[q]) && relaxed(ge[q], 2 * p[q]);
    check(ok, "C += A*B, C -= A*B, C = 2*(A*B)");
  }
  // K == 0 resets the target (DMatDMatMultExpr.h:527-530)
  {
    DynamicMatrix<double> A(3, 0), B(0, 4), C(3, 4, 7.0);
    C = A * B;
    bool ok = C.rows() == 3 && C.columns() == 4;
    for (double v : C.toVector()) ok = ok && v == 0.0;
    check(ok, "K == 0 product resets the target");
  }
  // exceptions of operator*
  {
    DynamicMatrix<double> A(3, 4), B(5, 2);
    DynamicVector<double> x(7);
    check(throws<std::invalid_argument>([&] { DynamicMatrix<double> C; C = A * B; }, "Matrix sizes do not match"),
          "A * B size mismatch throws invalid_argument");
    check(throws<std::invalid_argument>([&] { DynamicVector<double> y; y = A * x; },
                                        "Matrix and vector sizes do not match"),
          "A * x size mismatch throws invalid_argument");
  }
  // GEMV (MDaVDa.cpp shapes) and aliasing x = A * x
  for (auto mn : std::vector<std::pair<size_t, size_t>>{{0, 0}, {1, 1}, {6, 5}, {67, 127}, {127, 67}, {64, 128}}) {
    const size_t m = mn.first, n = mn.second;
    std::vector<int> a(m * n), x(n);
    for (auto& v : a) v = I(rng);
    for (auto& v : x) v = I(rng);
    DynamicMatrix<int> A(m, n, a.data());
    DynamicVector<int> X(n, x.data()), Y;
    Y = A * X;
    auto want = host_gemm(a, x, m, 1, n);
    check(Y.size() == m && Y.toVector() == want, "y = A * x int " + std::to_string(m) + "x" + std::to_string(n));
  }
  {
    const size_t n = 50;
    std::vector<double> a(n * n), x(n);
    for (auto& v : a) v = U(rng);
    for (auto& v : x) v = U(rng);
    DynamicMatrix<double> A(n, n, a.data());
    DynamicVector<double> X(n, x.data());
    X = A * X;  // aliased
    auto want = host_gemm(a, x, n, 1, n);
    auto got = X.toVector();
    bool ok = true;
    for (size_t i = 0; i < n; ++i) ok = ok && relaxed(got[i], want[i]);
    check(ok, "aliased x = A * x");
  }
  // CompressedMatrix: append / finalize, then y = S * x
  {
    // ( 1 0 2 0 )   ( 1 )   ( 7 )
    // ( 0 0 0 0 ) * ( 2 ) = ( 0 )
    // ( 0 3 0 4 )   ( 3 )   ( 22)
    //               ( 4 )
    CompressedMatrix<int> S(3, 4, 4);
    S.append(0, 0, 1); S.append(0, 2, 2); S.finalize(0);
    S.finalize(1);
    S.append(2, 1, 3); S.append(2, 3, 4); S.finalize(2);
    DynamicVector<int> x{1, 2, 3, 4}, y;
    y = S * x;
    check(S.nonZeros() == 4 && y.toVector() == std::vector<int>({7, 0, 22}), "CompressedMatrix<int> * vector known answer");
    check(throws<std::logic_error>([&] { CompressedMatrix<int> T(2, 2); T.append(0, 1, 1); T.append(0, 0, 1); }, ""),
          "append with non-increasing index throws");
    DynamicVector<int> z(3);
    check(throws<std::invalid_argument>([&] { DynamicVector<int> w; w = S * z; }, "Matrix and vector sizes do not match"),
          "S * x size mismatch throws invalid_argument");
  }
  {
    const size_t m = 300, n = 280;
    CompressedMatrix<double> S(m, n);
    std::vector<double> dense(m * n, 0.0), x(n);
    for (size_t i = 0; i < m; ++i) {
      for (size_t j = 0; j < n; ++j)
        if ((i * 7 + j * 3) % 11 == 0 || (i % 50 == 0 && j % 2 == 0)) {
          const double v = U(rng);
          S.append(i, j, v);
          dense[i * n + j] = v;
        }
      S.finalize(i);
    }
    for (auto& v : x) v = U(rng);
    DynamicVector<double> X(n, x.data()), Y;
    Y = S * X;
    auto got = Y.toVector();
    auto want = host_gemm(dense, x, m, 1, n);
    bool ok = true;
    for (size_t i = 0; i < m; ++i) ok = ok && relaxed(got[i], want[i], 1e-12);
    check(ok, "CompressedMatrix<double> * vector (mixed short and long rows)");
  }
  // solve(): lse/DenseTest.h -- random A with det != 0, A*x == b, N = 0..12 (+ larger)
  for (size_t n : {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 33, 100, 256}) {
    std::vector<double> a(n * n), b(n);
    for (auto& v : a) v = U(rng);
    for (size_t i = 0; i < n; ++i) a[i * n + i] += 2.0;
    for (auto& v : b) v = U(rng);
    DynamicMatrix<double> A(n, n, a.data());
    DynamicVector<double> B(n, b.data()), x1, x2;
    solve(A, x1, B);
    x2 = solve(A, B);
    DynamicVector<double> Ax;
    Ax = A * x1;
    auto ax = Ax.toVector(), v1 = x1.toVector(), v2 = x2.toVector();
    bool ok = v1 == v2;
    for (size_t i = 0; i < n; ++i) ok = ok && relaxed(ax[i], b[i]);
    check(ok, "solve(A,x,b) and x = solve(A,b), N = " + std::to_string(n));
  }
  {
    const size_t n = 20, r = 5;
    std::vector<double> a(n * n), b(n * r);
    for (auto& v : a) v = U(rng);
    for (size_t i = 0; i < n; ++i) a[i * n + i] += 3.0;
    for (auto& v : b) v = U(rng);
    DynamicMatrix<double> A(n, n, a.data()), B(n, r, b.data()), X, AX;
    solve(A, X, B);
    AX = A * X;
    auto g = AX.toVector();
    bool ok = X.rows() == n && X.columns() == r;
    for (size_t q = 0; q < n * r; ++q) ok = ok && relaxed(g[q], b[q]);
    check(ok, "solve(A,X,B) with 5 right-hand sides");
  }
  // products over the other storage orders: trans(A) * B, A * trans(B), trans(A) * trans(B)
  // (TDMatDMatMultExpr, DMatTDMatMultExpr, TDMatTDMatMultExpr) == the product of transposed copies
  {
    const size_t m = 37, n = 29, k = 45;
    std::vector<double> a(k * m), b(n * k), at(m * k), bt(k * n), c0(m * n);
    for (auto& v : a) v = U(rng);
    for (auto& v : b) v = U(rng);
    for (auto& v : c0) v = U(rng);
    for (size_t i = 0; i < k; ++i)
      for (size_t j = 0; j < m; ++j) at[j * k + i] = a[i * m + j];  // at = a^T (m x k)
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < k; ++j) bt[j * n + i] = b[i * k + j];  // bt = b^T (k x n)
    DynamicMatrix<double> A(k, m, a.data()), B(n, k, b.data()), At(m, k, at.data()), Bt(k, n, bt.data());
    DynamicMatrix<double> Bk(k, n, bt.data()), Am(m, k, at.data());
    DynamicMatrix<double> C1, C2, C3, R, C4(m, n, c0.data());
    C1 = trans(A) * Bk;
    C2 = Am * trans(B);
    C3 = trans(A) * trans(B);
    R = At * Bt;
    C4 += 2.0 * (trans(A) * trans(B));
    auto r = R.toVector(), g1 = C1.toVector(), g2 = C2.toVector(), g3 = C3.toVector(), g4 = C4.toVector();
    bool ok = C1.rows() == m && C1.columns() == n && g1 == r && g2 == r && g3 == r;
    for (size_t q = 0; q < m * n; ++q) ok = ok && relaxed(g4[q], c0[q] + 2.0 * r[q], 1e-12);
    ok = ok && throws<std::invalid_argument>([&] { C1 = trans(A) * B; }, "Matrix sizes do not match");
    check(ok, "trans(A)*B, A*trans(B), trans(A)*trans(B), C += s*(trans(A)*trans(B))");
  }
  // det (operations/determinant/DenseTest.cpp:285-470): 0 x 0 -> 0, 1 x 1 -> A(0,0) exactly,
  // unit-triangular 9 x 9 -> 1 exactly, non-square throws; random 8 x 8 against a host LU
  {
    bool ok = det(DynamicMatrix<double>(0, 0)) == 0.0;
    ok = ok && det(DynamicMatrix<double>(1, 1, -2.75)) == -2.75;
    const size_t n9 = 9;
    std::vector<double> lo(n9 * n9, 0.0), up(n9 * n9, 0.0);
    for (size_t i = 0; i < n9; ++i) {
      lo[i * n9 + i] = up[i * n9 + i] = 1.0;
      for (size_t j = 0; j < i; ++j) lo[i * n9 + j] = U(rng);
      for (size_t j = i + 1; j < n9; ++j) up[i * n9 + j] = U(rng);
    }
    ok = ok && det(DynamicMatrix<double>(n9, n9, lo.data())) == 1.0 &&
         det(DynamicMatrix<double>(n9, n9, up.data())) == 1.0;
    ok = ok && throws<std::invalid_argument>([&] { (void)det(DynamicMatrix<double>(2, 3, 0.0)); },
                                             "Invalid non-square matrix provided");
    // random 8 x 8: host Gaussian elimination with partial pivoting in long double
    const size_t n = 8;
    std::vector<double> a(n * n);
    for (auto& v : a) v = U(rng);
    std::vector<long double> w(a.begin(), a.end());
    long double ref = 1.0L;
    for (size_t k = 0; k < n; ++k) {
      size_t piv = k;
      for (size_t i = k + 1; i < n; ++i)
        if (std::fabs((double)w[i * n + k]) > std::fabs((double)w[piv * n + k])) piv = i;
      if (piv != k) {
        for (size_t j = 0; j < n; ++j) std::swap(w[k * n + j], w[piv * n + j]);
        ref = -ref;
      }
      ref *= w[k * n + k];
      for (size_t i = k + 1; i < n; ++i) {
        const long double l = w[i * n + k] / w[k * n + k];
        for (size_t j = k; j < n; ++j) w[i * n + j] -= l * w[k * n + j];
      }
    }
    ok = ok && relaxed(det(DynamicMatrix<double>(n, n, a.data())), (double)ref, 1e-12);
    check(ok, "det");
  }
  // inv / invert (Inversion.h:3337-3347): A * inv(A) = I, a power-of-two diagonal inverts
  // exactly, a singular matrix throws std::runtime_error, a non-square one std::invalid_argument
  {
    const size_t n = 37;
    std::vector<double> a(n * n);
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) a[i * n + j] = U(rng) + (i == j ? 4.0 : 0.0);
    DynamicMatrix<double> A(n, n, a.data());
    DynamicMatrix<double> X = inv(A);
    auto g = X.toVector();
    bool ok = X.rows() == n && X.columns() == n;
    for (size_t i = 0; i < n && ok; ++i)
      for (size_t j = 0; j < n; ++j) {
        long double s = 0.0L;
        for (size_t k = 0; k < n; ++k) s += (long double)a[i * n + k] * g[k * n + j];
        ok = ok && std::fabs((double)s - (i == j ? 1.0 : 0.0)) <= 1e-13;
      }
    std::vector<double> dg(16, 0.0);
    for (size_t i = 0; i < 4; ++i) dg[i * 4 + i] = std::ldexp(1.0, (int)i - 2);
    DynamicMatrix<double> D(4, 4, dg.data());
    invert(D);
    auto dv = D.toVector();
    for (size_t i = 0; i < 4; ++i) ok = ok && dv[i * 4 + i] == std::ldexp(1.0, 2 - (int)i);
    std::vector<double> sg(9, 1.0);  // rank one
    ok = ok && throws<std::runtime_error>([&] { DynamicMatrix<double> S(3, 3, sg.data()); invert(S); },
                                          "Inversion of singular matrix failed");
    ok = ok && throws<std::invalid_argument>([&] { DynamicMatrix<double> R(2, 3, 1.0); invert(R); },
                                             "Invalid non-square matrix provided");
    check(ok, "inv / invert");
  }
  // y = trans(S) * x (TSMatDVecMultExpr.h:329-383): bit-identical to the reference's loop --
  // y reset to +0, then rows j of S in order, y[i] += v * x[j] -- including an entry whose
  // products are all -0 (+0 as in the reference); size mismatch throws
  {
    const size_t m = 700, n = 450;
    CompressedMatrix<double> S(m, n);
    std::vector<std::vector<std::pair<size_t, double>>> rows(m);
    std::uniform_int_distribution<int> len(0, 9);
    for (size_t i = 0; i < m; ++i) {
      std::vector<size_t> cols;
      for (int q = len(rng); q > 0; --q) cols.push_back((size_t)(std::uniform_int_distribution<int>(0, (int)n - 2)(rng)));
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      for (size_t j : cols) {
        const double v = U(rng);
        S.append(i, j, v);
        rows[i].push_back({j, v});
      }
      if (i % 50 == 3) {  // column n - 1: only -0 products (x = 0 on these rows, v < 0)
        S.append(i, n - 1, -1.0);
        rows[i].push_back({n - 1, -1.0});
      }
      S.finalize(i);
    }
    std::vector<double> x(m);
    for (size_t i = 0; i < m; ++i) x[i] = (i % 50 == 3) ? 0.0 : U(rng);
    DynamicVector<double> X(m, x.data()), Y;
    Y = trans(S) * X;
    std::vector<double> want(n, 0.0);
    for (size_t i = 0; i < m; ++i)
      for (auto& e : rows[i]) {
        const double prod = e.second * x[i];
        want[e.first] = want[e.first] + prod;
      }
    auto g = Y.toVector();
    bool ok = Y.size() == n;
    for (size_t j = 0; j < n && ok; ++j)
      ok = std::memcmp(&g[j], &want[j], sizeof(double)) == 0;
    ok = ok && !std::signbit(g[n - 1]) && g[n - 1] == 0.0;
    ok = ok && throws<std::invalid_argument>([&] { DynamicVector<double> Z(n, 1.0); Y = trans(S) * Z; },
                                             "Matrix and vector sizes do not match");
    check(ok, "y = trans(S) * x bit-identical to the reference loop");
  }
  // y -= A * x (DMatDVecMultExpr subAssign)
  {
    const size_t m = 90, n = 70;
    std::vector<double> a(m * n), x(n), y0(m);
    for (auto& v : a) v = U(rng);
    for (auto& v : x) v = U(rng);
    for (auto& v : y0) v = U(rng);
    DynamicMatrix<double> A(m, n, a.data());
    DynamicVector<double> X(n, x.data()), Y(m, y0.data());
    Y -= A * X;
    auto g = Y.toVector();
    bool ok = true;
    for (size_t i = 0; i < m; ++i) {
      double s = 0.0;
      for (size_t j = 0; j < n; ++j) s += a[i * n + j] * x[j];
      ok = ok && relaxed(g[i], y0[i] - s, 1e-12);
    }
    check(ok, "y -= A * x");
  }
  // column-major containers: products over all storage-order combinations against a host
  // product (DMatTDMatMultExpr / TDMatDMatMultExpr / TDMatTDMatMultExpr, TDMatDVecMultExpr),
  // aliasing, solve / det / inv, and a column-built sparse matrix bit-exact to the reference's
  // column-order loop (TSMatDVecMultExpr.h:374-383)
  {
    const size_t m = 45, k = 38, n = 29;
    std::vector<double> a(m * k), b(k * n);  // row-major host copies
    for (auto& v : a) v = U(rng);
    for (auto& v : b) v = U(rng);
    auto colmajor = [](const std::vector<double>& r, size_t rows, size_t cols) {
      std::vector<double> c(rows * cols);
      for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < cols; ++j) c[j * rows + i] = r[i * cols + j];
      return c;
    };
    std::vector<double> want(m * n, 0.0);
    for (size_t i = 0; i < m; ++i)
      for (size_t j = 0; j < n; ++j)
        for (size_t q = 0; q < k; ++q) want[i * n + j] += a[i * k + q] * b[q * n + j];
    DynamicMatrix<double> Ar(m, k, a.data()), Br(k, n, b.data());
    DynamicMatrix<double, columnMajor> Ac(m, k, colmajor(a, m, k).data()), Bc(k, n, colmajor(b, k, n).data());
    bool ok = Ac.rows() == m && Ac.columns() == k && Ac(3, 5) == a[3 * k + 5];
    auto close_rm = [&](const std::vector<double>& g) {
      bool r = g.size() == m * n;
      for (size_t i = 0; i < m * n && r; ++i) r = relaxed(g[i], want[i], 1e-12);
      return r;
    };
    auto close_cm = [&](const std::vector<double>& g) {
      bool r = g.size() == m * n;
      for (size_t i = 0; i < m && r; ++i)
        for (size_t j = 0; j < n && r; ++j) r = relaxed(g[j * m + i], want[i * n + j], 1e-12);
      return r;
    };
    DynamicMatrix<double> C1;
    C1 = Ac * Br; ok = ok && close_rm(C1.toVector());
    C1 = Ar * Bc; ok = ok && close_rm(C1.toVector());
    C1 = Ac * Bc; ok = ok && close_rm(C1.toVector());
    DynamicMatrix<double, columnMajor> C2;
    C2 = Ar * Br; ok = ok && C2.rows() == m && C2.columns() == n && close_cm(C2.toVector());
    C2 = Ac * Bc; ok = ok && close_cm(C2.toVector());
    C2 = Ac * Br; ok = ok && close_cm(C2.toVector());
    // aliasing: B = A * B with column-major B (square A)
    std::vector<double> sq(k * k);
    for (auto& v : sq) v = U(rng);
    DynamicMatrix<double, columnMajor> Sq(k, k, colmajor(sq, k, k).data()), Bc2(k, n, colmajor(b, k, n).data());
    Bc2 = Sq * Bc2;
    {
      auto g = Bc2.toVector();
      for (size_t i = 0; i < k && ok; ++i)
        for (size_t j = 0; j < n && ok; ++j) {
          double s = 0.0;
          for (size_t q = 0; q < k; ++q) s += sq[i * k + q] * b[q * n + j];
          ok = relaxed(g[j * k + i], s, 1e-12);
        }
    }
    // y = A * x, column-major A
    std::vector<double> x(k);
    for (auto& v : x) v = U(rng);
    DynamicVector<double> X(k, x.data()), Y;
    Y = Ac * X;
    {
      auto g = Y.toVector();
      for (size_t i = 0; i < m && ok; ++i) {
        double s = 0.0;
        for (size_t q = 0; q < k; ++q) s += a[i * k + q] * x[q];
        ok = relaxed(g[i], s, 1e-12);
      }
    }
    // solve / det / inv with a column-major system matrix
    std::vector<double> sa(k * k);
    for (size_t i = 0; i < k; ++i)
      for (size_t j = 0; j < k; ++j) sa[i * k + j] = U(rng) + (i == j ? 5.0 : 0.0);
    DynamicMatrix<double> Sr(k, k, sa.data());
    DynamicMatrix<double, columnMajor> Sc(k, k, colmajor(sa, k, k).data());
    DynamicVector<double> xs, xr;
    solve(Sc, xs, X);
    solve(Sr, xr, X);
    {
      auto g1 = xs.toVector(), g2 = xr.toVector();
      for (size_t i = 0; i < k && ok; ++i) ok = relaxed(g1[i], g2[i], 1e-12);
    }
    ok = ok && relaxed(det(Sc), det(Sr), 1e-12);
    {
      auto Ic = inv(Sc);
      auto Ir = inv(Sr);
      auto g1 = Ic.toVector(), g2 = Ir.toVector();
      for (size_t i = 0; i < k && ok; ++i)
        for (size_t j = 0; j < k && ok; ++j) ok = relaxed(g1[j * k + i], g2[i * k + j], 1e-12);
    }
    check(ok, "column-major DynamicMatrix products, aliasing, solve / det / inv");
  }
  {
    const size_t m = 300, n = 260;
    CompressedMatrix<double, columnMajor> S(m, n);
    std::vector<std::vector<std::pair<size_t, double>>> cols(n);
    std::uniform_int_distribution<int> len(0, 8), row(0, (int)m - 1);
    for (size_t j = 0; j < n; ++j) {
      std::vector<size_t> rs;
      for (int q = len(rng); q > 0; --q) rs.push_back((size_t)row(rng));
      std::sort(rs.begin(), rs.end());
      rs.erase(std::unique(rs.begin(), rs.end()), rs.end());
      for (size_t i : rs) {
        const double v = U(rng);
        S.append(i, j, v);
        cols[j].push_back({i, v});
      }
      S.finalize(j);
    }
    std::vector<double> x(n);
    for (auto& v : x) v = U(rng);
    DynamicVector<double> X(n, x.data()), Y;
    Y = S * X;
    std::vector<double> want(m, 0.0);
    for (size_t j = 0; j < n; ++j)
      for (auto& e : cols[j]) {
        const double prod = e.second * x[j];
        want[e.first] = want[e.first] + prod;
      }
    auto g = Y.toVector();
    bool ok = Y.size() == m && S.rows() == m && S.columns() == n;
    for (size_t i = 0; i < m && ok; ++i) ok = std::memcmp(&g[i], &want[i], sizeof(double)) == 0;
    if (!cols[7].empty()) ok = ok && S(cols[7][0].first, 7) == cols[7][0].second;
    check(ok, "column-major CompressedMatrix: S * x bit-identical to the reference loop");
  }
  // readMatrixMarket: a symmetric coordinate file expanded to both triangles, then S * x
  {
    const char* path = "/tmp/bzg_frontend_test.mtx";
    FILE* f = std::fopen(path, "w");
    bool ok = f != nullptr;
    if (f) {
      std::fputs("%%MatrixMarket matrix coordinate real symmetric\n% comment\n3 3 4\n1 1 2\n2 1 -1\n3 2 -1\n3 3 2\n", f);
      std::fclose(f);
      CompressedMatrix<double> S;
      readMatrixMarket(path, S);
      DynamicVector<double> X(3, 1.0), Y;
      Y = S * X;
      ok = S.rows() == 3 && S.columns() == 3 && S.nonZeros() == 6 && S(0, 1) == -1.0 && S(1, 0) == -1.0 &&
           Y.toVector() == std::vector<double>{1.0, -2.0, 1.0};
      std::remove(path);
      ok = ok && throws<std::invalid_argument>([&] { CompressedMatrix<double> Z; readMatrixMarket(path, Z); }, "");
    }
    check(ok, "readMatrixMarket");
  }
  // smatdvecmult/AliasingTest.cpp:69-300 fixtures (sA3x4, sB3x3, da4, db3): products by hand,
  // aliased assignment x = S * x equal to the non-aliased result, row- and column-major S
  {
    CompressedMatrix<double> sA(3, 4), sB(3, 3);
    CompressedMatrix<double, columnMajor> tsA(3, 4), tsB(3, 3);
    const double a[3][4] = {{-1, 0, -2, 0}, {0, 2, -3, 1}, {0, 1, 2, 2}};
    const double b[3][3] = {{-1, 0, 0}, {1, -2, 2}, {0, 0, -3}};
    for (size_t i = 0; i < 3; ++i) {
      for (size_t j = 0; j < 4; ++j)
        if (a[i][j] != 0) { sA.set(i, j, a[i][j]); tsA.set(i, j, a[i][j]); }
      for (size_t j = 0; j < 3; ++j)
        if (b[i][j] != 0) { sB.set(i, j, b[i][j]); tsB.set(i, j, b[i][j]); }
    }
    const std::vector<double> da4 = {-1, 0, -3, 2}, db3 = {1, 2, 3};
    const std::vector<double> want_a = {7, 11, -2}, want_b = {-1, 3, -9};
    DynamicVector<double> x4(4, da4.data()), r;
    r = sA * x4;
    bool ok = r.toVector() == want_a;
    r = tsA * x4;
    ok = ok && r.toVector() == want_a;
    DynamicVector<double> x3(3, db3.data());
    x3 = sB * x3;  // aliased
    ok = ok && x3.toVector() == want_b;
    DynamicVector<double> y3(3, db3.data());
    y3 = tsB * y3;  // aliased, column-major
    ok = ok && y3.toVector() == want_b;
    check(ok, "smatdvecmult AliasingTest fixtures (row- and column-major, aliased)");
  }
  // ClassTest2.cpp:1244-1273 with column-major int operands (its row-major/column-major forms):
  // (1 0 2; 0 3 0; 4 0 5) * (0 2 0 0; 1 3 0 4; 0 0 0 5) = (0 2 0 10; 3 9 0 12; 0 8 0 25)
  {
    const std::vector<int> b_cm = {0, 1, 0, 2, 3, 0, 0, 0, 0, 0, 4, 5};  // 3 x 4, by columns
    const std::vector<int> a_cm = {1, 0, 4, 0, 3, 0, 2, 0, 5};            // 3 x 3, by columns
    const std::vector<int> a_rm = {1, 0, 2, 0, 3, 0, 4, 0, 5};
    const std::vector<int> b_rm = {0, 2, 0, 0, 1, 3, 0, 4, 0, 0, 0, 5};
    const std::vector<int> want = {0, 2, 0, 10, 3, 9, 0, 12, 0, 8, 0, 25};
    DynamicMatrix<int> Ar(3, 3, a_rm.data()), Br(3, 4, b_rm.data()), C;
    DynamicMatrix<int, columnMajor> Ac(3, 3, a_cm.data()), Bc(3, 4, b_cm.data());
    C = Ar * Bc;
    bool ok = C.toVector() == want;
    C = Ac * Br;
    ok = ok && C.toVector() == want;
    C = Ac * Bc;
    ok = ok && C.toVector() == want;
    DynamicVector<int> x(4, 1), y;
    y = Bc * x;  // row sums of B
    ok = ok && y.toVector() == std::vector<int>{2, 8, 5};
    check(ok, "ClassTest2 known answer with column-major int operands");
  }
  // lu() with a column-major A (lu/DenseTest.h: P * L * U == A for column-major storage)
  {
    const size_t n = 41;
    std::vector<double> a(n * n);  // by columns
    for (auto& v : a) v = U(rng);
    DynamicMatrix<double, columnMajor> A(n, n, a.data()), L, Um, P;
    lu(A, L, Um, P);
    auto l = L.toVector(), u = Um.toVector(), pm = P.toVector();
    bool ok = L.rows() == n && Um.columns() == n && P.rows() == n;
    for (size_t j = 0; j < n && ok; ++j)
      for (size_t i = 0; i < n && ok; ++i) {
        if (i < j) ok = l[j * n + i] == 0.0;                    // L unit lower
        if (i == j) ok = ok && l[j * n + i] == 1.0;
        if (i > j) ok = ok && u[j * n + i] == 0.0;               // U upper
      }
    for (size_t i = 0; i < n && ok; ++i) {                       // P a permutation
      double rs = 0.0;
      for (size_t j = 0; j < n; ++j) rs += pm[j * n + i];
      ok = rs == 1.0;
    }
    for (size_t i = 0; i < n && ok; ++i)                         // P * L * U == A
      for (size_t j = 0; j < n && ok; ++j) {
        double s = 0.0;
        for (size_t q = 0; q < n; ++q) {
          double lu_qj = 0.0;
          for (size_t r = 0; r < n; ++r) lu_qj += l[r * n + q] * u[j * n + r];
          s += pm[q * n + i] * lu_qj;
        }
        ok = relaxed(s, a[j * n + i], 1e-12);
      }
    check(ok, "lu() column-major: A == P * L * U");
  }
  // y = trans(A) * x (TDMatDVecMultExpr)
  {
    const size_t m = 300, n = 77;
    std::vector<double> a(m * n), x(m);
    for (auto& v : a) v = U(rng);
    for (auto& v : x) v = U(rng);
    DynamicMatrix<double> A(m, n, a.data());
    DynamicVector<double> X(m, x.data()), Y;
    Y = trans(A) * X;
    auto g = Y.toVector();
    bool ok = Y.size() == n;
    for (size_t j = 0; j < n; ++j) {
      double s = 0.0;
      for (size_t i = 0; i < m; ++i) s += a[i * n + j] * x[i];
      ok = ok && relaxed(g[j], s, 1e-12);
    }
    ok = ok && throws<std::invalid_argument>([&] { DynamicVector<double> Z(n, 1.0); Y = trans(A) * Z; },
                                             "Matrix and vector sizes do not match");
    check(ok, "y = trans(A) * x");
  }
  // set / insert (CompressedMatrix.h:2089-2135): random order, overwrite, duplicate insert
  {
    const size_t m = 50, n = 40;
    CompressedMatrix<double> S(m, n);
    std::vector<double> dense(m * n, 0.0), x(n);
    for (auto& v : x) v = U(rng);
    std::uniform_int_distribution<size_t> Ri(0, m - 1), Rj(0, n - 1);
    for (int t = 0; t < 300; ++t) {
      const size_t i = Ri(rng), j = Rj(rng);
      const double v = U(rng);
      S.set(i, j, v);
      dense[i * n + j] = v;
    }
    bool ok = throws<std::invalid_argument>([&] {
      size_t i = 0, j = 0;
      while (dense[i * n + j] == 0.0) { ++j; if (j == n) { j = 0; ++i; } }
      S.insert(i, j, 1.0);
    }, "Bad access index");
    size_t fi = 0, fj = 0;
    while (dense[fi * n + fj] != 0.0) { ++fj; if (fj == n) { fj = 0; ++fi; } }
    S.insert(fi, fj, 2.5);
    dense[fi * n + fj] = 2.5;
    DynamicVector<double> X(n, x.data()), Y;
    Y = S * X;
    auto got = Y.toVector();
    auto want = host_gemm(dense, x, m, 1, n);
    for (size_t i = 0; i < m; ++i) ok = ok && relaxed(got[i], want[i], 1e-12);
    for (size_t i = 0; i < m; ++i)
      for (size_t j = 0; j < n; ++j) ok = ok && S(i, j) == dense[i * n + j];
    check(ok, "CompressedMatrix set/insert in random order, then S * x");
  }
  // lu(A, L, U, P): lu/DenseTest.h:159-170 -- the decomposition reproduces A; for a row-major
  // matrix A = L * U * P with L lower, U unit upper and P a permutation (LU.h:166-172)
  for (size_t n : {1, 2, 3, 9, 40, 130}) {
    std::vector<double> a(n * n);
    for (auto& v : a) v = U(rng);
    DynamicMatrix<double> A(n, n, a.data()), L, Um, P, LU, LUP;
    lu(A, L, Um, P);
    LU = L * Um;
    LUP = LU * P;
    auto l = L.toVector(), u = Um.toVector(), pm = P.toVector(), r = LUP.toVector();
    bool ok = L.rows() == n && Um.columns() == n && P.rows() == n;
    for (size_t i = 0; i < n; ++i) {
      size_t ones = 0;
      for (size_t j = 0; j < n; ++j) {
        if (j > i) ok = ok && l[i * n + j] == 0.0;
        if (j < i) ok = ok && u[i * n + j] == 0.0;
        if (j == i) ok = ok && u[i * n + j] == 1.0;
        ok = ok && (pm[i * n + j] == 0.0 || pm[i * n + j] == 1.0);
        ones += pm[i * n + j] == 1.0;
        ok = ok && relaxed(r[i * n + j], a[i * n + j], 1e-12);
      }
      ok = ok && ones == 1;
    }
    check(ok, "lu(A,L,U,P), A == L*U*P, N = " + std::to_string(n));
  }
  {
    DynamicMatrix<double> S3{{1, 2, 3}, {2, 4, 6}, {0, 1, 1}};
    DynamicMatrix<double> S9(9, 9, 1.0);
    DynamicMatrix<double> R(3, 4);
    DynamicVector<double> b3(3, 1.0), b9(9, 1.0), b4(4, 1.0), x;
    check(throws<std::runtime_error>([&] { solve(S3, x, b3); }, "Solving LSE with singular system matrix failed"),
          "singular 3x3 throws runtime_error (closed-form message)");
    check(throws<std::runtime_error>([&] { solve(S9, x, b9); }, "Inversion of singular matrix failed"),
          "singular 9x9 throws runtime_error (LAPACK message)");
    check(throws<std::invalid_argument>([&] { solve(R, x, b3); }, "Invalid non-square system matrix provided"),
          "non-square system throws invalid_argument");
    check(throws<std::invalid_argument>([&] { solve(S3, x, b4); }, "Invalid right-hand side vector provided"),
          "rhs size mismatch throws invalid_argument");
  }
  // FP64-accurate GEMM routing
  {
    const size_t m = 300, n = 200, k = 700;
    std::vector<double> a(m * k), b(k * n);
    for (auto& v : a) v = U(rng);
    for (auto& v : b) v = U(rng);
    DynamicMatrix<double> A(m, k, a.data()), B(k, n, b.data()), C1, C2;
    C1 = A * B;
    setFp64AccurateGemm(true);
    C2 = A * B;
    setFp64AccurateGemm(false);
    auto g1 = C1.toVector(), g2 = C2.toVector();
    double num = 0, den = 0;
    for (size_t q = 0; q < m * n; ++q) { num += (g1[q] - g2[q]) * (g1[q] - g2[q]); den += g1[q] * g1[q]; }
    check(std::sqrt(num / den) < 1e-14, "FP64-accurate split GEMM agrees with native FP64 (rel. Frobenius < 1e-14)");
  }
  // orders beyond the batched kernels (n > 512) and more than 8 right-hand sides: solve(), lu(),
  // det(), inv() through the single-system LU (LSE.h:5229-5257, LU.h:98-140, Inversion.h:3337-3347)
  {
    const size_t n = 700, r = 12;
    std::vector<double> a(n * n), b(n * r);
    for (auto& v : a) v = U(rng);
    for (size_t i = 0; i < n; ++i) a[i * n + i] += 30.0;
    for (auto& v : b) v = U(rng);
    DynamicMatrix<double> A(n, n, a.data()), B(n, r, b.data()), X, AX;
    DynamicVector<double> bv(n, b.data()), x, Ax;
    solve(A, X, B);
    solve(A, x, bv);
    AX = A * X;
    Ax = A * x;
    auto g = AX.toVector(), gx = Ax.toVector();
    bool ok = X.rows() == n && X.columns() == r;
    for (size_t q = 0; q < n * r; ++q) ok = ok && relaxed(g[q], b[q], 1e-12);
    for (size_t i = 0; i < n; ++i) ok = ok && relaxed(gx[i], b[i], 1e-12);
    check(ok, "solve(A,X,B) with N = 700 and 12 right-hand sides, solve(A,x,b) with N = 700");
    const size_t m = 40;
    std::vector<double> c(m * m), d(m * 20);
    for (auto& v : c) v = U(rng);
    for (size_t i = 0; i < m; ++i) c[i * m + i] += 5.0;
    for (auto& v : d) v = U(rng);
    DynamicMatrix<double> C(m, m, c.data()), D(m, 20, d.data()), Y, CY;
    solve(C, Y, D);
    CY = C * Y;
    auto h = CY.toVector();
    ok = true;
    for (size_t q = 0; q < m * 20; ++q) ok = ok && relaxed(h[q], d[q], 1e-12);
    check(ok, "solve(A,X,B) with N = 40 and 20 right-hand sides");
    // det(A) for N = 700 against det(A^T) (a different factorisation: other pivots and
    // rounding, the same value) and inv(A) * A == I
    // (G = I + 0.02 U(-1, 1): |det| of order one)
    std::vector<double> gm(n * n), at(n * n);
    for (size_t q = 0; q < n * n; ++q) gm[q] = 0.02 * U(rng) + (q % (n + 1) == 0 ? 1.0 : 0.0);
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) at[j * n + i] = gm[i * n + j];
    DynamicMatrix<double> G(n, n, gm.data()), Gt(n, n, at.data());
    const double d1 = det(G), d2 = det(Gt);
    check(std::isfinite(d1) && d1 != 0.0 && relaxed(d1, d2, 1e-10), "det(G) for N = 700 (== det(G^T))");
    DynamicMatrix<double> Ai = inv(G), I1;
    I1 = Ai * G;
    auto e = I1.toVector();
    ok = true;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) ok = ok && std::fabs(e[i * n + j] - (i == j ? 1.0 : 0.0)) <= 1e-12;
    check(ok, "inv(G) * G == I for N = 700");
    DynamicMatrix<double> L, Uu, P, LU, LUP;
    lu(G, L, Uu, P);
    LU = L * Uu;
    LUP = LU * P;
    auto f = LUP.toVector();
    ok = true;
    for (size_t q = 0; q < n * n; ++q) ok = ok && relaxed(f[q], gm[q], 1e-12);
    check(ok, "lu(G,L,U,P), G == L*U*P, N = 700");
  }
  // aliased accumulation on the device: C += A * C equals C + (A * C) computed separately
  {
    const size_t n = 64;
    std::vector<double> a(n * n), c(n * n);
    for (auto& v : a) v = U(rng);
    for (auto& v : c) v = U(rng);
    DynamicMatrix<double> A(n, n, a.data()), C(n, n, c.data()), T;
    T = A * C;
    C -= A * C;
    auto t = T.toVector(), g = C.toVector();
    bool ok = true;
    for (size_t q = 0; q < n * n; ++q) ok = ok && g[q] == c[q] - t[q];
    DynamicVector<double> y(n, c.data()), z;
    z = A * y;
    y += A * y;
    auto zz = z.toVector(), yy = y.toVector();
    for (size_t i = 0; i < n; ++i) ok = ok && yy[i] == c[i] + zz[i];
    check(ok, "aliased C -= A * C and y += A * y (device add of the evaluated product)");
  }
  // complex<double> / complex<float> elements: products with trans() / ctrans(), scaled and
  // accumulated products, GEMV with ctrans(A), and solve() of a complex system
  {
    using cd = std::complex<double>;
    using cf = std::complex<float>;
    auto close = [](cd a, cd b, double eps) { return std::abs(a - b) <= eps * std::max(1.0, std::abs(b)); };
    const size_t m = 37, k = 53, n = 29;
    std::vector<cd> a(m * k), b(k * n), bh(n * k), at(k * m);
    for (auto& v : a) v = cd(U(rng), U(rng));
    for (auto& v : b) v = cd(U(rng), U(rng));
    for (size_t p = 0; p < k; ++p)
      for (size_t j = 0; j < n; ++j) bh[j * k + p] = std::conj(b[p * n + j]);
    for (size_t i = 0; i < m; ++i)
      for (size_t p = 0; p < k; ++p) at[p * m + i] = a[i * k + p];
    DynamicMatrix<cd> A(m, k, a.data()), B(k, n, b.data()), Bh(n, k, bh.data()), At(k, m, at.data()), C1, C2, C3;
    C1 = A * B;
    C2 = A * ctrans(Bh);
    C3 = trans(At) * B;
    const auto ref = host_gemm(a, b, m, n, k);
    auto c1 = C1.toVector(), c2 = C2.toVector(), c3 = C3.toVector();
    bool ok = C1.rows() == m && C1.columns() == n;
    for (size_t q = 0; q < m * n; ++q)
      ok = ok && close(c1[q], ref[q], 1e-13) && close(c2[q], ref[q], 1e-13) && close(c3[q], ref[q], 1e-13);
    check(ok, "complex<double>: A * B == A * ctrans(B^H) == trans(A^T) * B (37 x 53 x 29)");
    const cd s(0.5, -1.0);
    C2 = s * (A * B);
    C2 -= A * B;
    c2 = C2.toVector();
    ok = true;
    for (size_t q = 0; q < m * n; ++q) ok = ok && close(c2[q], (s - 1.0) * ref[q], 1e-13);
    check(ok, "complex<double>: C = s * (A * B); C -= A * B");
    std::vector<cd> xv(m);
    for (auto& v : xv) v = cd(U(rng), U(rng));
    DynamicVector<cd> x(m), y;
    x.upload(xv.data());
    y = ctrans(A) * x;
    auto yv = y.toVector();
    ok = y.size() == k;
    for (size_t p = 0; p < k; ++p) {
      cd t(0.0, 0.0);
      for (size_t i = 0; i < m; ++i) t += std::conj(a[i * k + p]) * xv[i];
      ok = ok && close(yv[p], t, 1e-13);
    }
    check(ok, "complex<double>: y = ctrans(A) * x");
    // solve(): N = 300 and N = 5 (a closed-form order in the reference), one and three RHS
    for (size_t N : {size_t(5), size_t(300)}) {
      std::vector<cd> g(N * N), r(N * 3);
      for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) g[i * N + j] = cd(U(rng), U(rng)) + (i == j ? cd(double(N), 0.0) : cd(0.0));
      for (auto& v : r) v = cd(U(rng), U(rng));
      DynamicMatrix<cd> G(N, N, g.data()), Rm(N, 3, r.data()), X;
      solve(G, X, Rm);
      auto xs = X.toVector();
      ok = X.rows() == N && X.columns() == 3;
      for (size_t i = 0; i < N && ok; ++i)
        for (size_t c = 0; c < 3; ++c) {
          cd t(0.0, 0.0);
          for (size_t j = 0; j < N; ++j) t += g[i * N + j] * xs[j * 3 + c];
          ok = ok && close(t, r[i * 3 + c], 1e-12);
        }
      DynamicVector<cd> bv(N), xx;
      std::vector<cd> r0(N);
      for (size_t i = 0; i < N; ++i) r0[i] = r[i * 3];
      bv.upload(r0.data());
      xx = solve(G, bv);
      auto x0 = xx.toVector();
      for (size_t i = 0; i < N; ++i) ok = ok && close(x0[i], xs[i * 3], 1e-12);
      check(ok, "complex<double>: solve(A, X, B) and x = solve(A, b), N = " + std::to_string(N));
    }
    {
      std::vector<cd> g(49, cd(0.0));
      for (size_t i = 0; i < 7; ++i) g[i * 7 + i] = cd(1.0, 1.0);
      for (size_t i = 0; i < 7; ++i) g[i * 7 + 3] = cd(0.0);  // a zero column: singular
      DynamicMatrix<cd> G(7, 7, g.data());
      DynamicVector<cd> bv(7, cd(1.0, 0.0)), xx;
      check(throws<std::runtime_error>([&] { solve(G, xx, bv); }, "Inversion of singular matrix failed"),
            "complex<double>: singular solve() throws (N = 7)");
    }
    {
      // column-major complex operands: C = A * B and solve() on a column-major system matrix
      DynamicMatrix<cd, columnMajor> Ac(m, k), Bc(k, n), Cc;
      std::vector<cd> as(m * k), bs(k * n);  // column-major storage = row-major transposes
      for (size_t i = 0; i < m; ++i)
        for (size_t p = 0; p < k; ++p) as[p * m + i] = a[i * k + p];
      for (size_t p = 0; p < k; ++p)
        for (size_t j = 0; j < n; ++j) bs[j * k + p] = b[p * n + j];
      Ac.storage().upload(as.data());
      Bc.storage().upload(bs.data());
      Cc = Ac * Bc;
      auto cs = Cc.storage().toVector();
      bool okc = Cc.rows() == m && Cc.columns() == n;
      for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j) okc = okc && close(cs[j * m + i], ref[i * n + j], 1e-13);
      check(okc, "complex<double> column-major: C = A * B");
      const size_t N = 40;
      std::vector<cd> g(N * N), gs(N * N), r(N);
      for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) g[i * N + j] = cd(U(rng), U(rng)) + (i == j ? cd(double(N), 0.0) : cd(0.0));
      for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) gs[j * N + i] = g[i * N + j];
      for (auto& v : r) v = cd(U(rng), U(rng));
      DynamicMatrix<cd, columnMajor> G(N, N);
      G.storage().upload(gs.data());
      DynamicVector<cd> bv(N), xx;
      bv.upload(r.data());
      solve(G, xx, bv);
      auto xs = xx.toVector();
      okc = xx.size() == N;
      for (size_t i = 0; i < N; ++i) {
        cd t(0.0, 0.0);
        for (size_t j = 0; j < N; ++j) t += g[i * N + j] * xs[j];
        okc = okc && close(t, r[i], 1e-12);
      }
      check(okc, "complex<double> column-major: solve(A, x, b), N = 40");
    }
    std::vector<cf> af(m * k), bf(k * n);
    for (size_t q = 0; q < m * k; ++q) af[q] = cf(a[q]);
    for (size_t q = 0; q < k * n; ++q) bf[q] = cf(b[q]);
    DynamicMatrix<cf> Af(m, k, af.data()), Bf(k, n, bf.data()), Cf;
    Cf = Af * Bf;
    auto cfv = Cf.toVector();
    const auto reff = host_gemm(af, bf, m, n, k);
    ok = true;
    for (size_t q = 0; q < m * n; ++q) ok = ok && std::abs(cfv[q] - reff[q]) <= 1e-4f * std::max(1.0f, std::abs(reff[q]));
    check(ok, "complex<float>: A * B (37 x 53 x 29)");
  }
  std::printf("%d failure(s)\n", failures);
  return failures;
}
