// CPU ORACLE (compiled reference, LAPACK-backed paths) -- test and benchmark infrastructure only.
//
// Thin extern "C" driver around the reference's dense solve() (blaze/math/dense/LSE.h:6452-6482:
// closed forms for N <= 6, solveNxN -> gesv -> dgesv_ for N >= 7, LSE.h:5733-5770), used as the
// CPU baseline of the batched solves.  The reference leaves LAPACK to the build
// (CMakeLists.txt:72, find_package(LAPACK)); oracle/Makefile links the LAPACK the image provides
// (MKL's single dynamic library) when it exists, and skips this driver otherwise.  Shared-memory
// parallelisation inside Blaze is switched off (BLAZE_USE_SHARED_MEMORY_PARALLELIZATION=0 on the
// command line): the systems are distributed over threads by an outer OpenMP loop instead, with
// LAPACK itself sequential.
#include <blaze/Math.h>
#include <omp.h>

#include <chrono>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
using CMat = blaze::CustomMatrix<double, blaze::unaligned, blaze::unpadded, blaze::rowMajor>;

// Solves systems [first, last) with solve(A, X, B) on views of the pools; returns the number of
// systems whose solve threw (singular).
int64_t solve_range(int64_t first, int64_t last, const int32_t* n, const int32_t* nrhs, const int64_t* a_off,
                    const int64_t* b_off, const double* A, const double* B, double* X) {
  int64_t fails = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : fails)
  for (int64_t s = first; s < last; ++s) {
    const size_t ns = (size_t)n[s], nr = (size_t)nrhs[s];
    // views of the pools (the reference's solve() only reads A and B: it copies both into
    // column-major temporaries before dgesv_)
    const CMat a(const_cast<double*>(A) + a_off[s], ns, ns);
    const CMat b(const_cast<double*>(B) + b_off[s], ns, nr);
    blaze::DynamicMatrix<double, blaze::rowMajor> x;
    try {
      blaze::solve(a, x, b);
      CMat out(X + b_off[s], ns, nr);
      out = x;
    } catch (const std::exception&) {
      ++fails;
    }
  }
  return fails;
}
}  // namespace

extern "C" {

// Batched solves A_i X_i = B_i (row-major n_i x n_i and n_i x nrhs_i at a_off[i], b_off[i]) with
// the reference's solve(), `threads` OpenMP threads over the systems.  Returns the best of
// `reps` timed passes (seconds) after one untimed pass; *fails = systems whose solve threw.
double ref_time_dsolve_batch(int64_t count, const int32_t* n, const int32_t* nrhs, const int64_t* a_off,
                             const int64_t* b_off, const double* A, const double* B, double* X, int threads,
                             int reps, int64_t* fails) {
  if (threads > 0) omp_set_num_threads(threads);
  *fails = solve_range(0, count, n, nrhs, a_off, b_off, A, B, X);
  double best = 1e300;
  for (int r = 0; r < reps; ++r) {
    const double t0 = now();
    solve_range(0, count, n, nrhs, a_off, b_off, A, B, X);
    const double t = now() - t0;
    if (t < best) best = t;
  }
  return best;
}

// X = solve(A, B) for one complex<double> system (row-major n x n and n x nrhs) with the
// reference's solve() (closed forms for n <= 6, zgesv_ beyond); returns 1 when it threw.
int ref_zsolve(int64_t n, int64_t nrhs, const std::complex<double>* A, const std::complex<double>* B,
               std::complex<double>* X) {
  using ZMat = blaze::CustomMatrix<std::complex<double>, blaze::unaligned, blaze::unpadded, blaze::rowMajor>;
  const ZMat a(const_cast<std::complex<double>*>(A), (size_t)n, (size_t)n);
  const ZMat b(const_cast<std::complex<double>*>(B), (size_t)n, (size_t)nrhs);
  blaze::DynamicMatrix<std::complex<double>, blaze::rowMajor> x;
  try {
    blaze::solve(a, x, b);
  } catch (const std::exception&) {
    return 1;
  }
  ZMat out(X, (size_t)n, (size_t)nrhs);
  out = x;
  return 0;
}

}  // extern "C"
