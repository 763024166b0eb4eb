/*
 * blaze_gpu.h -- C ABI of the MI355X (gfx950) multiplication & solve library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, throws nothing and returns
 * a bzg_status.  Matrix/vector pointers passed to compute calls are DEVICE pointers (HBM);
 * dense matrices are row-major with an explicit leading dimension (Blaze's `spacing()`).
 * Work is enqueued on the context's HIP stream and is asynchronous unless stated otherwise.
 *
 * Status convention (mirrors LAPACK `info`):
 *   0      success
 *   > 0    numerical failure: 1-based index of the first exactly-zero pivot (singular matrix)
 *   < 0    invalid argument / runtime error (see bzg_status_string)
 *
 * Reference interfaces replaced (paths relative to the Blaze source tree):
 *   bzg_dgemm / bzg_sgemm / bzg_igemm  <- DMatDMatMultExpr assign (blaze/math/expressions/
 *        DMatDMatMultExpr.h:516-569) and its BLAS seam cblas_dgemm (blaze/math/blas/cblas/gemm.h:161)
 *   bzg_dgemm_ex / bzg_sgemm_ex / bzg_igemm_ex <- the same products over column-major (transposed) operands
 *        (DMatTDMatMultExpr.h, TDMatDMatMultExpr.h, TDMatTDMatMultExpr.h; CBLAS transposes)
 *   bzg_dgemv / bzg_sgemv / bzg_igemv  <- DMatDVecMultExpr assign (DMatDVecMultExpr.h:408-458)
 *        and cblas_dgemv (blaze/math/blas/cblas/gemv.h:153)
 *   bzg_dcsrmv / bzg_scsrmv / bzg_icsrmv <- SMatDVecMultExpr element evaluation
 *        (SMatDVecMultExpr.h:204-207 -> SVecDVecInnerExpr.h:127-135)
 *   bzg_csr_plan_set_sum_start + bzg_csc_to_csr_host + bzg_dcsrmv <- TSMatDVecMultExpr assign
 *        (TSMatDVecMultExpr.h:329-383: column-major sparse matrix or trans(S) times a vector)
 *   bzg_dcg                            <- the CG loop of blazemark/src/blaze/CG.cpp:114-135 (its
 *        SpMV, dot products and vector updates, resident on the device)
 *   bzg_dgesv_batched                  <- solve() (blaze/math/dense/LSE.h:6343-6373) ->
 *        gesv (blaze/math/lapack/gesv.h:136-175) -> dgesv_ (blaze/math/lapack/clapack/gesv.h:207)
 *   bzg_dgetrf_batched / bzg_dgetrs_batched <- getrf / getrs (blaze/math/lapack/clapack/getrf.h:190-197,
 *        clapack/getrs.h:207-215), the factorisation under lu() (blaze/math/dense/LU.h:98-140)
 *   bzg_ddet_batched                   <- det() (blaze/math/expressions/DMatDetExpr.h:384-424)
 *   bzg_dinv_batched                   <- invert() / inv() by LU (blaze/math/dense/Inversion.h:3337-3347)
 *   bzg_dsgesv_batched                 <- same call site; LAPACK dsgesv semantics (mixed precision)
 *   bzg_dgesv / bzg_dgetrf / bzg_dgetrs <- the same LAPACK calls for one system of any order
 *        (clapack/gesv.h:207, getrf.h:190-197, getrs.h:207-215; solveNxN, LSE.h:5229-5257)
 *   bzg_ddet_lu / bzg_dgetri            <- det() (DMatDetExpr.h:324-357) and invert() by LU
 *        (Inversion.h:3337-3347) from those factors
 *   bzg_zgemm_ex / bzg_cgemm_ex        <- DMatDMatMultExpr for complex<double> / complex<float>
 *        elements and its BLAS call cblas_zgemm / cblas_cgemm (blaze/math/blas/cblas/gemm.h:197-283)
 *   bzg_zgemv_ex / bzg_cgemv_ex        <- DMatDVecMultExpr for complex elements and cblas_zgemv /
 *        cblas_cgemv (blaze/math/blas/cblas/gemv.h:186-245)
 *   bzg_zgesv                          <- solve() of a complex<double> system -> gesv -> zgesv_
 *        (blaze/math/lapack/clapack/gesv.h:317-333)
 *   bzg_dgemm_fp64acc[_i8]             <- no reference function: FP64-accurate GEMM from BF16
 *        (INT8) residue slices; drop-in for bzg_dgemm at the same call site
 *        (DMatDMatMultExpr.h:516)
 *   bzg_server_* / bzg_submit_*        <- no reference function: a request server over the same
 *        operations for concurrent clients
 */
#ifndef BLAZE_GPU_H
#define BLAZE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t bzg_status;

#define BZG_OK 0
#define BZG_ERR_INVALID_ARGUMENT (-1) /* bad size/pointer/leading dimension */
#define BZG_ERR_SIZE_MISMATCH (-2)    /* operand shapes do not conform */
#define BZG_ERR_HIP (-3)              /* a HIP runtime call failed */
#define BZG_ERR_OUT_OF_MEMORY (-4)    /* device allocation failed or budget exceeded */
#define BZG_ERR_NOT_SQUARE (-5)       /* solve(): non-square system matrix */
#define BZG_ERR_UNSUPPORTED (-6)      /* shape/option outside the implemented range */
#define BZG_ERR_NO_DEVICE (-7)        /* no gfx950 device visible */
#define BZG_ERR_BUSY (-8)             /* server: queue full / not admitted yet */
#define BZG_ERR_CANCELLED (-9)        /* server: request cancelled */

typedef struct bzg_context bzg_context;

/* ---- context, stream, memory ------------------------------------------------------------ */

/* Creates a context on the current HIP device with its own (default-priority) stream. */
bzg_status bzg_create(bzg_context** ctx);
bzg_status bzg_destroy(bzg_context* ctx);
/* Enqueue all later work of this context on the given hipStream_t (passed as void*);
 * NULL selects the device's default (null) stream.  The context's own stream is
 * bzg_own_stream(ctx). */
bzg_status bzg_set_stream(bzg_context* ctx, void* hip_stream);
void* bzg_get_stream(bzg_context* ctx);
void* bzg_own_stream(bzg_context* ctx);
/* Blocks until all work enqueued on the context's stream has finished. */
bzg_status bzg_synchronize(bzg_context* ctx);
const char* bzg_status_string(bzg_status s);
/* Text of the last error detected on this thread (HIP error name, failing argument ...). */
const char* bzg_last_error(void);

/* Device-time measurement of named kernels with HIP events recorded on the context stream
 * around their launches: "dgemm", "fp64acc_gemm" (the residue slice GEMMs), "fp64acc_total",
 * "csrmv", "dgesv", "dsgesv".  bzg_profile_read waits for the recorded events, returns the
 * summed milliseconds and the number of recorded launches, and clears them. */
bzg_status bzg_profile_enable(bzg_context* ctx, int32_t on);
bzg_status bzg_profile_read(bzg_context* ctx, const char* kernel, double* total_ms,
                            int64_t* launches);

/* Device memory (hipMalloc/hipFree) for hosts that do not include HIP headers. */
bzg_status bzg_malloc(void** dptr, size_t bytes);
bzg_status bzg_free(void* dptr);
/* Synchronous copies on the context stream (kind: 0 = H2D, 1 = D2H, 2 = D2D). */
bzg_status bzg_memcpy(bzg_context* ctx, void* dst, const void* src, size_t bytes, int kind);
bzg_status bzg_memset_zero(bzg_context* ctx, void* dptr, size_t bytes);

/* ---- dense GEMM:  C = alpha * A * B + beta * C  (row-major, device pointers) --------------
 * A is m x k (lda >= k), B is k x n (ldb >= n), C is m x n (ldc >= n).  beta == 0 never
 * reads C.  k == 0 sets C = beta*C (Blaze resets C: DMatDMatMultExpr.h:527-530). */
bzg_status bzg_dgemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, double alpha,
                     const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                     double* C, int64_t ldc);
bzg_status bzg_sgemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, float alpha,
                     const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                     float* C, int64_t ldc);
/* int32 GEMM: wrap-around (mod 2^32) arithmetic, bit-exact against any summation order. */
bzg_status bzg_igemm(bzg_context* ctx, int64_t m, int64_t n, int64_t k, int32_t alpha,
                     const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb,
                     int32_t beta, int32_t* C, int64_t ldc);
/* bzg_igemm over transposed operands (transa / transb as bzg_dgemm_ex): INT32 wrap-around
 * products are exact, so every form equals the product of transposed copies bit for bit.
 * bzg_igemm and bzg_igemm_ex run products with m, n >= 128 and either k >= 512
 * and m * n >= 256 * 256, or k >= 256 and m * n >= 2048 * 2048 (digit workspace <= 8 GiB), as
 * bzg_igemm_digits; the rest on the vector ALUs. */
bzg_status bzg_igemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        int32_t alpha, const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb,
                        int32_t beta, int32_t* C, int64_t ldc);
/* The same product (same arguments and result bits as bzg_igemm_ex) on the INT8 matrix cores:
 * each operand is split into 4 balanced base-256 digits, the 10 digit-pair products with
 * i + j <= 3 run as 4 INT8 GEMMs with exact INT32 sums, and C = alpha * sum_c 2^(8c) P_c +
 * beta * C is formed modulo 2^32.  One pass covers k <= 32767 (the class sums stay below
 * 2^31); longer K runs in chunks of at most 32640, accumulated with beta = 1 (exact: wrap-
 * around sums do not depend on the grouping).  k == 0 sets C = beta * C.  Uses 16 bytes of
 * workspace per element of the 256-padded C plus 4 bytes per element of each padded operand
 * chunk.  Replaces the same Blaze integer products as bzg_igemm_ex. */
bzg_status bzg_igemm_digits(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                            int32_t alpha, const int32_t* A, int64_t lda, const int32_t* B, int64_t ldb,
                            int32_t beta, int32_t* C, int64_t ldc);
/* C = alpha * op(A) * op(B) + beta * C with op(A) = A or A^T (transa != 0: A is stored k x m,
 * lda >= m) and op(B) = B or B^T (transb != 0: B is stored n x k, ldb >= k).  Replaces the
 * products over the other storage orders (DMatTDMatMultExpr.h, TDMatDMatMultExpr.h,
 * TDMatTDMatMultExpr.h): a column-major m x k matrix is a row-major k x m one.  The
 * transposed slabs are reordered on the way into LDS, so every form sums each C element in
 * the same order as bzg_dgemm on explicitly transposed copies (bit-identical results). */
bzg_status bzg_dgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        double alpha, const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc);
bzg_status bzg_sgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        float alpha, const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                        float* C, int64_t ldc);


/* ---- grouped GEMM: many independent products of uneven sizes in one launch ------------
 * No single reference function: a batch of DMatDMatMultExpr assignments
 * (DMatDMatMultExpr.h:516-569), each C_p = alpha_p A_p B_p + beta_p C_p with bzg_dgemm's
 * rules.  `problems` is a DEVICE array of `count` descriptors (so that the call can be captured
 * in a HIP graph and replayed with new descriptor contents); `status` (device, may be NULL)
 * receives 0 or -1 (descriptor rejected, product skipped) per problem.  One persistent launch
 * walks the 64 x 64 tiles of all products; each C element is summed in the same k order as
 * bzg_dgemm, so every product is bit-identical to its own bzg_dgemm call. */
typedef struct {
  int64_t m, n, k;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
  double alpha, beta;
} bzg_dgemm_problem;
bzg_status bzg_dgemm_grouped(bzg_context* ctx, int64_t count, const bzg_dgemm_problem* problems,
                             int32_t* status);

/* Tile-shape autotuner (no reference counterpart; the reference's blocking is fixed at
 * compile time, blaze/math/dense/MMM.h + Blocking.h:58-61).  Times the MFMA tile shapes
 * (0: 32 x 32, 1: 64 x 64, 2: 128 x 128, and for float 3: 128 x 256) on an m x n x k product
 * (fp32 != 0: float, else double) and records the fastest for the size class (ceil(log2) of
 * m, n, k); later bzg_dgemm / bzg_sgemm calls of that class use it.  All shapes give
 * bit-identical results.  Writes the chosen shape and its time in ms (either may be NULL).
 * Synchronises the stream. */
bzg_status bzg_gemm_autotune(bzg_context* ctx, int32_t fp32, int64_t m, int64_t n, int64_t k,
                             int32_t* shape, float* ms);
/* Forgets every tuned size class (back to the built-in heuristic). */
bzg_status bzg_gemm_tuning_clear(void);

/* ---- dense GEMV:  y = alpha * A * x + beta * y  (row-major A, device pointers) ---------- */
/* B = A^T for a row-major m x n A into the n x m B (ldb >= m), not in place.  No reference
 * function of its own: the storage-order change lu() makes when it hands a row-major matrix
 * to the column-major getrf (blaze/math/dense/LU.h:98-140, lapack/getrf.h). */
bzg_status bzg_dtranspose(bzg_context* ctx, int64_t m, int64_t n, const double* A, int64_t lda, double* B,
                          int64_t ldb);
/* y = alpha * A^T * x + beta * y for the row-major m x n A (x: m entries, y: n entries): the
 * product with a column-major matrix (TDMatDVecMultExpr, blaze/math/expressions/
 * TDMatDVecMultExpr.h; cblas_dgemv with a transpose).  Deterministic two-stage column sums.
 * beta == 0 never reads y.  Scratch is cached in the context. */
bzg_status bzg_dgemv_t(bzg_context* ctx, int64_t m, int64_t n, double alpha, const double* A, int64_t lda,
                       const double* x, double beta, double* y);
bzg_status bzg_sgemv_t(bzg_context* ctx, int64_t m, int64_t n, float alpha, const float* A, int64_t lda,
                       const float* x, float beta, float* y);
bzg_status bzg_igemv_t(bzg_context* ctx, int64_t m, int64_t n, int32_t alpha, const int32_t* A, int64_t lda,
                       const int32_t* x, int32_t beta, int32_t* y);
bzg_status bzg_dgemv(bzg_context* ctx, int64_t m, int64_t n, double alpha, const double* A,
                     int64_t lda, const double* x, double beta, double* y);
bzg_status bzg_sgemv(bzg_context* ctx, int64_t m, int64_t n, float alpha, const float* A,
                     int64_t lda, const float* x, float beta, float* y);
bzg_status bzg_igemv(bzg_context* ctx, int64_t m, int64_t n, int32_t alpha, const int32_t* A,
                     int64_t lda, const int32_t* x, int32_t beta, int32_t* y);

/* y += alpha * x over n contiguous elements (device pointers; INT32 wraps modulo 2^32).  The
 * element-wise update of an accumulation whose operand aliases its target: the product is
 * evaluated into a temporary and added (blaze/math/dense/DynamicMatrix.h:1670-1672,
 * blaze/math/dense/DynamicVector.h:1412-1414). */
bzg_status bzg_daxpy(bzg_context* ctx, int64_t n, double alpha, const double* x, double* y);
bzg_status bzg_saxpy(bzg_context* ctx, int64_t n, float alpha, const float* x, float* y);
bzg_status bzg_iaxpy(bzg_context* ctx, int64_t n, int32_t alpha, const int32_t* x, int32_t* y);

/* ---- CSR SpMV:  y = A * x  (row-major CompressedMatrix packed as CSR) --------------------
 * rowptr has n_rows+1 int32 entries (rowptr[0] == 0, nnz = rowptr[n_rows] < 2^31); colidx
 * int32 strictly increasing per row.  Every row's sum runs sequentially in storage order
 * (the reference order, SVecDVecInnerExpr.h:127-135) or as a fixed-shape tree; results are
 * deterministic run to run.
 *
 * A plan analyses the row-length histogram once (host side, O(n_rows)) and picks the kernels:
 * rows of up to 32 nonzeros are summed sequentially, either in nnz-balanced row blocks staged
 * through LDS (2048 or 4096 nonzeros per block) or by lane groups (2..16 lanes per row);
 * rows of 33..256 nonzeros get a wavefront each, longer rows a workgroup (fixed-shape tree
 * sums).  The plan owns device metadata; the matrix arrays stay owned by the caller. */
typedef struct bzg_csr_plan bzg_csr_plan;
bzg_status bzg_csr_plan_create(bzg_context* ctx, int64_t n_rows, int64_t n_cols,
                               const int32_t* rowptr_host, bzg_csr_plan** plan);
/* positive_zero != 0: every y entry is the sum started from +0 (column-major products,
 * TSMatDVecMultExpr's reset-then-accumulate, blaze/math/expressions/TSMatDVecMultExpr.h:329-383)
 * instead of from the first product (SMatDVecMultExpr, SVecDVecInnerExpr.h:127-135).  The two
 * differ only in the sign of an all-zero sum; the kernels take the start value, so the cost is
 * the same. */
bzg_status bzg_csr_plan_set_sum_start(bzg_csr_plan* plan, int32_t positive_zero);
/* Host CSC -> CSR pack (a column-major CompressedMatrix, CompressedMatrix.h:3118, into the
 * layout the device products read): colptr (n_cols + 1), rowidx, val (elem_bytes 4 or 8) in;
 * rowptr (n_rows + 1), colidx, val_out (colptr[n_cols] entries) out.  Within every row the
 * columns come out ascending, so y = A x over the result sums each row in column order -- the
 * reference's order for a column-major A.  The CSR of trans(S) for a row-major S is this pack
 * applied to S's CSR read as CSC. */
bzg_status bzg_csc_to_csr_host(int64_t n_rows, int64_t n_cols, const int32_t* colptr, const int32_t* rowidx,
                               const void* val, int32_t elem_bytes, int32_t* rowptr, int32_t* colidx,
                               void* val_out);
/* ---- Matrix Market input (host) ------------------------------------------------------------
 * No reference function (the reference points to an external Matrix Market project for its
 * sparse inputs).  Reads a "coordinate" file -- real / double / integer / pattern values (pattern
 * entries are 1.0), general / symmetric / skew-symmetric (expanded to both triangles) -- into a
 * host CSR: rows in order, columns ascending, duplicate entries summed in file order.  The
 * arrays are malloc'ed by the library; bzg_csr_host_free releases them.  Errors: unreadable
 * file, bad header or entry, index out of range, fewer entries than declared
 * (BZG_ERR_INVALID_ARGUMENT); dense "array" files, complex / hermitian fields, sizes >= 2^31
 * (BZG_ERR_UNSUPPORTED). */
typedef struct {
  int64_t n_rows, n_cols, nnz;
  int32_t* rowptr; /* n_rows + 1 */
  int32_t* colidx; /* nnz */
  double* val;     /* nnz */
} bzg_csr_host;
bzg_status bzg_mm_read_csr(const char* path, bzg_csr_host* out);
bzg_status bzg_csr_host_free(bzg_csr_host* matrix);

bzg_status bzg_csr_plan_destroy(bzg_csr_plan* plan);
/* Sliced-ELL copy of an all-short-row matrix (every row <= 32 nonzeros; BZG_ERR_UNSUPPORTED
 * otherwise) held by the plan: slices of 64 consecutive rows stored slot by slot, padding slots
 * with column -1.  dtype: 0 double, 1 float, 2 int32; rowptr / colidx / val are the device CSR
 * arrays of the matrix the plan was made for (copied, not kept).  Then y = S * x with
 * bzg_dsellmv / bzg_ssellmv / bzg_isellmv reads only the copy and x: bit-identical to
 * bzg_*csrmv (every row is the sequential sum in storage order, SVecDVecInnerExpr.h:127-135),
 * with coalesced value and index loads and no row pointers.  Pack again after the values
 * change.  Memory: (4 + sizeof element) bytes per slot. */
bzg_status bzg_csr_plan_pack_sell(bzg_context* ctx, bzg_csr_plan* plan, const int32_t* rowptr, const int32_t* colidx,
                                  const void* val, int32_t dtype);
bzg_status bzg_dsellmv(bzg_context* ctx, const bzg_csr_plan* plan, const double* x, double* y);
bzg_status bzg_ssellmv(bzg_context* ctx, const bzg_csr_plan* plan, const float* x, float* y);
bzg_status bzg_isellmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* x, int32_t* y);
/* Overrides the vector-CSR shape chosen from the histogram: lanes per row in {1,2,4,8,16,32},
 * rows per lane group in {1,2}, nonzeros loaded per lane per batch in {1,2,4,8}; used by the
 * autotuner.  Results stay bit-identical (every row is still summed in storage order). */
bzg_status bzg_csr_plan_tune(bzg_csr_plan* plan, int32_t lanes, int32_t rows_per_group,
                             int32_t cap);
/* Autotuner: times the short-row kernel configurations (row blocks of 2048 / 4096 nonzeros
 * with cached or non-temporal loads; lane groups of 2..16 lanes, 1-2 rows, 4-8 loads in
 * flight) on this matrix (fp32 != 0: float values / x / y, else double) and keeps the fastest
 * in the plan; y is overwritten.  Results are identical for every configuration.  Writes the
 * best time in ms (may be NULL).  Synchronises the stream. */
bzg_status bzg_csr_plan_autotune(bzg_context* ctx, bzg_csr_plan* plan, int32_t fp32,
                                 const int32_t* rowptr, const int32_t* colidx, const void* val,
                                 const void* x, void* y, float* ms);
/* Describes the plan: writes a short text (kernel choice, lanes per row, blocks) to buf. */
bzg_status bzg_csr_plan_describe(const bzg_csr_plan* plan, char* buf, size_t len);
bzg_status bzg_dcsrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                      const int32_t* colidx, const double* val, const double* x, double* y);
bzg_status bzg_scsrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                      const int32_t* colidx, const float* val, const float* x, float* y);
bzg_status bzg_icsrmv(bzg_context* ctx, const bzg_csr_plan* plan, const int32_t* rowptr,
                      const int32_t* colidx, const int32_t* val, const int32_t* x, int32_t* y);

/* ---- conjugate gradients on a resident CSR matrix ----------------------------------------
 * The loop of the reference's CG benchmark (blazemark/src/blaze/CG.cpp:114-135):
 *   r = A x - b;  delta = r.r;  d = -r;  then `iterations` times:
 *   h = A d;  alpha = delta / (d.h);  x += alpha d;  r += alpha h;  beta = r.r;
 *   d = (beta / delta) d - r;  delta = beta.
 * A is n x n CSR (device arrays) with a plan from bzg_csr_plan_create; b and x are device
 * vectors (x: start in, result out).  The products are bzg_dcsrmv's; dot products are
 * deterministic (fixed partial-sum order), and no scalar leaves the device inside the loop, so
 * the call can be captured in a HIP graph.  rr_out (host, may be NULL) receives r.r after the
 * last iteration (one stream synchronisation).  Scratch (3 n doubles) is cached in the context. */
bzg_status bzg_dcg(bzg_context* ctx, const bzg_csr_plan* plan, int64_t n, const int32_t* rowptr,
                   const int32_t* colidx, const double* val, const double* b, double* x, int32_t iterations,
                   double* rr_out);

/* ---- batched dense solves (LAPACK dgesv / dsgesv semantics per system) ------------------
 * `count` independent systems A_i X_i = B_i of order n[i] with nrhs[i] right-hand sides.
 * All per-system arrays are DEVICE arrays.  A_i is stored row-major n x n at
 * A + a_off[i] (leading dimension n[i]); B_i / X_i are stored row-major n x nrhs at
 * B + b_off[i] (leading dimension nrhs[i]).  ipiv_i (1-based, LAPACK convention) at
 * ipiv + p_off[i] (n[i] entries).  info[i] = 0 or the 1-based index of an exactly-zero
 * pivot.  Row-major storage means the factorisation is of A itself (not of A^T), with
 * partial pivoting by rows exactly as dgetf2: the first row of maximal |a|.
 * 1 <= n[i] <= 512, 0 <= nrhs[i] <= 8; a system outside that range gets info[i] = -1.
 * Orders up to 256 run on persistent workgroups (up to 4 resident per CU; the grid size is
 * set by bzg_solver_autotune, default 8 per CU), orders 257..512 on a second kernel instance
 * with twice the LDS (2 per CU).
 * `order` (device, may be NULL) lists the systems in the order the persistent workgroups
 * should take them -- bzg_batched_order() gives longest-first; NULL = index order.
 * ipiv may be NULL (then p_off is ignored). */
bzg_status bzg_dgesv_batched(bzg_context* ctx, int64_t count, const int32_t* n,
                             const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                             const int64_t* p_off, double* A, double* B, int32_t* ipiv,
                             int32_t* info, const int32_t* order);
/* No reference counterpart: bzg_dgesv_batched for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results (bit for bit) as bzg_dgesv_batched; it calls no hipMalloc,
 * hipFree or allocating routine and does not synchronise.  Its scratch (work counters and the
 * list of systems of order > 256: 256 B + 4 B per system) is a fixed device array of
 * 256 B + 4 * BZG_NOALLOC_MAX_SYSTEMS bytes (about 4 MiB) that the library holds on every GPU it
 * runs on until the process ends; count > BZG_NOALLOC_MAX_SYSTEMS returns BZG_ERR_UNSUPPORTED.
 * Every call on one GPU shares that array (counters are cleared by a kernel at the start of each
 * call), so calls that overlap on one GPU (several streams or host threads) race: serialise
 * them, e.g. on one stream. */
#define BZG_NOALLOC_MAX_SYSTEMS 1048576
bzg_status bzg_dgesv_batched_noalloc(bzg_context* ctx, int64_t count, const int32_t* n,
                                     const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                                     const int64_t* p_off, double* A, double* B, int32_t* ipiv,
                                     int32_t* info, const int32_t* order);
/* Factorisation only (LAPACK dgetrf per system, replaces getrf, clapack/getrf.h:190-197):
 * A_i := L U in place (unit-lower L below the diagonal), ipiv_i 1-based, info as above.
 * The factors and pivots are exactly those bzg_dgesv_batched computes for the same A_i. */
bzg_status bzg_dgetrf_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                              const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info,
                              const int32_t* order);
/* Determinants (replaces det(), blaze/math/expressions/DMatDetExpr.h:384-424, whose orders >= 7
 * take detNxN, :324-357): factors every A_i in place exactly as bzg_dgetrf_batched (ipiv and
 * info as there), then det[i] = product over k of U_i(k,k), negated for every k with
 * ipiv_i[k] != k + 1, multiplied left to right from 1; 0 when info[i] > 0.  n[i] = 0 gives
 * det 0 and info 0 (the reference's 0 x 0 case); an order outside 0..512 gives info -1 and NaN.
 * The reference's closed forms for orders 2..6 agree with this to rounding. */
bzg_status bzg_ddet_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                            const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info, double* det,
                            const int32_t* order);
/* Inverses in place (replaces invert()/inv() by LU, blaze/math/dense/Inversion.h:3337-3347:
 * getrf then getri): A_i := A_i^-1, computed as the factorisation of bzg_dgetrf_batched
 * (ipiv, info as there) followed by dgetrs solves against the identity: every eight-column
 * chunk of every system is one system of a single batched dgetrs launch (longest first).
 * A system with info[i] > 0 (singular; the reference throws "Inversion of singular matrix
 * failed") keeps its LU factors in A_i; n[i] = 0 gives info 0.  Reads n, a_off and p_off back
 * to the host (synchronisations) and needs scratch of two copies of the factors' size (the
 * factors and the right-hand sides) plus 44 bytes per chunk (cached in the context). */
bzg_status bzg_dinv_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int64_t* a_off,
                            const int64_t* p_off, double* A, int32_t* ipiv, int32_t* info,
                            const int32_t* order);
/* Solve with factors (LAPACK dgetrs per system, replaces getrs, clapack/getrs.h:207-215):
 * LU / ipiv as written by bzg_dgetrf_batched; the interchanges are applied to B_i in pivot
 * order, then B_i := U^-1 L^-1 B_i in place.  info[i] = 0, or -1 outside the size range.
 * A zero on U's diagonal is not checked (as in dgetrs).  Same result, bit for bit, as the
 * solve phase of bzg_dgesv_batched. */
bzg_status bzg_dgetrs_batched(bzg_context* ctx, int64_t count, const int32_t* n, const int32_t* nrhs,
                              const int64_t* a_off, const int64_t* b_off, const int64_t* p_off,
                              const double* LU, const int32_t* ipiv, double* B, int32_t* info,
                              const int32_t* order);
/* Mixed precision: FP32 LU with partial pivoting + FP64 iterative refinement.  A and B are
 * read only; X (same layout as B) receives the solution.  Per system (LAPACK dsgesv):
 * iter[i] >= 0  converged after iter[i] refinement steps;
 * iter[i] <  0  fell back to the FP64 LU:  -2 FP32 factorisation hit a zero pivot,
 *               -3 refinement did not converge in max_iter steps (-31 in LAPACK).
 * Stopping test per RHS column: ||r||_inf <= ||x||_inf * ||A||_inf * eps * sqrt(n) with
 * eps = tol (pass 0 to use LAPACK's 2^-53).  Refinement also stops (-> FP64) when the
 * largest residual fails to halve on two consecutive steps.  info[i] as in
 * bzg_dgesv_batched.  Orders 1..256 are factored with the FP32 matrix held on chip (one
 * multi-wavefront workgroup per system, in four order classes), orders 257..512 by a streamed
 * FP32 kernel; the FP64 fallback runs after them in the same call.  Scratch (class lists,
 * counters, the streamed kernel's FP32 matrices, the fallback's FP64 matrices) is cached in
 * the context. */
bzg_status bzg_dsgesv_batched(bzg_context* ctx, int64_t count, const int32_t* n,
                              const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                              const double* A, const double* B, double* X, int32_t* iter,
                              int32_t* info, double tol, int32_t max_iter, const int32_t* order);
/* Autotuner of the batched solvers' persistent grids (no reference counterpart): times
 * 1, 2, 3, 4, 6 and 8 workgroups per CU on the given batch and keeps the fastest for later
 * calls: the grid of bzg_dgesv_batched (mixed == 0), or of bzg_dsgesv_batched's FP64 fallback
 * kernel (mixed != 0; its on-chip kernels always run at full residency: 12, 6, 2 and 1
 * workgroups per CU for the orders 1..32, 33..64, 65..128, 129..256).  A and B are not modified
 * (the FP64 solver runs on copies; a_elems / b_elems are the pool sizes in doubles).  Results
 * do not depend on the grid.  Writes the choice and its time in ms (either may be NULL).
 * Synchronises the stream. */
bzg_status bzg_solver_autotune(bzg_context* ctx, int32_t mixed, int64_t count, const int32_t* n,
                               const int32_t* nrhs, const int64_t* a_off, const int64_t* b_off,
                               const double* A, int64_t a_elems, const double* B, int64_t b_elems,
                               const int32_t* order, int32_t* wg_per_cu, float* ms);
/* Longest-first work order for the batched solvers: writes to order_dev (device, count
 * entries) the system indices sorted by descending n (stable), computed from n_host. */
bzg_status bzg_batched_order(bzg_context* ctx, int64_t count, const int32_t* n_host,
                             int32_t* order_dev);

/* ---- single-system dense LU (any order up to 32768, any number of right-hand sides) -------
 * A is a row-major n x n DEVICE matrix (leading dimension lda), B row-major n x nrhs (ldb),
 * ipiv (n entries) and info (one entry) DEVICE arrays.  P A = L U as LAPACK dgetrf computes it
 * for A: ipiv 1-based in interchange order, info = 0 or the 1-based column of the first exactly
 * zero pivot (the factorisation completes).  The products inside run on bzg_dgemm; workspace
 * (n ints + n * nrhs doubles for the solves) is cached in the context.  Orders above 32768 return
 * BZG_ERR_UNSUPPORTED. */
/* dgetrf: A := L \ U in place (replaces getrf, blaze/math/lapack/clapack/getrf.h:190-197 and
 * the factorisation under lu(), blaze/math/dense/LU.h:98-140). */
bzg_status bzg_dgetrf(bzg_context* ctx, int64_t n, double* A, int64_t lda, int32_t* ipiv, int32_t* info);
/* dgetrs ('N'): B := A^-1 B from bzg_dgetrf's factors (replaces getrs, clapack/getrs.h:207-215). */
bzg_status bzg_dgetrs(bzg_context* ctx, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                      const int32_t* ipiv, double* B, int64_t ldb);
/* dgesv = dgetrf + dgetrs (replaces solve() for one system: blaze/math/dense/LSE.h:6343-6373 ->
 * solveNxN LSE.h:5229-5257 -> dgesv_, clapack/gesv.h:207).  B is left unchanged when info > 0;
 * reading info back synchronises the context stream once. */
bzg_status bzg_dgesv(bzg_context* ctx, int64_t n, int64_t nrhs, double* A, int64_t lda, int32_t* ipiv,
                     double* B, int64_t ldb, int32_t* info);
/* A := A^-1 from bzg_dgetrf's factors (info must have been 0): the inverse by LU of invert()
 * (blaze/math/dense/Inversion.h:3337-3347), as the solution of A X = I. */
bzg_status bzg_dgetri(bzg_context* ctx, int64_t n, double* A, int64_t lda, const int32_t* ipiv);
/* *det = det(A) from bzg_dgetrf's factors by detNxN's rule (blaze/math/expressions/
 * DMatDetExpr.h:324-357): 0 when info > 0, else the product, left to right from 1, of U's
 * diagonal with the sign flipped at every interchange (det is a DEVICE pointer). */
bzg_status bzg_ddet_lu(bzg_context* ctx, int64_t n, const double* A, int64_t lda, const int32_t* ipiv,
                       const int32_t* info, double* det);

/* ---- complex element types (interleaved re, im: std::complex<T>, C99 T _Complex) -----------
 * Pointers are to the first real of the first element; leading dimensions count complex
 * elements; alpha and beta point to (re, im) pairs in host memory.  trans codes: 0 = N,
 * 1 = T (transpose), 2 = C (conjugate transpose), as CBLAS_TRANSPOSE.
 *
 * C = alpha op(A) op(B) + beta C, row-major, m x n x k (replaces cblas_zgemm / cblas_cgemm,
 * blaze/math/blas/cblas/gemm.h:197-283, the BLAS path of DMatDMatMultExpr for complex
 * elements).  op(A) is m x k: A stored m x k (N) or k x m (T, C); op(B) k x n likewise.
 * beta = 0 never reads C.  Runs on the real MFMA GEMM through the 2 x 2 real form of the
 * product (8 m n k flops); uses 32 k n bytes of workspace (complex<double>; half for float),
 * plus 16 m k for a transposed A. */
bzg_status bzg_zgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        const double* alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                        const double* beta, double* C, int64_t ldc);
bzg_status bzg_cgemm_ex(bzg_context* ctx, int32_t transa, int32_t transb, int64_t m, int64_t n, int64_t k,
                        const float* alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                        const float* beta, float* C, int64_t ldc);
/* y = alpha op(A) x + beta y, A stored m x n row-major (replaces cblas_zgemv / cblas_cgemv,
 * blaze/math/blas/cblas/gemv.h:186-245, for DMatDVecMultExpr with complex elements): op = N
 * takes x of n and y of m entries, op = T / C x of m and y of n.  beta = 0 never reads y. */
bzg_status bzg_zgemv_ex(bzg_context* ctx, int32_t trans, int64_t m, int64_t n, const double* alpha,
                        const double* A, int64_t lda, const double* x, const double* beta, double* y);
bzg_status bzg_cgemv_ex(bzg_context* ctx, int32_t trans, int64_t m, int64_t n, const float* alpha,
                        const float* A, int64_t lda, const float* x, const float* beta, float* y);
/* Solves A X = B for one complex<double> system of order n <= 16384 (replaces solve() of a
 * complex system: gesv -> zgesv_, blaze/math/lapack/clapack/gesv.h:317-333).  B (n x nrhs, ldb)
 * is overwritten with X; A is not modified.  Solved as the real 2n x 2n system
 * [[Re A, -Im A], [Im A, Re A]] [Re X; Im X] = [Re B; Im B] by bzg_dgesv, so there are no complex
 * pivots to return (no ipiv).  *info (DEVICE pointer): 0, or > 0 when A is singular (an index
 * of the real system's elimination, <= 2n; B is left unchanged then).  Workspace: 32 n^2 +
 * 16 n nrhs bytes. */
bzg_status bzg_zgesv(bzg_context* ctx, int64_t n, int64_t nrhs, const double* A, int64_t lda, double* B,
                     int64_t ldb, int32_t* info);
/* B = A^T for a row-major m x n complex<double> A into the n x m B (ldb >= m), not in place (no
 * conjugation).  No reference function of its own: the storage-order change of a column-major
 * complex system matrix before bzg_zgesv, as bzg_dtranspose for double. */
bzg_status bzg_ztranspose(bzg_context* ctx, int64_t m, int64_t n, const double* A, int64_t lda, double* B,
                          int64_t ldb);

/* ---- FP64-accurate GEMM from BF16 residue slices (Ozaki scheme II on MFMA) --------------
 * Same contract as bzg_dgemm with alpha = 1, beta = 0.  `moduli` selects the number of
 * coprime moduli (slice GEMMs), 2..24.  0 = automatic: row / column maxima and sums of squares
 * of A and B give an estimate of the relative Frobenius error for each count,
 * 2^-P sqrt((sum_i D_i^2 ||B||_F^2 + sum_j E_j^2 ||A||_F^2) / (12 sum_q ||A_:q||^2 ||B_q:||^2))
 * with P the bits per operand the count allows and D_i (E_j) the power of two above row i's
 * (column j's) largest |entry|; the smallest count whose estimate is <= 8e-15 is used (24 if
 * none is), or <= 3.2e-15 when single entries dominate their rows or columns (k sum_i D_i^2 >
 * 12 ||A||_F^2 or the same for B), where the realised error scatters more widely around the
 * estimate.  This reads 5 numbers back to the host (one stream synchronisation), so capture a
 * graph with an explicit count.  A NaN or an infinity in A or B (seen by that pass) makes the
 * call the plain FP64 GEMM (bzg_dgemm), so such values propagate exactly as there; with an
 * explicit count the rows of A and columns of B that hold them are recomputed in FP64 on the
 * device.  Workspace is allocated from the context on first use and cached. */
bzg_status bzg_dgemm_fp64acc(bzg_context* ctx, int64_t m, int64_t n, int64_t k,
                             const double* A, int64_t lda, const double* B, int64_t ldb,
                             double* C, int64_t ldc, int32_t moduli);
/* Same contract and method with INT8 residues (moduli <= 255) on the INT8 matrix cores with
 * exact INT32 accumulation (k <= 2^17); roughly twice the slice-GEMM rate of BF16. */
bzg_status bzg_dgemm_fp64acc_i8(bzg_context* ctx, int64_t m, int64_t n, int64_t k,
                                const double* A, int64_t lda, const double* B, int64_t ldb,
                                double* C, int64_t ldc, int32_t moduli);
/* Plan of the last FP64-accurate GEMM on this context: moduli count, operand bits P and the
 * error estimate used by the automatic choice (0 for an explicit count); any may be NULL. */
bzg_status bzg_fp64acc_last_plan(bzg_context* ctx, int32_t* moduli, int32_t* bits, double* est_rel_err);
/* Number of BF16 slice GEMMs that bzg_dgemm_fp64acc runs for this k and an explicit moduli
 * count (0: the count giving P >= 50 bits, the choice without operand statistics). */
int32_t bzg_fp64acc_slices(int64_t k, int32_t moduli);
/* Bits P kept per scaled operand (|A'|,|B'| <= 2^P) for this k / moduli choice; the
 * elementwise error is bounded by k * 2^-P * max|A_i,:| * max|B_:,j|.  -1 if invalid. */
int32_t bzg_fp64acc_operand_bits(int64_t k, int32_t moduli, int32_t int8_slices);


/* ---- continuous-batching server -----------------------------------------------------------
 * No reference counterpart (the reference evaluates one expression per call); it serves the
 * same three operations to many concurrent clients.  Operands are DEVICE pointers that stay
 * valid until the request completes; CSR matrices are uploaded once (bzg_server_add_csr) and
 * stay resident.  Results (C, y or X, row-major, ld = n / 1 / nrhs) are written to server-owned
 * HBM and stay valid until bzg_release.  A request is admitted when fewer than max_inflight
 * launches are running and its result (plus staging) fits in hbm_budget_bytes together with
 * everything not yet released and the resident matrices; admission is FIFO, and all admissible
 * solve requests waiting behind the head join its batch (one bzg_dsgesv_batched launch, FP32 LU +
 * FP64 refinement to LAPACK dsgesv's test, FP64 fallback); likewise GEMMs of up to 2^31 flops
 * (one bzg_dgemm_grouped launch).  Work is spread over `streams` HIP
 * streams so that requests overlap on the device.  Each result is bit-identical to the same
 * request run alone.  All functions are thread-safe. */
typedef struct bzg_server bzg_server;
typedef struct {
  size_t hbm_budget_bytes; /* admission budget: unreleased results + staging + resident CSR */
  int32_t max_inflight;    /* launches on the device at once (>= 1) */
  int32_t streams;         /* HIP streams / library contexts used by the server (1..16) */
  int32_t solve_batch_max; /* largest number of solve (or GEMM) requests fused into one launch */
} bzg_server_config;
typedef struct {
  bzg_status status;      /* launch status (0 ok) */
  int32_t kind;           /* 0 GEMM, 1 SpMV, 2 solve */
  int32_t info;           /* solve: LAPACK info (0 or first zero pivot; -1 out of range) */
  int32_t iterations;     /* solve: refinement steps (>= 0) or < 0 when the FP64 LU was used */
  double* result;         /* device pointer of the result */
  int64_t result_elems;
  double t_submit_us, t_admit_us, t_done_us; /* steady-clock microseconds */
} bzg_request_result;
typedef struct {
  int64_t submitted, admitted, completed, released;
  int64_t launches, solve_batches, max_solve_batch, gemm_batches, max_gemm_batch, inflight;
  int64_t bytes_in_use, bytes_peak, budget;
} bzg_server_stats;

bzg_status bzg_server_create(const bzg_server_config* cfg, bzg_server** server);
/* Waits for every admitted request, then frees all server memory (results included). */
bzg_status bzg_server_destroy(bzg_server* server);
/* Uploads a double CSR matrix (host arrays, same format as bzg_dcsrmv) and builds its plan;
 * it counts against the budget for the server's lifetime.  A matrix whose rows all hold <= 32
 * nonzeros is kept as the plan's sliced-ELL copy instead (bzg_csr_plan_pack_sell; same results):
 * both forms are reserved while the copy is made, then only the copy's bytes. */
bzg_status bzg_server_add_csr(bzg_server* server, int64_t n_rows, int64_t n_cols,
                              const int32_t* rowptr, const int32_t* colidx, const double* val,
                              int32_t* handle);
/* C = A * B (A m x k, B k x n, device, row-major with leading dimensions). */
bzg_status bzg_submit_gemm(bzg_server* server, int64_t m, int64_t n, int64_t k, const double* A,
                           int64_t lda, const double* B, int64_t ldb, int64_t* id);
/* y = S * x for a resident matrix (x device, n_cols entries). */
bzg_status bzg_submit_spmv(bzg_server* server, int32_t csr, const double* x, int64_t* id);
/* X = A^-1 B (A n x n, B n x nrhs, device, row-major, ld n / nrhs), 1 <= n <= 512,
 * 1 <= nrhs <= 8. */
bzg_status bzg_submit_solve(bzg_server* server, int32_t n, int32_t nrhs, const double* A,
                            const double* B, int64_t* id);
/* hold != 0: requests are still accepted and queued but none is admitted until a call with
 * hold = 0, so a burst submitted in between is batched as if it had arrived at once (all of its
 * solves, and all of its small GEMMs, join one launch each as far as the budget allows).  A
 * wait on a held request returns only after another thread lifts the hold; destroy lifts it. */
bzg_status bzg_server_hold(bzg_server* server, int32_t hold);
bzg_status bzg_poll(bzg_server* server, int64_t id, int32_t* done);
/* Blocks until the request has completed and describes it. */
bzg_status bzg_wait(bzg_server* server, int64_t id, bzg_request_result* out);
/* Frees the request's result (its bytes return to the budget); BZG_ERR_BUSY while running. */
bzg_status bzg_release(bzg_server* server, int64_t id);
bzg_status bzg_server_get_stats(bzg_server* server, bzg_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* BLAZE_GPU_H */
