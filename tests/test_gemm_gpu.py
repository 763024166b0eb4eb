"""GPU parity: dense GEMM (C = A * B) against the CPU oracle and the reference's known answers.

Mirrors blazetest's dmatdmatmult operation tests (blazetest/src/mathtest/operations/
dmatdmatmult/MDaMDa.cpp:77-92 size grid; OperationTest.h checks with relaxed equality) and the
known answer of matrices/dynamicmatrix/ClassTest2.cpp:1244-1273.
Tolerances: int32 and integer-valued doubles bit-exact; random doubles within Blaze's relaxed
equality (|a-b| <= 1e-8 * max(1,|a|), Equal.h:214-224) and within the Higham bound
gamma_k * (|A||B|)_ij of the exact product; float within 1e-6 relative.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, dev, empty, full, have_gpu, host, p

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

FN = {np.float64: "bzg_dgemm", np.float32: "bzg_sgemm", np.int32: "bzg_igemm"}


def gpu_gemm(A, B, alpha=1, beta=0, Cinit=None, lda=None, ldb=None, ldc=None):
    dt = A.dtype.type
    m, k = A.shape
    n = B.shape[1]
    lda = lda or max(k, 1); ldb = ldb or max(n, 1); ldc = ldc or max(n, 1)
    Ap = np.zeros((m, lda), A.dtype); Ap[:, :k] = A
    Bp = np.zeros((k, ldb), A.dtype); Bp[:, :n] = B
    dA = dev(Ap) if m * lda else empty((1,), A.dtype)
    dB = dev(Bp) if k * ldb else empty((1,), A.dtype)
    if Cinit is not None:
        Cp = np.zeros((m, ldc), A.dtype); Cp[:, :n] = Cinit
        dC = dev(Cp)
    else:
        dC = full((max(m, 1), ldc), A.dtype, float("nan") if dt != np.int32 else -7)
    ctx().call(FN[dt], m, n, k, alpha, p(dA), lda, p(dB), ldb, beta, p(dC), ldc)
    return host(dC)[:m, :n]


def relaxed_equal(a, b, eps):
    return np.all(np.abs(a - b) <= eps * np.maximum(1.0, np.abs(b)))


def test_known_answer_classtest2():
    # ClassTest2.cpp:1244-1273: (1 0 2; 0 3 0; 4 0 5) * (0 2 0 0; 1 3 0 4; 0 0 0 5)
    A = np.array([[1, 0, 2], [0, 3, 0], [4, 0, 5]])
    B = np.array([[0, 2, 0, 0], [1, 3, 0, 4], [0, 0, 0, 5]])
    expect = np.array([[0, 2, 0, 10], [3, 9, 0, 12], [0, 8, 0, 25]])
    for dt in (np.int32, np.float64, np.float32):
        got = gpu_gemm(A.astype(dt), B.astype(dt))
        assert np.array_equal(got, expect.astype(dt)), dt


def test_aliasing_fixture_products():
    # AliasingTest.cpp:1803-1850: dA3x4 * dB4x3
    A = np.array([[-1, 0, -2, 0], [0, 2, -3, 1], [0, 1, 2, 2]])
    B = np.array([[1, 0, -3], [0, -1, 0], [0, 2, 1], [2, 1, -2]])
    for dt in (np.int32, np.float64):
        assert np.array_equal(gpu_gemm(A.astype(dt), B.astype(dt)), (A @ B).astype(dt))


SIZES = list(range(0, 7)) + [15, 16, 32, 37, 63, 64]


@pytest.mark.parametrize("dt", [np.int32, np.float64])
def test_blazetest_size_grid(dt):
    rng = np.random.default_rng(7)
    for m in SIZES:
        for n in SIZES:
            for k in SIZES:
                if dt == np.int32:
                    A = rng.integers(-10, 11, (m, k)).astype(np.int32)
                    B = rng.integers(-10, 11, (k, n)).astype(np.int32)
                else:
                    A = rng.uniform(-1, 1, (m, k)); B = rng.uniform(-1, 1, (k, n))
                got = gpu_gemm(A, B)
                ref = oracle.gemm(A, B)
                if dt == np.int32:
                    assert np.array_equal(got, ref), (m, n, k)
                else:
                    assert relaxed_equal(got, ref, 1e-8), (m, n, k)


def test_float_grid():
    rng = np.random.default_rng(8)
    for (m, n, k) in [(1, 1, 1), (5, 3, 7), (37, 63, 15), (64, 64, 64), (130, 70, 257)]:
        A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
        B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
        got = gpu_gemm(A, B)
        ref = A.astype(np.float64) @ B.astype(np.float64)
        bound = 2 * k * 2.0 ** -24 * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)) + 1e-30
        assert np.all(np.abs(got - ref) <= bound), (m, n, k)


def test_integer_valued_doubles_bit_exact_all_tiles():
    # Integer entries in [-10,10]: every partial sum is an integer < 2^53, so every summation
    # order is exact -> bit-exact against the oracle regardless of the tile shape chosen.
    rng = np.random.default_rng(9)
    for (m, n, k) in [(512, 512, 512), (1000, 1100, 300), (2048, 2048, 64), (129, 257, 1031)]:
        A = rng.integers(-10, 11, (m, k)).astype(np.float64)
        B = rng.integers(-10, 11, (k, n)).astype(np.float64)
        assert np.array_equal(gpu_gemm(A, B), A @ B), (m, n, k)


def test_config1_512_mt19937_64():
    # 512^3, A then B uniform[-1,1) from one MT19937-64 stream (seed 128753984).
    import blaze_amd as bz
    m = n = k = 512
    A = np.empty((m, k)); B = np.empty((k, n))
    bz.gen().bzg_gen_uniform_mt(C.c_uint64(128753984), C.c_int64(A.size), A.ctypes.data_as(C.c_void_p),
                                C.c_int64(B.size), B.ctypes.data_as(C.c_void_p))
    got = gpu_gemm(A, B)
    ref = oracle.gemm(A, B)
    assert relaxed_equal(got, ref, 1e-8)
    # Higham: |C_hat - C| <= gamma_k |A||B|; check against the double-double exact entries.
    rows = np.arange(0, 512, 7); cols = (rows * 13) % 512
    hi, lo = oracle.gemm_dd_entries(A, B, rows, cols)
    absAB = np.abs(A) @ np.abs(B)
    gam = k * 2.0 ** -53 / (1 - k * 2.0 ** -53)
    assert np.all(np.abs(got[rows, cols] - hi) <= gam * absAB[rows, cols])
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel < 1e-14


def test_alpha_beta_and_padding():
    rng = np.random.default_rng(10)
    m, n, k = 70, 90, 33
    A = rng.uniform(-1, 1, (m, k)); B = rng.uniform(-1, 1, (k, n)); C0 = rng.uniform(-1, 1, (m, n))
    got = gpu_gemm(A, B, alpha=2.5, beta=-0.5, Cinit=C0, lda=40, ldb=96, ldc=97)
    assert relaxed_equal(got, 2.5 * (A @ B) - 0.5 * C0, 1e-12)


def test_k_zero_resets():
    # DMatDMatMultExpr.h:527-530: with K == 0 the target is reset.
    A = np.zeros((5, 0)); B = np.zeros((0, 4))
    got = gpu_gemm(A, B, Cinit=np.full((5, 4), 3.0))
    assert np.array_equal(got, np.zeros((5, 4)))


def test_beta_zero_never_reads_nan_target():
    rng = np.random.default_rng(11)
    A = rng.uniform(-1, 1, (300, 200)); B = rng.uniform(-1, 1, (200, 260))
    got = gpu_gemm(A, B)  # target pre-filled with NaN
    assert np.isfinite(got).all()


def test_invalid_arguments_rejected():
    import blaze_amd as bz
    with pytest.raises(bz.BlazeGpuError):
        ctx().call("bzg_dgemm", 4, 4, 4, 1.0, C.c_void_p(0), 4, C.c_void_p(0), 4, 0.0, C.c_void_p(0), 4)
    with pytest.raises(bz.BlazeGpuError):
        d = empty((16,), np.float64)
        ctx().call("bzg_dgemm", 4, 4, 4, 1.0, p(d), 3, p(d), 4, 0.0, p(d), 4)


def test_deterministic_repeat():
    rng = np.random.default_rng(12)
    A = rng.uniform(-1, 1, (700, 900)); B = rng.uniform(-1, 1, (900, 650))
    a = gpu_gemm(A, B); b = gpu_gemm(A, B)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_transposed_operands_bit_identical(dt):
    # the four storage-order forms (DMatTDMatMultExpr / TDMatDMatMultExpr / TDMatTDMatMultExpr):
    # bzg_*gemm_ex on transposed storage == bzg_*gemm on explicitly transposed copies, bit for
    # bit (same k order), and within Higham's bound of the oracle
    fn = "bzg_dgemm_ex" if dt == np.float64 else "bzg_sgemm_ex"
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    rng = np.random.default_rng(33)
    for (m, n, k, pad) in [(1, 1, 1, 0), (7, 5, 3, 1), (32, 32, 16, 0), (65, 130, 47, 2), (300, 260, 129, 1),
                           (1024, 1024, 256, 0)]:
        A = rng.uniform(-1, 1, (m, k)).astype(dt)
        B = rng.uniform(-1, 1, (k, n)).astype(dt)
        ref = gpu_gemm(A, B)
        want = oracle.gemm(A.astype(np.float64), B.astype(np.float64))
        bound = 2 * k * u * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))) + 1e-300
        assert np.all(np.abs(ref - want) <= bound)
        for ta in (0, 1):
            for tb in (0, 1):
                As = np.zeros((k, m + pad), dt) if ta else np.zeros((m, k + pad), dt)
                Bs = np.zeros((n, k + pad), dt) if tb else np.zeros((k, n + pad), dt)
                if ta: As[:, :m] = A.T
                else: As[:, :k] = A
                if tb: Bs[:, :k] = B.T
                else: Bs[:, :n] = B
                dA, dB = dev(As), dev(Bs)
                dC = full((m, n), dt, 3.0)
                ctx().call(fn, ta, tb, m, n, k, 1.0, p(dA), As.shape[1], p(dB), Bs.shape[1], 0.0, p(dC), n)
                got = host(dC)
                assert np.array_equal(got, ref), (m, n, k, ta, tb)


def test_transposed_alpha_beta_and_argument_checks():
    rng = np.random.default_rng(34)
    m, n, k = 50, 40, 30
    A = rng.uniform(-1, 1, (k, m)); B = rng.uniform(-1, 1, (n, k)); C0 = rng.uniform(-1, 1, (m, n))
    dA, dB, dC = dev(A), dev(B), dev(C0)
    ctx().call("bzg_dgemm_ex", 1, 1, m, n, k, -2.0, p(dA), m, p(dB), k, 0.5, p(dC), n)
    want = -2.0 * (A.T @ B.T) + 0.5 * C0
    bound = (2 * k + 4) * 2.0 ** -53 * (2 * np.abs(A.T) @ np.abs(B.T) + 0.5 * np.abs(C0))
    assert np.all(np.abs(host(dC) - want) <= bound)
    import blaze_amd as bz
    with pytest.raises(bz.BlazeGpuError):   # transposed A needs lda >= m
        ctx().call("bzg_dgemm_ex", 1, 0, m, n, k, 1.0, p(dA), m - 1, p(dB), n, 0.0, p(dC), n)


def test_int32_transposed_operands_exact():
    # bzg_igemm_ex: all four storage-order forms, full-range int32 (wrap-around products),
    # alpha / beta: bit-identical to the oracle's exact product of the logical operands
    rng = np.random.default_rng(34)
    for (m, n, k, pad) in [(1, 1, 1, 0), (7, 5, 3, 1), (67, 45, 33, 2), (130, 70, 200, 0)]:
        A = rng.integers(-2 ** 31, 2 ** 31, (m, k), dtype=np.int64).astype(np.int32)
        B = rng.integers(-2 ** 31, 2 ** 31, (k, n), dtype=np.int64).astype(np.int32)
        C0 = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
        AB = oracle.gemm(A, B)
        want = (3 * AB.astype(np.int64) - 2 * C0.astype(np.int64)).astype(np.uint64).astype(np.uint32).view(np.int32)
        for ta in (0, 1):
            for tb in (0, 1):
                As = np.zeros((k, m + pad), np.int32) if ta else np.zeros((m, k + pad), np.int32)
                Bs = np.zeros((n, k + pad), np.int32) if tb else np.zeros((k, n + pad), np.int32)
                if ta: As[:, :m] = A.T
                else: As[:, :k] = A
                if tb: Bs[:, :k] = B.T
                else: Bs[:, :n] = B
                dA, dB, dC = dev(As), dev(Bs), dev(C0)
                ctx().call("bzg_igemm_ex", ta, tb, m, n, k, 3, p(dA), As.shape[1], p(dB), Bs.shape[1], -2, p(dC), n)
                assert np.array_equal(host(dC), want), (m, n, k, ta, tb)


def test_int32_transposed_gemv_exact():
    # bzg_igemv_t: y = A^T x over a row-major A, wrap-around exact against the oracle's gemv
    # of the transposed copy
    rng = np.random.default_rng(35)
    for (m, n, pad) in [(1, 1, 0), (3, 7, 1), (500, 300, 2), (20000, 64, 0), (64, 20000, 0)]:
        A = np.zeros((m, n + pad), np.int32)
        A[:, :n] = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
        x = rng.integers(-2 ** 31, 2 ** 31, m, dtype=np.int64).astype(np.int32)
        want = oracle.gemv(np.ascontiguousarray(A[:, :n].T), x)
        dA, dx, dy = dev(A), dev(x), full((n,), np.int32, 5)  # held: the pointers must stay valid
        ctx().call("bzg_igemv_t", m, n, 1, p(dA), n + pad, p(dx), 0, p(dy))
        assert np.array_equal(host(dy), want), (m, n)


def test_fp32_large_tile_bit_identical_to_small_tiles():
    # a 4096 x 4096 FP32 product runs the 128 x 256 tile; rows 0..99 of it must equal, bit for
    # bit, a 100-row product of the same operands (small tiles), and all four storage-order
    # forms of the large product must be identical (every tile sums in the same k order)
    rng = np.random.default_rng(36)
    m, n, k = 4096, 4096, 200
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    big = gpu_gemm(A, B)
    small = gpu_gemm(A[:100].copy(), B)
    assert np.array_equal(big[:100], small)
    for ta in (0, 1):
        for tb in (0, 1):
            if not ta and not tb:
                continue
            dA = dev(np.ascontiguousarray(A.T) if ta else A)
            dB = dev(np.ascontiguousarray(B.T) if tb else B)
            dC = full((m, n), np.float32, 0.0)
            ctx().call("bzg_sgemm_ex", ta, tb, m, n, k, 1.0, p(dA), m if ta else k, p(dB), k if tb else n, 0.0,
                       p(dC), n)
            assert np.array_equal(host(dC), big), (ta, tb)
