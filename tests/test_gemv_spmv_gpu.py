"""GPU parity: dense GEMV (y = A * x) and CSR SpMV (y = S * x).

GEMV mirrors blazetest dmatdvecmult (MDaVDa.cpp sizes {0..6}^2, 67x127, 127x67, 64x128,
128x64); SpMV mirrors smatdvecmult (MCaVDa.cpp:79,86: {0..6}^2 with every nnz count) and the
reference's row product order (SVecDVecInnerExpr.h:127-135).  The vector-CSR path adds every
row's products in storage order, so it is compared bit-for-bit with the oracle (and with the
compiled reference where present); wavefront-per-row rows are compared within
gamma_n * sum|v||x| of the exact sum.
"""
import ctypes as C

import numpy as np
import pytest

import blaze_amd as bz
import oracle
from tests.gpu_util import ctx, dev, empty, full, have_gpu, host, p

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

GEMV = {np.float64: "bzg_dgemv", np.float32: "bzg_sgemv", np.int32: "bzg_igemv"}
CSRMV = {np.float64: "bzg_dcsrmv", np.float32: "bzg_scsrmv", np.int32: "bzg_icsrmv"}


def gpu_gemv(A, x, alpha=1, beta=0, y0=None, lda=None):
    m, n = A.shape
    lda = lda or max(n, 1)
    Ap = np.zeros((m, lda), A.dtype); Ap[:, :n] = A
    dA = dev(Ap) if Ap.size else empty((1,), A.dtype)
    dx = dev(x) if n else empty((1,), A.dtype)
    dy = dev(y0) if y0 is not None else full((max(m, 1),), A.dtype, -3)
    ctx().call(GEMV[A.dtype.type], m, n, alpha, p(dA), lda, p(dx), beta, p(dy))
    return host(dy)[:m]


def test_gemv_blazetest_grid():
    rng = np.random.default_rng(20)
    shapes = [(m, n) for m in range(7) for n in range(7)] + [(67, 127), (127, 67), (64, 128), (128, 64)]
    for (m, n) in shapes:
        Ai = rng.integers(-10, 11, (m, n)).astype(np.int32); xi = rng.integers(-10, 11, n).astype(np.int32)
        assert np.array_equal(gpu_gemv(Ai, xi), oracle.gemv(Ai, xi)), (m, n)
        Ad = rng.uniform(-1, 1, (m, n)); xd = rng.uniform(-1, 1, n)
        got = gpu_gemv(Ad, xd); ref = oracle.gemv(Ad, xd)
        assert np.all(np.abs(got - ref) <= 1e-8 * np.maximum(1, np.abs(ref))), (m, n)


def test_gemv_large_and_unaligned():
    rng = np.random.default_rng(21)
    for (m, n, lda) in [(4096, 4096, 4096), (1000, 3001, 3001), (513, 777, 780), (3, 20000, 20000)]:
        A = rng.uniform(-1, 1, (m, n)); x = rng.uniform(-1, 1, n)
        got = gpu_gemv(A, x, lda=lda)
        ref = A @ x
        bound = n * 2.0 ** -53 * 1.01 * (np.abs(A) @ np.abs(x))
        assert np.all(np.abs(got - ref) <= bound + 1e-300), (m, n)
        Af = A.astype(np.float32); xf = x.astype(np.float32)
        gf = gpu_gemv(Af, xf, lda=lda)
        bf = n * 2.0 ** -24 * 1.01 * (np.abs(Af).astype(np.float64) @ np.abs(xf))
        assert np.all(np.abs(gf - Af.astype(np.float64) @ xf) <= bf), (m, n)


def test_gemv_alpha_beta():
    rng = np.random.default_rng(22)
    A = rng.uniform(-1, 1, (300, 500)); x = rng.uniform(-1, 1, 500); y0 = rng.uniform(-1, 1, 300)
    got = gpu_gemv(A, x, alpha=-1.5, beta=2.0, y0=y0)
    assert np.allclose(got, -1.5 * (A @ x) + 2.0 * y0, rtol=1e-13, atol=1e-13)


def test_gemv_wide_column_split():
    # few rows, long rows: the columns run as chunks whose sums are added in chunk order
    # (aligned and unaligned leading dimensions, tails shorter than a vector, alpha/beta,
    # int32 wrap-around exact, identical results run to run)
    rng = np.random.default_rng(23)
    for (m, n, lda) in [(1, 4096, 4096), (64, 100000, 100000), (64, 100000, 100003), (700, 9000, 9001),
                        (17, 65537, 65540)]:
        A = rng.uniform(-1, 1, (m, n)); x = rng.uniform(-1, 1, n); y0 = rng.uniform(-1, 1, m)
        absv = np.abs(A) @ np.abs(x)
        got = gpu_gemv(A, x, lda=lda)
        assert np.all(np.abs(got - A @ x) <= (n + 2) * 2.0 ** -53 * absv + 1e-300), (m, n, lda)
        assert np.array_equal(gpu_gemv(A, x, lda=lda), got)
        g2 = gpu_gemv(A, x, alpha=0.5, beta=-2.0, y0=y0, lda=lda)
        assert np.all(np.abs(g2 - (0.5 * (A @ x) - 2.0 * y0)) <= (n + 4) * 2.0 ** -53 * (absv + 2 * np.abs(y0)))
        Af = A.astype(np.float32); xf = x.astype(np.float32)
        gf = gpu_gemv(Af, xf, lda=lda)
        bf = (n + 2) * 2.0 ** -24 * (np.abs(Af).astype(np.float64) @ np.abs(xf))
        assert np.all(np.abs(gf - Af.astype(np.float64) @ xf) <= bf), (m, n, lda)
        Ai = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
        xi = rng.integers(-2 ** 31, 2 ** 31, n, dtype=np.int64).astype(np.int32)
        assert np.array_equal(gpu_gemv(Ai, xi, lda=lda), oracle.gemv(Ai, xi)), (m, n, lda)


# ---- SpMV ---------------------------------------------------------------------------------
class Plan:
    def __init__(self, rowptr, n_cols):
        self.h = C.c_void_p()
        bz.check("bzg_csr_plan_create", bz.lib().bzg_csr_plan_create(
            ctx().handle, len(rowptr) - 1, n_cols, rowptr.ctypes.data_as(C.c_void_p), C.byref(self.h)))

    def describe(self):
        buf = C.create_string_buffer(256)
        bz.lib().bzg_csr_plan_describe(self.h, buf, 256)
        return buf.value.decode()

    def __del__(self):
        bz.lib().bzg_csr_plan_destroy(self.h)


def gpu_csrmv(rowptr, col, val, x, n_cols):
    plan = Plan(rowptr, n_cols)
    n = len(rowptr) - 1
    d = [dev(a) if a.size else empty((1,), a.dtype) for a in (rowptr, col, val, x)]
    dy = full((max(n, 1),), val.dtype, -5)
    ctx().call(CSRMV[val.dtype.type], plan.h, *[p(t) for t in d], p(dy))
    return host(dy)[:n], plan.describe()


LONG_ROW = 32  # rows longer than this use the tree-sum (wavefront / workgroup per row) kernels


def check_spmv(got, rp, ci, va, x):
    """Rows <= LONG_ROW: bit-identical to the sequential storage-order sum; longer rows: within
    2 * gamma_len * sum|v||x| of it (fixed-shape tree order)."""
    ref = oracle.csrmv(rp, ci, va, x)
    lens = np.diff(rp)
    short = lens <= LONG_ROW
    assert np.array_equal(got[short], ref[short])
    if (~short).any():
        absx = np.abs(va.astype(np.float64)) * np.abs(x[ci].astype(np.float64))
        csum = np.concatenate([[0.0], np.cumsum(absx)])
        bound = csum[rp[1:]] - csum[rp[:-1]]
        u = 2.0 ** -53 if va.dtype == np.float64 else 2.0 ** -24
        tol = 2 * lens * u * 1.01 * bound
        assert np.all(np.abs(got[~short] - ref[~short]) <= tol[~short])
    return ref


def random_csr(rng, m, n, nnz, dtype):
    # blazetest creator: random (row, col) positions until nnz distinct entries exist
    pos = set()
    while len(pos) < nnz:
        pos.add((int(rng.integers(0, m)), int(rng.integers(0, n))))
    pos = sorted(pos)
    rowptr = np.zeros(m + 1, np.int32)
    for (r, _) in pos:
        rowptr[r + 1] += 1
    rowptr = np.cumsum(rowptr).astype(np.int32)
    col = np.array([c for (_, c) in pos], np.int32)
    if dtype == np.int32:
        val = rng.integers(-10, 11, len(pos)).astype(np.int32)
    else:
        val = rng.uniform(-1, 1, len(pos)).astype(dtype)
    return rowptr, col, val


def test_spmv_blazetest_grid():
    rng = np.random.default_rng(30)
    for m in range(7):
        for n in range(7):
            for nnz in range(0, m * n + 1):
                for dt in (np.int32, np.float64):
                    rp, ci, va = random_csr(rng, m, n, nnz, dt)
                    x = (rng.integers(-10, 11, n).astype(np.int32) if dt == np.int32 else rng.uniform(-1, 1, n))
                    got, _ = gpu_csrmv(rp, ci, va, x, n)
                    assert np.array_equal(got, oracle.csrmv(rp, ci, va, x)), (m, n, nnz, dt)


def test_spmv_67x127():
    rng = np.random.default_rng(31)
    for nnz in (0, 13, 67, 1000, 67 * 127):
        rp, ci, va = random_csr(rng, 67, 127, nnz, np.float64)
        x = rng.uniform(-1, 1, 127)
        got, _ = gpu_csrmv(rp, ci, va, x, 127)
        ref = check_spmv(got, rp, ci, va, x)
        if oracle.have_ref():  # the reference's own kernel gives the oracle's sums bit for bit
            assert np.array_equal(ref, oracle.ref_csrmv(rp, ci, va, x, 127))


def laplace(nx):
    g = bz.gen()
    n = nx ** 3
    nnz = g.bzg_gen_laplace3d_nnz(C.c_int64(nx))
    rp = np.zeros(n + 1, np.int32); ci = np.zeros(nnz, np.int32); va = np.zeros(nnz)
    g.bzg_gen_laplace3d(C.c_int64(nx), rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                        va.ctypes.data_as(C.c_void_p))
    return rp, ci, va


def test_spmv_laplacian_bit_exact():
    rng = np.random.default_rng(32)
    for nx in (16, 64):
        rp, ci, va = laplace(nx)
        x = rng.uniform(-1, 1, nx ** 3)
        got, desc = gpu_csrmv(rp, ci, va, x, nx ** 3)
        assert "wavefront-per-row rows=0" in desc
        assert np.array_equal(got, oracle.csrmv(rp, ci, va, x)), nx
        # fp32 variant
        got32, _ = gpu_csrmv(rp, ci, va.astype(np.float32), x.astype(np.float32), nx ** 3)
        assert np.array_equal(got32, oracle.csrmv(rp, ci, va.astype(np.float32), x.astype(np.float32)))


def test_spmv_laplacian_128_integer_exact():
    # Full 128^3 size: integer x -> every row sum exact; 6*x_i - sum of neighbours.
    rp, ci, va = laplace(128)
    rng = np.random.default_rng(33)
    x = rng.integers(-1000, 1000, 128 ** 3).astype(np.float64)
    got, _ = gpu_csrmv(rp, ci, va, x, 128 ** 3)
    ref = oracle.csrmv(rp, ci, va, x)
    assert np.array_equal(got, ref)
    ones, _ = gpu_csrmv(rp, ci, va, np.ones(128 ** 3), 128 ** 3)
    # row sums: 6 - (number of neighbours) = number of missing neighbours
    assert ones.min() == 0 and ones.max() == 3


def zipf(n, s=1.5, kmax=2000, seed=5):
    g = bz.gen()
    rp = np.zeros(n + 1, np.int32)
    g.bzg_gen_zipf_rowptr(C.c_uint64(seed), C.c_int64(n), C.c_double(s), C.c_int32(kmax),
                          rp.ctypes.data_as(C.c_void_p))
    nnz = int(rp[-1])
    ci = np.zeros(nnz, np.int32); va = np.zeros(nnz)
    g.bzg_gen_zipf_fill(C.c_uint64(seed), C.c_int64(n), C.c_int64(n), rp.ctypes.data_as(C.c_void_p),
                        ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p))
    return rp, ci, va


def test_spmv_powerlaw_rows():
    rng = np.random.default_rng(34)
    n = 20000
    rp, ci, va = zipf(n)
    x = rng.uniform(-1, 1, n)
    got, desc = gpu_csrmv(rp, ci, va, x, n)
    assert "wavefront-per-row" in desc
    check_spmv(got, rp, ci, va, x)
    # integer-valued values: exact in any order
    vi = np.round(va * 10); xi = np.round(x * 10)
    goti, _ = gpu_csrmv(rp, ci, vi, xi, n)
    assert np.array_equal(goti, oracle.csrmv(rp, ci, vi, xi))


def test_spmv_empty_rows_and_int32():
    rp = np.array([0, 0, 2, 2, 3], np.int32); ci = np.array([1, 3, 0], np.int32)
    va = np.array([2, -3, 5], np.int32); x = np.array([1, 2, 3, 4], np.int32)
    got, _ = gpu_csrmv(rp, ci, va, x, 4)
    assert got.tolist() == [0, 2 * 2 - 3 * 4, 0, 5]


def test_spmv_no_nonzeros_null_arrays():
    # rows without any nonzero: y = 0, and x / colidx / val may be null (nothing is gathered)
    for dt, fn in ((np.float64, "bzg_dcsrmv"), (np.float32, "bzg_scsrmv")):
        rp = np.zeros(1001, np.int32)
        plan = Plan(rp, 1000)
        drp = dev(rp)
        dy = full((1000,), dt, -5)
        ctx().call(fn, plan.h, p(drp), C.c_void_p(0), C.c_void_p(0), C.c_void_p(0), p(dy))
        assert not host(dy).any()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gemv_transposed(dt):
    # y = alpha A^T x + beta y (TDMatDVecMultExpr): within Higham's bound of the oracle's
    # gemv on the transposed copy, deterministic, beta = 0 never reads y
    fn = "bzg_dgemv_t" if dt == np.float64 else "bzg_sgemv_t"
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    rng = np.random.default_rng(35)
    for (m, n, pad) in [(1, 1, 0), (3, 7, 1), (100, 513, 2), (4096, 300, 0), (20000, 1030, 1), (64, 20000, 0),
                        (100000, 1, 0), (50000, 3, 1), (30000, 64, 0), (30000, 65, 0), (9000, 130, 3)]:
        A = np.zeros((m, n + pad), dt); A[:, :n] = rng.uniform(-1, 1, (m, n))
        x = rng.uniform(-1, 1, m).astype(dt)
        y0 = rng.uniform(-1, 1, n).astype(dt)
        dA, dx = dev(A), dev(x)
        want = oracle.gemv(np.ascontiguousarray(A[:, :n].T).astype(np.float64), x.astype(np.float64))
        absv = np.abs(A[:, :n].T).astype(np.float64) @ np.abs(x).astype(np.float64)
        dy = full((n,), dt, float("nan"))
        ctx().call(fn, m, n, 1.0, p(dA), n + pad, p(dx), 0.0, p(dy))
        got = host(dy)
        assert np.all(np.abs(got - want) <= (m + 2) * u * absv + 1e-300), (m, n)
        dy2 = full((n,), dt, float("nan"))
        ctx().call(fn, m, n, 1.0, p(dA), n + pad, p(dx), 0.0, p(dy2))
        assert np.array_equal(host(dy2), got)
        dy3 = dev(y0)
        ctx().call(fn, m, n, -0.5, p(dA), n + pad, p(dx), 2.0, p(dy3))
        w3 = -0.5 * want + 2.0 * y0.astype(np.float64)
        assert np.all(np.abs(host(dy3) - w3) <= (m + 4) * u * (0.5 * absv + 2 * np.abs(y0)) + 1e-300)
    # m = 0: y = beta y
    dy = dev(np.ones(5, dt))
    dA0, dx0 = dev(np.zeros(5, dt)), dev(np.zeros(1, dt))
    ctx().call(fn, 0, 5, 1.0, p(dA0), 5, p(dx0), 3.0, p(dy))
    assert np.array_equal(host(dy), np.full(5, 3.0, dt))
