"""GPU parity of the complex element types through the C ABI: bzg_zgemm_ex / bzg_cgemm_ex,
bzg_zgemv_ex / bzg_cgemv_ex and bzg_zgesv (reference: DMatDMatMultExpr and DMatDVecMultExpr
with complex elements, cblas gemm.h:197-283 / gemv.h:186-245; solve() -> zgesv_,
clapack/gesv.h:317-333).

Checks: every product against the reference's own complex kernels (tests/golden/
complex_blazetest.npz) and the textbook product within the rounding bound
oracle.gemm_op_bound (summation order differs from the reference's, so entries agree to
rounding, not bit for bit); complex alpha and beta, beta = 0 over a NaN-filled C, k = 0,
alpha = 0, leading dimensions beyond the row length; solves against the reference's solve()
and by normwise backward error, and a singular system.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, dev, have_gpu, host, p

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from inputs import (COMPLEX_GEMM_SHAPES, COMPLEX_GEMV_SHAPES, COMPLEX_SOLVE_CASES, complex_gemm_case,  # noqa: E402
                    complex_gemm_ops, complex_gemv_case, complex_solve_case, load)

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

EPS = {np.dtype(np.complex128): 2.0 ** -53, np.dtype(np.complex64): 2.0 ** -24}
REAL = {np.dtype(np.complex128): np.float64, np.dtype(np.complex64): np.float32}


def _d(z):
    """Device copy of a complex array as its interleaved real view (at least one element)."""
    z = np.ascontiguousarray(z)
    r = z.view(REAL[z.dtype]).ravel()
    return dev(r if r.size else np.zeros(2, r.dtype))


def _scalar(v, dtype):
    a = np.array([v.real, v.imag], REAL[np.dtype(dtype)])
    return a, a.ctypes.data_as(C.c_void_p)


def _gemm(opa, opb, A, B, m, n, k, alpha=1.0, beta=0.0, Cin=None, ldc=None):
    dt = A.dtype
    ldc = ldc or max(n, 1)
    lda = A.shape[1] if A.ndim == 2 and A.shape[1] else max(k if opa == 0 else m, 1)
    ldb = B.shape[1] if B.ndim == 2 and B.shape[1] else max(n if opb == 0 else k, 1)
    Cst = np.zeros((m, ldc), dt) if Cin is None else np.array(Cin, dt)
    dA, dB, dC = _d(A), _d(B), _d(Cst)
    al, alp = _scalar(complex(alpha), dt)
    be, bep = _scalar(complex(beta), dt)
    fn = "bzg_zgemm_ex" if dt == np.complex128 else "bzg_cgemm_ex"
    ctx().call(fn, opa, opb, m, n, k, alp, p(dA), lda, p(dB), ldb, bep, p(dC), ldc)
    out = host(dC)[:2 * m * ldc].view(dt).reshape(m, ldc)
    return out


def _bound(opa, opb, A, B, dt, scale=1.0):
    return scale * oracle.gemm_op_bound(opa, opb, A, B, EPS[np.dtype(dt)])


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_gemm_vs_reference_fixture(dtype):
    fx = load("complex_blazetest.npz")
    seed = int(fx["seed"])
    key = "gemm_z" if dtype == np.complex128 else "gemm_c"
    ops_of = complex_gemm_ops if dtype == np.complex128 else (lambda idx: [(0, 0), (2, 1)])
    ref_all, off = fx[key], 0
    for idx, (m, n, k) in enumerate(COMPLEX_GEMM_SHAPES):
        for opa, opb in ops_of(idx):
            A, B = complex_gemm_case(seed, idx, opa, opb, dtype)
            ref = ref_all[off:off + m * n].reshape(m, n)
            off += m * n
            if m == 0 or n == 0:
                continue
            got = _gemm(opa, opb, A, B, m, n, k)
            assert np.all(np.abs(got - ref) <= 2 * _bound(opa, opb, A, B, dtype)), (idx, opa, opb)
    assert off == len(ref_all)


@pytest.mark.parametrize("opa,opb", [(0, 0), (1, 0), (0, 1), (2, 2), (2, 0), (0, 2), (1, 2)])
@pytest.mark.parametrize("shape", [(257, 190, 333), (1, 300, 7), (512, 512, 512)])
def test_complex_gemm_ops_alpha_beta(opa, opb, shape):
    m, n, k = shape
    rng = np.random.default_rng(m * 7 + n + k + 31 * opa + 5 * opb)
    z = lambda *s: (rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s))  # noqa: E731
    A = z(*((k, m) if opa else (m, k)))
    B = z(*((n, k) if opb else (k, n)))
    C0 = z(m, n + 3)  # ldc = n + 3: the 3 padding columns stay untouched
    alpha, beta = 0.75 - 0.5j, -0.25 + 1.5j
    got = _gemm(opa, opb, A, B, m, n, k, alpha, beta, C0, ldc=n + 3)
    want = alpha * oracle.gemm_op(opa, opb, A, B) + beta * C0[:, :n]
    bnd = 2 * abs(alpha) * _bound(opa, opb, A, B, np.complex128) + 4 * 2.0 ** -53 * abs(beta) * np.abs(C0[:, :n]) * 2
    assert np.all(np.abs(got[:, :n] - want) <= bnd)
    assert np.array_equal(got[:, n:], C0[:, n:])


def test_complex_gemm_beta_zero_ignores_nan_target_and_edge_sizes():
    rng = np.random.default_rng(5)
    m, n, k = 70, 45, 33
    A = rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))
    Cn = np.full((m, n), np.nan + 1j * np.nan)
    got = _gemm(0, 0, A, B, m, n, k, 1.0, 0.0, Cn)
    assert np.isfinite(got).all()
    assert np.all(np.abs(got - A @ B) <= 2 * _bound(0, 0, A, B, np.complex128))
    # k = 0 and alpha = 0: C := beta C (beta = 0 writes zeros over NaN)
    C0 = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    got = _gemm(0, 0, A[:, :0].copy(), B[:0].copy(), m, n, 0, 1.0, 2.0 - 1.0j, C0)
    assert np.allclose(got, (2.0 - 1.0j) * C0, rtol=4 * 2.0 ** -53, atol=0)
    got = _gemm(0, 0, A, B, m, n, k, 0.0, 0.0, Cn)
    assert np.array_equal(got, np.zeros((m, n), np.complex128))


def test_complex_gemm_rejects_bad_arguments():
    import blaze_amd as bz
    one = np.array([1.0, 0.0])
    op = one.ctypes.data_as(C.c_void_p)
    L = bz.lib()
    assert L.bzg_zgemm_ex(ctx().handle, 3, 0, 4, 4, 4, op, C.c_void_p(1), 4, C.c_void_p(1), 4, op,
                          C.c_void_p(1), 4) != 0
    assert L.bzg_zgemm_ex(ctx().handle, 0, 0, 4, 4, 4, op, C.c_void_p(1), 3, C.c_void_p(1), 4, op,
                          C.c_void_p(1), 4) != 0


def _gemv(op, A, x, alpha=1.0, beta=0.0, y0=None):
    dt = A.dtype
    m, n = A.shape
    L = n if op else m
    y = np.zeros(L, dt) if y0 is None else np.array(y0, dt)
    dA, dx, dy = _d(A), _d(x), _d(y)
    al, alp = _scalar(complex(alpha), dt)
    be, bep = _scalar(complex(beta), dt)
    fn = "bzg_zgemv_ex" if dt == np.complex128 else "bzg_cgemv_ex"
    ctx().call(fn, op, m, n, alp, p(dA), max(n, 1), p(dx), bep, p(dy))
    return host(dy)[:2 * L].view(dt)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_gemv_vs_reference_fixture(dtype):
    fx = load("complex_blazetest.npz")
    seed = int(fx["seed"])
    ref_all = fx["gemv_z" if dtype == np.complex128 else "gemv_c"]
    off = 0
    for idx, (m, n) in enumerate(COMPLEX_GEMV_SHAPES):
        for op in (0, 1, 2):
            A, x = complex_gemv_case(seed, idx, op, dtype)
            L = n if op else m
            ref = ref_all[off:off + L]
            off += L
            if L == 0:
                continue
            got = _gemv(op, A, x)
            bnd = 2 * oracle.gemm_op_bound(op, 0, A, x[:, None], EPS[np.dtype(dtype)])[:, 0]
            assert np.all(np.abs(got - ref) <= bnd), (idx, op)
    assert off == len(ref_all)


@pytest.mark.parametrize("op", [0, 1, 2])
@pytest.mark.parametrize("shape", [(4000, 300), (300, 4000), (1, 5000), (5000, 1)])
def test_complex_gemv_alpha_beta(op, shape):
    m, n = shape
    rng = np.random.default_rng(m + 3 * n + op)
    A = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    L, X = (n, m) if op else (m, n)
    x = rng.uniform(-1, 1, X) + 1j * rng.uniform(-1, 1, X)
    y0 = rng.uniform(-1, 1, L) + 1j * rng.uniform(-1, 1, L)
    alpha, beta = -1.25 + 0.5j, 0.5 - 0.75j
    got = _gemv(op, A, x, alpha, beta, y0)
    want = alpha * (oracle.op(A, op) @ x) + beta * y0
    bnd = 2 * abs(alpha) * oracle.gemm_op_bound(op, 0, A, x[:, None], 2.0 ** -53)[:, 0] + 8 * 2.0 ** -53 * abs(beta) * np.abs(y0)
    assert np.all(np.abs(got - want) <= bnd)


def _zgesv(A, B):
    import torch
    n, nrhs = B.shape
    dA, dB = _d(A), _d(B)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx().call("bzg_zgesv", n, nrhs, p(dA), n, p(dB), nrhs, p(info))
    X = host(dB)[:2 * n * nrhs].view(np.complex128).reshape(n, nrhs)
    Aout = host(dA)[:2 * n * n].view(np.complex128).reshape(n, n)
    return X, int(info.cpu()[0]), Aout


def test_zgesv_vs_reference_fixture():
    fx = load("complex_blazetest.npz")
    seed = int(fx["seed"])
    xs, off = fx["solve_z"], 0
    for idx, (n, nrhs) in enumerate(COMPLEX_SOLVE_CASES):
        A, B = complex_solve_case(seed, idx)
        ref = xs[off:off + n * nrhs].reshape(n, nrhs)
        off += n * nrhs
        X, info, Aout = _zgesv(A, B)
        assert info == 0
        assert np.array_equal(Aout, A)  # A is not modified
        scale = np.abs(A).sum(1).max() * np.abs(ref).max()
        assert np.abs(X - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), idx
        assert np.abs(A @ X - B).max() / scale < 1e-14, idx


@pytest.mark.parametrize("n,nrhs", [(300, 5), (1000, 2), (2100, 1)])
def test_zgesv_backward_error(n, nrhs):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs))
    X, info, _ = _zgesv(A, B)
    assert info == 0
    berr = np.abs(A @ X - B).max() / (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
    assert berr < 1e-14, berr


def test_zgesv_singular_leaves_b_and_reports():
    fx = load("complex_blazetest.npz")
    A, B = complex_solve_case(int(fx["seed"]), 7)
    A[:, 5] = 0  # the reference's solve() throws on this system (fixture: singular_throws)
    assert int(fx["singular_throws"]) == 1
    X, info, _ = _zgesv(A, B)
    assert info > 0
    assert np.array_equal(X, B)
    # order 0 and no right-hand sides
    X, info, _ = _zgesv(np.zeros((0, 0), np.complex128), np.zeros((0, 1), np.complex128))
    assert info == 0


@pytest.mark.parametrize("m,n,ldb_pad", [(37, 53, 0), (300, 7, 5), (1, 1, 0), (128, 256, 3)])
def test_ztranspose(m, n, ldb_pad):
    rng = np.random.default_rng(m * n + ldb_pad)
    A = rng.uniform(-1, 1, (m, n + 2)) + 1j * rng.uniform(-1, 1, (m, n + 2))  # lda = n + 2
    ldb = m + ldb_pad
    Bst = np.full((n, ldb), 9.5 - 2.0j)
    dA, dB = _d(A), _d(Bst)
    ctx().call("bzg_ztranspose", m, n, p(dA), n + 2, p(dB), ldb)
    got = host(dB)[:2 * n * ldb].view(np.complex128).reshape(n, ldb)
    assert np.array_equal(got[:, :m], A[:, :n].T)
    assert np.array_equal(got[:, m:], Bst[:, m:])
