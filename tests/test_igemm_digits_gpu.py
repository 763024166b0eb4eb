"""GPU parity: INT32 GEMM on the INT8 matrix cores (bzg_igemm_digits, and bzg_igemm's large
products routed to it) against the oracle's wrap-around product.

Blaze's integer products wrap modulo 2^32 (DMatDMatMultExpr.h with int elements; the known
answer of matrices/dynamicmatrix/ClassTest2.cpp:1244-1273 is checked in test_gemm_gpu.py).
Every result here is bit-exact against oracle.gemm (a plain C triple loop in unsigned 32-bit
arithmetic), including the edge of the exact range: k = 32767 with every digit at -128, where
the largest class sum is 4 * 32767 * 2^14 = 2^31 - 65536.
"""
import numpy as np
import pytest

import blaze_amd as bz
import oracle
from tests.gpu_util import ctx, dev, full, have_gpu, host, p

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def _wrap(v):
    return np.asarray(v, np.int64).astype(np.uint64).astype(np.uint32).view(np.int32)


def _stored(A, B, ta, tb, pad):
    m, k = A.shape
    n = B.shape[1]
    As = np.zeros((k, m + pad), np.int32) if ta else np.zeros((m, k + pad), np.int32)
    Bs = np.zeros((n, k + pad), np.int32) if tb else np.zeros((k, n + pad), np.int32)
    if ta: As[:, :m] = A.T
    else: As[:, :k] = A
    if tb: Bs[:, :k] = B.T
    else: Bs[:, :n] = B
    return As, Bs


def test_digits_all_forms_alpha_beta_exact():
    rng = np.random.default_rng(71)
    for (m, n, k, pad) in [(1, 1, 1, 0), (7, 5, 3, 1), (67, 45, 33, 2), (300, 260, 700, 3), (257, 513, 129, 0)]:
        A = rng.integers(-2 ** 31, 2 ** 31, (m, k), dtype=np.int64).astype(np.int32)
        B = rng.integers(-2 ** 31, 2 ** 31, (k, n), dtype=np.int64).astype(np.int32)
        A.flat[0] = -2 ** 31
        B.flat[-1] = 2 ** 31 - 1
        C0 = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
        AB = oracle.gemm(A, B)
        want = _wrap(3 * AB.astype(np.int64) - 2 * C0.astype(np.int64))
        for ta in (0, 1):
            for tb in (0, 1):
                As, Bs = _stored(A, B, ta, tb, (ta + tb + pad) % 3)
                dA, dB, dC = dev(As), dev(Bs), dev(C0)
                ctx().call("bzg_igemm_digits", ta, tb, m, n, k, 3, p(dA), As.shape[1], p(dB), Bs.shape[1], -2,
                           p(dC), n)
                assert np.array_equal(host(dC), want), (m, n, k, ta, tb)
        # alpha = 1, beta = 0: C is written without being read
        dA, dB = dev(A), dev(B)
        dC = full((m, n), np.int32, -7)
        ctx().call("bzg_igemm_digits", 0, 0, m, n, k, 1, p(dA), k, p(dB), n, 0, p(dC), n)
        assert np.array_equal(host(dC), AB), (m, n, k)


def test_digits_exact_at_largest_k():
    # 0x7F7F7F80 has the base-256 digits (-128, -128, -128, -128): every class sum is at its
    # largest, (c + 1) * k * 2^14, and still below 2^31 for k = 32767
    x = np.int32(0x7F7F7F80)
    m, n, k = 64, 96, 32767
    A = np.full((m, k), x, np.int32)
    B = np.full((k, n), x, np.int32)
    A[1, ::3] = -2 ** 31
    B[::5, 2] = 2 ** 31 - 1
    want = oracle.gemm(A, B)
    dA, dB, dC = dev(A), dev(B), full((m, n), np.int32, 0)
    ctx().call("bzg_igemm_digits", 0, 0, m, n, k, 1, p(dA), k, p(dB), n, 0, p(dC), n)
    assert np.array_equal(host(dC), want)
    # the transposed forms read the same values through the other digit kernels
    dAt, dBt = dev(np.ascontiguousarray(A.T)), dev(np.ascontiguousarray(B.T))
    dC2 = full((m, n), np.int32, 0)
    ctx().call("bzg_igemm_digits", 1, 1, m, n, k, 1, p(dAt), m, p(dBt), k, 0, p(dC2), n)
    assert np.array_equal(host(dC2), want)


def test_digits_long_k_in_chunks_and_k_zero_scales():
    # k beyond one pass (32767) runs in K chunks accumulated with beta = 1: 70001 = 3 chunks,
    # full-range values, alpha / beta, both storage orders of each operand
    rng = np.random.default_rng(75)
    m, n, k = 40, 72, 70001
    A = rng.integers(-2 ** 31, 2 ** 31, (m, k), dtype=np.int64).astype(np.int32)
    B = rng.integers(-2 ** 31, 2 ** 31, (k, n), dtype=np.int64).astype(np.int32)
    C0 = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
    want = _wrap((-3 * oracle.gemm(A, B).astype(np.int64)) % 2 ** 32 + (5 * C0.astype(np.int64)) % 2 ** 32)
    for ta, tb in ((0, 0), (1, 1)):
        As, Bs = _stored(A, B, ta, tb, 1)
        dA, dB, dC = dev(As), dev(Bs), dev(C0)
        ctx().call("bzg_igemm_digits", ta, tb, m, n, k, -3, p(dA), As.shape[1], p(dB), Bs.shape[1], 5, p(dC), n)
        assert np.array_equal(host(dC), want), (ta, tb)
    # routed: 256 x 256 x 33000 (two chunks) through bzg_igemm
    m2, n2, k2 = 256, 256, 33000
    A2 = rng.integers(-2 ** 31, 2 ** 31, (m2, k2), dtype=np.int64).astype(np.int32)
    B2 = rng.integers(-2 ** 31, 2 ** 31, (k2, n2), dtype=np.int64).astype(np.int32)
    dA2, dB2, dC2 = dev(A2), dev(B2), full((m2, n2), np.int32, -7)
    ctx().call("bzg_igemm", m2, n2, k2, 1, p(dA2), k2, p(dB2), n2, 0, p(dC2), n2)
    assert np.array_equal(host(dC2), oracle.gemm(A2, B2))
    # k == 0: C = beta * C
    dC3 = full((m, n), np.int32, 5)
    ctx().call("bzg_igemm_digits", 0, 0, m, n, 0, 1, p(dA), 1, p(dB), n, 3, p(dC3), n)
    assert np.array_equal(host(dC3), np.full((m, n), 15, np.int32))


def test_routed_products_exact():
    # bzg_igemm_ex takes the digit path for k >= 512 and m * n >= 256^2: all four storage-order
    # forms at 300 x 260 x 700 and a 2048 x 2048 x 600 product, exact against the oracle (and the
    # explicit entry point gives the same bits)
    rng = np.random.default_rng(73)
    m, n, k = 300, 260, 700
    A = rng.integers(-2 ** 31, 2 ** 31, (m, k), dtype=np.int64).astype(np.int32)
    B = rng.integers(-2 ** 31, 2 ** 31, (k, n), dtype=np.int64).astype(np.int32)
    C0 = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
    want = _wrap(-5 * oracle.gemm(A, B).astype(np.int64) + 7 * C0.astype(np.int64))
    for ta in (0, 1):
        for tb in (0, 1):
            As, Bs = _stored(A, B, ta, tb, 1)
            dA, dB, dC = dev(As), dev(Bs), dev(C0)
            ctx().call("bzg_igemm_ex", ta, tb, m, n, k, -5, p(dA), As.shape[1], p(dB), Bs.shape[1], 7, p(dC), n)
            assert np.array_equal(host(dC), want), (ta, tb)
    rng = np.random.default_rng(72)
    m, n, k = 2048, 2048, 600
    A = rng.integers(-2 ** 31, 2 ** 31, (m, k), dtype=np.int64).astype(np.int32)
    B = rng.integers(-2 ** 31, 2 ** 31, (k, n), dtype=np.int64).astype(np.int32)
    want = oracle.gemm(A, B)
    dA, dB = dev(A), dev(B)
    dC = full((m, n), np.int32, -7)
    ctx().call("bzg_igemm", m, n, k, 1, p(dA), k, p(dB), n, 0, p(dC), n)
    assert np.array_equal(host(dC), want)
    dC2 = full((m, n), np.int32, -7)
    ctx().call("bzg_igemm_digits", 0, 0, m, n, k, 1, p(dA), k, p(dB), n, 0, p(dC2), n)
    assert np.array_equal(host(dC2), want)


def test_digits_random_shapes_exact():
    # seeded sweep over ragged shapes (k crossing the 128-byte stage and 256-row tile edges),
    # storage orders, leading-dimension padding and alpha / beta, bit-exact against the oracle
    rng = np.random.default_rng(74)
    for case in range(24):
        m, n = (int(v) for v in rng.integers(1, 700, 2))
        k = int(rng.integers(1, 3000))
        ta, tb, pad = int(rng.integers(0, 2)), int(rng.integers(0, 2)), int(rng.integers(0, 5))
        alpha, beta = (int(v) for v in rng.integers(-2 ** 31, 2 ** 31, 2, dtype=np.int64))
        if case % 4 == 0:
            beta = 0
        A = rng.integers(-2 ** 31, 2 ** 31, (m, k), dtype=np.int64).astype(np.int32)
        B = rng.integers(-2 ** 31, 2 ** 31, (k, n), dtype=np.int64).astype(np.int32)
        C0 = rng.integers(-2 ** 31, 2 ** 31, (m, n), dtype=np.int64).astype(np.int32)
        AB = oracle.gemm(A, B).astype(np.int64)
        want = _wrap((alpha * AB) % 2 ** 32 + (beta * C0.astype(np.int64)) % 2 ** 32)
        As, Bs = _stored(A, B, ta, tb, pad)
        dA, dB, dC = dev(As), dev(Bs), dev(C0)
        ctx().call("bzg_igemm_digits", ta, tb, m, n, k, alpha, p(dA), As.shape[1], p(dB), Bs.shape[1], beta,
                   p(dC), n)
        assert np.array_equal(host(dC), want), (case, m, n, k, ta, tb, pad, alpha, beta)
