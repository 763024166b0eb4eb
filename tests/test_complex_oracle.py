"""CPU pins of the complex-element checker (oracle.gemm_op, oracle.op, oracle.gemm_op_bound).

The fixtures in tests/golden/complex_blazetest.npz are the reference's own complex kernels
(DMatDMatMultExpr / DMatDVecMultExpr with complex<double> and complex<float> elements, and
solve() -> zgesv_), compiled from its headers into oracle/_ref (make_golden.py).  Here they are
held against the textbook complex product in numpy under a rounding bound that a wrong
transpose code, a dropped conjugation or a swapped operand exceeds, and the solve fixture
against the residual of the system and LAPACK's zgesv (numpy.linalg.solve).  No GPU is used.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import oracle

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from inputs import (COMPLEX_GEMM_SHAPES, COMPLEX_GEMV_SHAPES, COMPLEX_SOLVE_CASES, complex_gemm_case,  # noqa: E402
                    complex_gemm_ops, complex_gemv_case, complex_solve_case, load)

FIX = "complex_blazetest.npz"


@pytest.fixture(scope="module")
def fx():
    return load(FIX)


def _gemm_cases(seed, dtype, ops_of):
    off = 0
    for idx, (m, n, k) in enumerate(COMPLEX_GEMM_SHAPES):
        for opa, opb in ops_of(idx):
            A, B = complex_gemm_case(seed, idx, opa, opb, dtype)
            yield idx, opa, opb, A, B, (m, n), off
            off += m * n


def test_complex_double_products_match_textbook(fx):
    z, seed, last = fx["gemm_z"], int(fx["seed"]), 0
    for idx, opa, opb, A, B, (m, n), off in _gemm_cases(seed, np.complex128, complex_gemm_ops):
        ref = z[off:off + m * n].reshape(m, n)
        got = oracle.gemm_op(opa, opb, A, B)
        bound = oracle.gemm_op_bound(opa, opb, A, B, 2.0 ** -53)
        assert np.all(np.abs(got - ref) <= bound), (idx, opa, opb)
        last = off + m * n
    assert last == len(z)


def test_complex_float_products_match_textbook(fx):
    c, seed, last = fx["gemm_c"], int(fx["seed"]), 0
    ops = lambda idx: [(0, 0), (2, 1)]  # noqa: E731
    for idx, opa, opb, A, B, (m, n), off in _gemm_cases(seed, np.complex64, ops):
        ref = c[off:off + m * n].reshape(m, n)
        got = oracle.gemm_op(opa, opb, A, B)
        bound = oracle.gemm_op_bound(opa, opb, A, B, 2.0 ** -24)
        assert np.all(np.abs(got - ref) <= bound), (idx, opa, opb)
        last = off + m * n
    assert last == len(c)


def test_complex_bound_rejects_a_wrong_op(fx):
    # the checker is sharp: the product with op(B) = B^T where the reference computed B^H, or
    # with A's conjugation dropped, lies far outside the bound
    z, seed = fx["gemm_z"], int(fx["seed"])
    hits = 0
    for idx, opa, opb, A, B, (m, n), off in _gemm_cases(seed, np.complex128, complex_gemm_ops):
        if m * n == 0 or COMPLEX_GEMM_SHAPES[idx][2] < 3 or (opa, opb) not in ((0, 2), (2, 1)):
            continue
        ref = z[off:off + m * n].reshape(m, n)
        wrong = oracle.gemm_op(opa if opb == 2 else 1, 1 if opb == 2 else opb, A, B)
        bound = oracle.gemm_op_bound(opa, opb, A, B, 2.0 ** -53)
        assert np.any(np.abs(wrong - ref) > bound), (idx, opa, opb)
        hits += 1
    assert hits > 10


def test_complex_gemv_fixture(fx):
    gz, gc, seed = fx["gemv_z"], fx["gemv_c"], int(fx["seed"])
    off = 0
    for idx, (m, n) in enumerate(COMPLEX_GEMV_SHAPES):
        for op in (0, 1, 2):
            A, x = complex_gemv_case(seed, idx, op)
            A32, x32 = complex_gemv_case(seed, idx, op, np.complex64)
            L = n if op else m
            want = oracle.op(A, op) @ x
            bnd = oracle.gemm_op_bound(op, 0, A, x[:, None], 2.0 ** -53)[:, 0]
            assert np.all(np.abs(gz[off:off + L] - want) <= bnd), (idx, op)
            want32 = oracle.op(A32.astype(np.complex128), op) @ x32.astype(np.complex128)
            bnd32 = oracle.gemm_op_bound(op, 0, A32, x32[:, None], 2.0 ** -24)[:, 0]
            assert np.all(np.abs(gc[off:off + L] - want32) <= bnd32), (idx, op)
            off += L
    assert off == len(gz) == len(gc)


def test_complex_solve_fixture(fx):
    xs, seed = fx["solve_z"], int(fx["seed"])
    off = 0
    for idx, (n, nrhs) in enumerate(COMPLEX_SOLVE_CASES):
        A, B = complex_solve_case(seed, idx)
        X = xs[off:off + n * nrhs].reshape(n, nrhs)
        off += n * nrhs
        res = np.abs(A @ X - B).max() / (np.abs(A).sum(1).max() * np.abs(X).max())
        assert res < 64 * 2.0 ** -53, (idx, res)
        assert np.allclose(X, np.linalg.solve(A, B), rtol=0, atol=1e-13 * max(1.0, np.abs(X).max()))
    assert off == len(xs)
    assert int(fx["singular_throws"]) == 1
