"""Seeded inputs shared by the golden-fixture generator and the tests that read the fixtures.

Every input is regenerated from a stored seed by the package's generator library
(blaze_amd/csrc/gen.c: MT19937-64 and SplitMix64 counter streams), so the fixtures hold
only seeds and expected outputs.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import blaze_amd as bz

HERE = Path(__file__).resolve().parent
vp, u64, i64, i32 = C.c_void_p, C.c_uint64, C.c_int64, C.c_int32


def _p(a):
    return a.ctypes.data_as(vp)


def ints(seed, stream, n, lo=-10, hi=10):
    out = np.empty(n, np.int32)
    bz.gen().bzg_gen_int_ctr(u64(seed), u64(stream), i64(n), i32(lo), i32(hi), _p(out))
    return out


def uniform(seed, stream, n):
    out = np.empty(n, np.float64)
    bz.gen().bzg_gen_uniform_ctr(u64(seed), u64(stream), i64(n), _p(out))
    return out


def uniform_mt(seed, *shapes):
    """Matrices filled one after the other from a single MT19937-64 stream (A first, then B)."""
    outs = [np.empty(s) for s in shapes]
    if len(outs) == 1:
        bz.gen().bzg_gen_uniform_mt(u64(seed), i64(outs[0].size), _p(outs[0]), i64(0), vp(0))
    else:
        bz.gen().bzg_gen_uniform_mt(u64(seed), i64(outs[0].size), _p(outs[0]), i64(outs[1].size), _p(outs[1]))
    return outs


def laplace3d(nx):
    nnz = bz.gen().bzg_gen_laplace3d_nnz(i64(nx))
    n = nx ** 3
    rp = np.empty(n + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
    bz.gen().bzg_gen_laplace3d(i64(nx), _p(rp), _p(ci), _p(v))
    return rp, ci, v


def zipf(seed, n, s, kmax):
    rp = np.empty(n + 1, np.int32)
    bz.gen().bzg_gen_zipf_rowptr(u64(seed), i64(n), C.c_double(s), i32(kmax), _p(rp))
    nnz = int(rp[-1])
    ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
    bz.gen().bzg_gen_zipf_fill(u64(seed), i64(n), i64(n), _p(rp), _p(ci), _p(v))
    return rp, ci, v


GEMM_SHAPES = [(m, n, k) for m in range(7) for n in range(7) for k in range(7)] + [
    (15, 37, 63), (37, 15, 63), (16, 32, 64), (64, 64, 64), (63, 65, 67), (128, 128, 128)]


def gemm_case(seed, idx, dtype):
    """Operands of case idx of the size grid: int32 in [-10, 10] or double uniform[-1, 1)."""
    m, n, k = GEMM_SHAPES[idx]
    if dtype == np.int32:
        return ints(seed, 2 * idx, m * k).reshape(m, k), ints(seed, 2 * idx + 1, k * n).reshape(k, n)
    return (uniform(seed, 2 * idx, m * k).reshape(m, k), uniform(seed, 2 * idx + 1, k * n).reshape(k, n))


def gemv_case(seed, idx, m, n, dtype):
    """Operands of GEMV case idx: A (m x n), x (n) for A * x and xt (m) for trans(A) * x;
    int32 in [-10, 10] or uniform[-1, 1) (float: the double values rounded)."""
    if dtype == np.int32:
        return (ints(seed, 3 * idx, m * n).reshape(m, n), ints(seed, 3 * idx + 1, n), ints(seed, 3 * idx + 2, m))
    A = uniform(seed, 3 * idx, m * n).reshape(m, n).astype(dtype)
    return A, uniform(seed, 3 * idx + 1, n).astype(dtype), uniform(seed, 3 * idx + 2, m).astype(dtype)


def solve_small_case(n, t):
    """System t (0..3) of size n of the closed-form fixture; t == 3 is exactly singular
    (a zero last row, so every term of a determinant expansion is an exact zero)."""
    A = uniform(0x501E, 100 * n + 2 * t, n * n).reshape(n, n) + n * np.eye(n)
    if t == 3:
        A[-1] = 0.0
    b = uniform(0x501E, 100 * n + 2 * t + 1, n)
    return A, b


def load(name):
    with np.load(HERE / name, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def complex_uniform(seed, stream, shape, dtype=np.complex128):
    """Complex entries with real and imaginary parts uniform[-1, 1) from counter streams 2s, 2s + 1
    (complex64: the double values rounded)."""
    size = int(np.prod(shape))
    z = uniform(seed, 2 * stream, size) + 1j * uniform(seed, 2 * stream + 1, size)
    return z.reshape(shape).astype(dtype)


COMPLEX_GEMM_SHAPES = [(m, n, k) for m in (0, 1, 2, 3, 5) for n in (0, 1, 2, 3, 5) for k in (0, 1, 2, 3, 5)] + [
    (15, 37, 63), (37, 15, 63), (64, 64, 64), (33, 35, 67)]
COMPLEX_GEMM_OPS = [(0, 0), (1, 0), (0, 1), (0, 2), (2, 0), (2, 1), (1, 2), (2, 2)]
COMPLEX_GEMM_OPS_LARGE = [(0, 0), (2, 1), (1, 2)]  # shapes beyond 5 x 5 x 5


def complex_gemm_ops(idx):
    """The op pairs of case idx: all eight on the small grid, three on the larger shapes."""
    return COMPLEX_GEMM_OPS if max(COMPLEX_GEMM_SHAPES[idx]) <= 5 else COMPLEX_GEMM_OPS_LARGE


def complex_gemm_case(seed, idx, opa, opb, dtype=np.complex128):
    """Stored operands of complex GEMM case idx (shape from COMPLEX_GEMM_SHAPES, op pair):
    A is m x k (op 0) or k x m, B is k x n (op 0) or n x k."""
    m, n, k = COMPLEX_GEMM_SHAPES[idx]
    sa = (k, m) if opa else (m, k)
    sb = (n, k) if opb else (k, n)
    return (complex_uniform(seed, 2 * idx, sa, dtype), complex_uniform(seed, 2 * idx + 1, sb, dtype))


COMPLEX_GEMV_SHAPES = [(m, n) for m in (0, 1, 2, 4) for n in (0, 1, 3, 4)] + [(67, 127), (127, 67), (300, 1000)]


def complex_gemv_case(seed, idx, op, dtype=np.complex128):
    """A (m x n) and x (n entries for op 0, m for op 1 / 2) of complex GEMV case idx."""
    m, n = COMPLEX_GEMV_SHAPES[idx]
    return (complex_uniform(seed, 2 * idx, (m, n), dtype), complex_uniform(seed, 2 * idx + 1, (m if op else n,), dtype))


COMPLEX_SOLVE_CASES = [(1, 1), (2, 1), (3, 2), (5, 1), (6, 3), (7, 1), (16, 2), (50, 4), (129, 3)]


def complex_solve_case(seed, idx):
    """A (n x n, complex uniform plus n on the diagonal) and B (n x nrhs) of solve case idx."""
    n, nrhs = COMPLEX_SOLVE_CASES[idx]
    A = complex_uniform(seed, 2 * idx, (n, n)) + n * np.eye(n)
    return A, complex_uniform(seed, 2 * idx + 1, (n, nrhs))
