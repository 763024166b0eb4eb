"""GPU parity: batched determinants (bzg_ddet_batched), the reference's det()
(blaze/math/expressions/DMatDetExpr.h:384-424; orders >= 7 via detNxN, :324-357).

Pins, none of them a restatement of the kernel's own product:
* exact determinants of integer matrices by fraction-free (Bareiss) elimination in Python
  integers: the FP64 LU result within 1e-12 relative;
* numpy.linalg.det (LAPACK dgetrf) on well-conditioned random matrices up to n = 512;
* exact special cases: permutation matrices give their sign, unit-triangular matrices give 1
  (the reference's test, operations/determinant/DenseTest.cpp:322-355), a zero column gives
  0 with info = its 1-based index, 1 x 1 gives the element, 0 x 0 gives 0 with info 0, and an
  order beyond 512 gives info -1 and NaN;
* the factors agree bit for bit with bzg_dgetrf_batched's.
"""
import numpy as np
import pytest

from tests.gpu_util import have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def _ctx():
    from tests.gpu_util import ctx
    return ctx()


def bareiss(M):
    """Exact determinant of an integer matrix (fraction-free elimination, Python ints)."""
    a = [[int(v) for v in row] for row in M]
    n = len(a)
    sign, prev = 1, 1
    for k in range(n - 1):
        if a[k][k] == 0:
            sw = next((i for i in range(k + 1, n) if a[i][k] != 0), None)
            if sw is None:
                return 0
            a[k], a[sw] = a[sw], a[k]
            sign = -sign
        for i in range(k + 1, n):
            for j in range(k + 1, n):
                a[i][j] = (a[i][j] * a[k][k] - a[i][k] * a[k][j]) // prev
        prev = a[k][k]
    return sign * a[n - 1][n - 1] if n else 0


def test_det_exact_integer_matrices():
    from blaze_amd.solve import det_batched
    rng = np.random.default_rng(81)
    As = [rng.integers(-5, 6, (n, n)).astype(np.float64) for n in list(range(1, 13)) * 3]
    dets, infos = det_batched(_ctx(), As)
    for A, d, info in zip(As, dets, infos):
        exact = bareiss(A.astype(np.int64))
        if exact == 0:
            assert abs(d) <= 1e-9 * np.prod(np.abs(A).sum(1) + 1), (A.shape, d)
        else:
            assert info == 0
            assert abs(d - exact) <= 1e-12 * abs(exact), (A.shape, d, exact)


def test_det_against_lapack_and_special_cases():
    from blaze_amd.solve import det_batched, getrf_batched
    rng = np.random.default_rng(82)
    sizes = [2, 3, 5, 7, 16, 64, 100, 255, 256, 257, 400, 512]
    As = [np.eye(n) + rng.uniform(-1, 1, (n, n)) / n for n in sizes]
    perms = [np.eye(n)[rng.permutation(n)] for n in (2, 3, 9, 40, 300)]
    lower = np.tril(rng.uniform(-1, 1, (9, 9)), -1) + np.eye(9)
    upper = np.triu(rng.uniform(-1, 1, (9, 9)), 1) + np.eye(9)
    zc = rng.uniform(-1, 1, (20, 20)); zc[:, 6] = 0.0
    one = np.array([[-2.75]])
    batch = As + perms + [lower, upper, zc, one, np.zeros((0, 0)), np.zeros((513, 513))]
    dets, infos = det_batched(_ctx(), batch)
    k = len(As)
    for A, d, info in zip(As, dets[:k], infos[:k]):
        ref = np.linalg.det(A)
        assert info == 0 and abs(d - ref) <= 1e-10 * abs(ref), (A.shape, d, ref)
    for P, d in zip(perms, dets[k:k + len(perms)]):
        assert d == round(np.linalg.det(P)) and abs(d) == 1.0
    k += len(perms)
    assert dets[k] == 1.0 and dets[k + 1] == 1.0          # unit lower / unit upper, exactly
    assert dets[k + 2] == 0.0 and infos[k + 2] == 7        # first zero pivot in column 7 (1-based)
    assert dets[k + 3] == -2.75 and infos[k + 3] == 0      # 1 x 1: the element itself
    assert dets[k + 4] == 0.0 and infos[k + 4] == 0        # 0 x 0
    assert np.isnan(dets[k + 5]) and infos[k + 5] == -1    # beyond 512
    # the factors are bzg_dgetrf_batched's: the product over its diagonal, in order, is det
    LUs, ipivs, _ = getrf_batched(_ctx(), As)
    for LU, piv, d in zip(LUs, ipivs, dets[:len(As)]):
        p = 1.0
        for i in range(LU.shape[0]):
            p *= LU[i, i] if piv[i] == i + 1 else -LU[i, i]
        assert p == d
