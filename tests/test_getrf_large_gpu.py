"""GPU parity of the single-system LU for any order (bzg_dgetrf / bzg_dgesv / bzg_dgetri /
bzg_ddet_lu), the path of solve(), lu(), det() and inv() beyond the batched kernels' n <= 512
and 8 right-hand sides (reference: LSE.h:5229-5257 -> dgesv_, LU.h:98-140, Inversion.h:3337-3347,
DMatDetExpr.h:324-357).

Pins: the pivots equal the oracle's dgetf2 pivots (oracle/oracle.c orc_getrf) for n = 1 .. 2048
and LAPACK dgetrf's (scipy.linalg.lu_factor) at n = 4096; the factors reproduce P A = L U to
n * 2^-52 relative; solutions have normwise backward error <= 1e-14; the first exactly-zero
pivot matches the oracle's info; the determinant matches the oracle's factors under the same
rule; inverses satisfy A X = I.
"""
import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def _perm_apply(ipiv, A):
    P = np.array(A, copy=True)
    for i, p in enumerate(ipiv - 1):
        if p != i:
            P[[i, p]] = P[[p, i]]
    return P


def _factor_residual(A, LU, ipiv):
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    return np.abs(_perm_apply(ipiv, A) - L @ U).max() / max(np.abs(A).max(), 1e-300)


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 100, 513, 1000, 2048])
def test_getrf_pivots_match_oracle(n):
    from blaze_amd import solve
    rng = np.random.default_rng(1000 + n)
    A = rng.uniform(-1, 1, (n, n))
    print(f"getrf n={n}", flush=True)
    LU, ip, info = solve.getrf(ctx(), A)
    LUo, ipo, info_o = oracle.getrf(A)
    assert info == info_o == 0
    assert np.array_equal(ip, ipo), (n, np.nonzero(ip != ipo)[0][:5])
    assert _factor_residual(A, LU, ip) <= n * 2.0 ** -52 * 4
    assert np.abs(LU - LUo).max() <= 1e-9 * max(1.0, np.abs(LUo).max())


def test_getrf_4096_pivots_match_lapack():
    import scipy.linalg
    from blaze_amd import solve
    n = 4096
    rng = np.random.default_rng(4096)
    A = rng.standard_normal((n, n))
    print("getrf n=4096", flush=True)
    LU, ip, info = solve.getrf(ctx(), A)
    lu_s, piv_s = scipy.linalg.lu_factor(A)
    assert info == 0
    assert np.array_equal(ip - 1, piv_s)
    assert _factor_residual(A, LU, ip) <= n * 2.0 ** -52 * 4
    assert np.abs(LU - lu_s).max() <= 1e-8 * np.abs(lu_s).max()


def test_getrf_leading_dimension_and_alignment():
    from blaze_amd import solve
    rng = np.random.default_rng(77)
    A = rng.uniform(-1, 1, (601, 601))
    LU1, ip1, _ = solve.getrf(ctx(), A)
    LU2, ip2, _ = solve.getrf(ctx(), A, lda=611)
    assert np.array_equal(ip1, ip2)
    assert np.abs(LU1 - LU2).max() <= 1e-12 * np.abs(LU1).max()


@pytest.mark.parametrize("n,nrhs", [(1000, 64), (300, 16), (513, 1), (40, 9), (2048, 3)])
def test_gesv_any_order_and_rhs_count(n, nrhs):
    from blaze_amd import solve
    rng = np.random.default_rng(n * 7 + nrhs)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, nrhs))
    print(f"gesv n={n} nrhs={nrhs}", flush=True)
    X, _, ip, info = solve.gesv(ctx(), A, B)
    assert info == 0
    assert oracle.backward_error(A, B, X) <= 1e-14
    if n <= 1000:
        Xo, _, ipo, info_o = oracle.gesv(A, B)
        assert np.array_equal(ip, ipo)
        cond = np.linalg.cond(A)
        assert np.abs(X - Xo).max() <= 1e-13 * cond * max(1.0, np.abs(Xo).max())


def test_singular_reports_first_zero_pivot_and_keeps_b():
    from blaze_amd import solve
    rng = np.random.default_rng(700)
    n = 700
    A = rng.uniform(-1, 1, (n, n))
    A[:, 123] = 0.0                       # exactly singular: zero column
    B = rng.uniform(-1, 1, (n, 2))
    X, LU, ip, info = solve.gesv(ctx(), A, B)
    _, ipo, info_o = oracle.getrf(A)
    assert info == info_o == 124
    assert np.array_equal(ip[:123], ipo[:123])
    assert np.array_equal(X, B)           # dgesv leaves B unchanged when info > 0
    d, dinfo = solve.det(ctx(), A)
    assert d == 0.0 and dinfo == 124


def test_nan_on_the_diagonal_stays_the_pivot():
    from blaze_amd import solve
    rng = np.random.default_rng(50)
    A = rng.uniform(-1, 1, (50, 50))
    A[0, 0] = np.nan
    _, ip, _ = solve.getrf(ctx(), A)
    _, ipo, _ = oracle.getrf(A)
    assert np.array_equal(ip, ipo)


def test_det_matches_oracle_factors():
    from blaze_amd import solve
    rng = np.random.default_rng(600)
    for n in (600, 1100):
        A = np.eye(n) + 0.02 * rng.uniform(-1, 1, (n, n))
        A[[3, 7]] = A[[7, 3]]              # an odd permutation somewhere
        d, info = solve.det(ctx(), A)
        LUo, ipo, info_o = oracle.getrf(A)
        do = 1.0
        for i in range(n):                 # detNxN's rule on the oracle's factors
            do *= LUo[i, i]
            if ipo[i] != i + 1:
                do = -do
        assert info == info_o == 0
        assert abs(d - do) <= 1e-11 * abs(do), (n, d, do)
        sign, logdet = np.linalg.slogdet(A)
        assert np.sign(d) == sign and abs(np.log(abs(d)) - logdet) <= 1e-10


def test_inverse_any_order():
    from blaze_amd import solve
    rng = np.random.default_rng(900)
    for n in (513, 900):
        A = rng.uniform(-1, 1, (n, n)) + n ** 0.5 * np.eye(n)
        X, info = solve.inv(ctx(), A)
        assert info == 0
        assert np.abs(A @ X - np.eye(n)).max() <= 1e-12
        Xn = np.linalg.inv(A)
        assert np.abs(X - Xn).max() <= 1e-12 * np.linalg.cond(A) * np.abs(Xn).max()
