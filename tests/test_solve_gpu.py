"""GPU parity: batched LU solves (dgesv semantics) and the mixed-precision solver (dsgesv).

Mirrors the reference's lse/lu tests (blazetest/src/mathtest/operations/lse/DenseTest.cpp:71
runs N = 0..12; lse/DenseTest.h:140-200 checks A*x == b; lu/DenseTest.h:159-170 checks the
reconstruction of A from L, U and the permutation) against the CPU oracle, which follows
LAPACK's dgetf2 pivot rule (first row of maximal |a|).
Tolerances: pivots (ipiv) identical; LU factors within 1e-12 * max|A| (rounding order of the
blocked update differs from the unblocked oracle); solutions within 1e-10 relative for the
diagonally dominant systems; backward error <= 1e-12 for the mixed-precision solver
(its convergence target); exact singularity reported as info = first zero pivot.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def systems(rng, sizes, nrhs=2, cond=False):
    As, Bs = [], []
    for n in sizes:
        A = rng.uniform(-1, 1, (n, n)) + (0 if cond else n) * np.eye(n)
        As.append(A)
        Bs.append(rng.uniform(-1, 1, (n, nrhs)))
    return As, Bs


def test_gesv_small_sizes_match_oracle_pivots():
    from blaze_amd import solve
    rng = np.random.default_rng(40)
    sizes = list(range(1, 13)) + [15, 16, 17, 31, 32, 33, 64, 100, 127, 128, 129, 200, 255, 256]
    # general (not diagonally dominant) matrices exercise the pivoting
    As = [rng.uniform(-1, 1, (n, n)) for n in sizes]
    Bs = [rng.uniform(-1, 1, (n, 3)) for n in sizes]
    Xs, infos, LUs, ipivs = solve.gesv_batched(ctx(), As, Bs, with_ipiv=True)
    for A, B, X, info, LU, ip in zip(As, Bs, Xs, infos, LUs, ipivs):
        n = A.shape[0]
        Xo, LUo, ipo, info_o = oracle.gesv(A, B)
        assert info == info_o == 0, n
        assert np.array_equal(ip, ipo), n
        assert np.abs(LU - LUo).max() <= 1e-12 * max(1, np.abs(LUo).max()) * n, n
        be = oracle.backward_error(A, B, X)
        assert be <= 1e-14 * n, (n, be)
        # P*A = L*U reconstruction (lu/DenseTest.h:159-170)
        L = np.tril(LU, -1) + np.eye(n); U = np.triu(LU)
        PA = A.copy()
        for i, p in enumerate(ip - 1):
            PA[[i, p]] = PA[[p, i]]
        assert np.abs(L @ U - PA).max() <= 1e-13 * n * np.abs(A).max()


def test_gesv_exact_integer_system():
    # Integer matrix with a known integer solution: x = (1, -2, 3, ...).
    from blaze_amd import solve
    rng = np.random.default_rng(41)
    for n in (7, 20, 50):
        A = rng.integers(-5, 6, (n, n)).astype(np.float64) + 30 * np.eye(n)
        x = np.arange(1, n + 1, dtype=np.float64) * (-1) ** np.arange(n)
        B = (A @ x)[:, None]
        Xs, infos = solve.gesv_batched(ctx(), [A], [B])
        assert infos[0] == 0
        assert np.abs(Xs[0][:, 0] - x).max() < 1e-12 * n


def test_gesv_singular_reports_info():
    from blaze_amd import solve
    A1 = np.array([[1.0, 2.0], [2.0, 4.0]])
    A2 = np.zeros((9, 9)); A2[np.arange(9), np.arange(9)] = 1.0; A2[4, 4] = 0.0
    A3 = np.ones((12, 12))
    Xs, infos = solve.gesv_batched(ctx(), [A1, A2, A3], [np.ones((2, 1)), np.ones((9, 1)), np.ones((12, 1))])
    for A, info in zip([A1, A2, A3], infos):
        _, _, _, info_o = oracle.gesv(A, np.ones((A.shape[0], 1)))
        assert info == info_o and info > 0


def test_gesv_uneven_batch_longest_first():
    from blaze_amd import solve
    rng = np.random.default_rng(42)
    sizes = rng.integers(4, 257, 600)
    nr = rng.integers(1, 9, 600)
    As = [rng.uniform(-1, 1, (n, n)) + n * np.eye(n) for n in sizes]
    Bs = [rng.uniform(-1, 1, (n, r)) for n, r in zip(sizes, nr)]
    Xs, infos = solve.gesv_batched(ctx(), As, Bs)
    assert (infos == 0).all()
    for i in range(0, 600, 37):
        Xo, _, _, _ = oracle.gesv(As[i], Bs[i])
        assert np.abs(Xs[i] - Xo).max() <= 1e-12 * max(1, np.abs(Xo).max())
    Xs2, _ = solve.gesv_batched(ctx(), As, Bs, lpt=False)
    assert all(np.array_equal(a, b) for a, b in zip(Xs, Xs2))  # order does not change results


def test_dsgesv_well_and_ill_conditioned():
    from blaze_amd import solve
    rng = np.random.default_rng(43)
    As, Bs = systems(rng, [4, 9, 33, 100, 180, 256], nrhs=3)
    # kappa = 1e8 and 1e8 (refinement may still converge: kappa*2^-24 ~ 6 is a worst-case
    # bound) and kappa = 1e13 (cannot converge: must fall back to the FP64 LU)
    for n, lo in ((40, -8), (128, -8), (64, -13)):
        Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
        Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
        As.append(Q1 @ np.diag(np.logspace(0, lo, n)) @ Q2.T)
        Bs.append(rng.uniform(-1, 1, (n, 2)))
    Xs, iters, infos = solve.dsgesv_batched(ctx(), As, Bs)
    assert (infos == 0).all()
    for A, B, X, it in zip(As, Bs, Xs, iters):
        be = oracle.backward_error(A, B, X)
        assert be <= 1e-12, (A.shape, be, it)
        _, it_o, _ = oracle.dsgesv(A, B)
        assert (it >= 0) == (it_o >= 0), (A.shape, it, it_o)
    assert (iters[:6] >= 0).all() and (iters[:6] <= 5).all()
    assert iters[8] == -3  # kappa = 1e13: refinement stalls -> FP64 fallback


def test_orders_above_256_and_out_of_range():
    # n in (256, 512] runs on the second kernel instance; both instances share one batch.
    from blaze_amd import solve
    rng = np.random.default_rng(44)
    sizes = [257, 300, 5, 384, 511, 64, 512, 256]
    As = [rng.uniform(-1, 1, (n, n)) for n in sizes]        # general: exercises pivoting
    Bs = [rng.uniform(-1, 1, (n, 2)) for n in sizes]
    Xs, infos, LUs, ipivs = solve.gesv_batched(ctx(), As, Bs, with_ipiv=True)
    for A, B, X, info, LU, ip in zip(As, Bs, Xs, infos, LUs, ipivs):
        n = A.shape[0]
        Xo, LUo, ipo, info_o = oracle.gesv(A, B)
        assert info == info_o == 0, n
        assert np.array_equal(ip, ipo), n
        assert np.abs(LU - LUo).max() <= 1e-12 * max(1, np.abs(LUo).max()) * n, n
        assert oracle.backward_error(A, B, X) <= 1e-14 * n, n
    # mixed precision on the same mix (diagonally dominant so FP32 refinement converges)
    Ad = [A + A.shape[0] * np.eye(A.shape[0]) for A in As]
    Xm, iters, infom = solve.dsgesv_batched(ctx(), Ad, Bs)
    assert (infom == 0).all() and (iters >= 0).all()
    for A, B, X in zip(Ad, Bs, Xm):
        assert oracle.backward_error(A, B, X) <= 1e-12
    # orders beyond 512 and more than 8 right-hand sides are rejected per system (info = -1)
    bad_A = [np.eye(513), np.eye(3), np.eye(4)]
    bad_B = [np.ones((513, 1)), np.ones((3, 9)), np.ones((4, 1))]
    _, infob = solve.gesv_batched(ctx(), bad_A, bad_B)
    assert list(infob) == [-1, -1, 0]
    _, itb, infob2 = solve.dsgesv_batched(ctx(), bad_A, bad_B)
    assert list(infob2) == [-1, -1, 0] and itb[0] == 0 and itb[1] == 0


def test_getrf_getrs_match_gesv_and_oracle():
    # dgetrf factors / pivots == the oracle's dgetf2 (pivots identical, factors within the
    # blocked-update rounding), dgetrs with those factors == the gesv solution bit for bit
    # (same kernels), and == the oracle's dgetrs on the same factors within rounding
    from blaze_amd import solve
    rng = np.random.default_rng(46)
    sizes = [1, 2, 3, 7, 16, 33, 64, 100, 255, 256, 300, 512]
    As = [rng.uniform(-1, 1, (n, n)) for n in sizes]
    Bs = [rng.uniform(-1, 1, (n, int(rng.integers(1, 9)))) for n in sizes]
    LUs, ipivs, infos = solve.getrf_batched(ctx(), As)
    Xg, infog, LUg, ipg = solve.gesv_batched(ctx(), As, Bs, with_ipiv=True)
    Xs, infos2 = solve.getrs_batched(ctx(), LUs, ipivs, Bs)
    for A, B, LU, ip, X, Xr, LU2, ip2, n in zip(As, Bs, LUs, ipivs, Xs, Xg, LUg, ipg, sizes):
        LUo, ipo, info_o = oracle.getrf(A)
        assert info_o == 0
        assert np.array_equal(ip, ipo), n
        assert np.array_equal(LU, LU2) and np.array_equal(ip, ip2), n   # same factors as gesv
        assert np.abs(LU - LUo).max() <= 1e-12 * max(1, np.abs(LUo).max()) * n, n
        assert np.array_equal(X, Xr), n                                  # same solve as gesv
        Xo = oracle.getrs(LU, ip, B)
        assert np.allclose(X, Xo, rtol=1e-10, atol=1e-12), n
    assert (infos == 0).all() and (infos2 == 0).all() and (infog == 0).all()


def test_getrf_singular_and_out_of_range():
    from blaze_amd import solve
    A = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])  # rank 2
    LUs, ipivs, infos = solve.getrf_batched(ctx(), [A, np.eye(4)])
    _, ipo, info_o = oracle.getrf(A)
    assert infos[0] == info_o > 0 and infos[1] == 0
    assert np.array_equal(ipivs[0], ipo)


def test_lu_rowmajor_convention():
    # lu(A, L, U, P) for a row-major matrix (LU.h:98-140, 222-330): A = L * U * P, L lower,
    # U unit upper, P a column permutation -- the factorisation of A^T by the oracle's dgetf2
    from blaze_amd import solve
    rng = np.random.default_rng(47)
    for n in (1, 2, 5, 17, 64, 200):
        A = rng.uniform(-1, 1, (n, n))
        L, U, P, info = solve.lu_rowmajor(ctx(), A)
        assert info == 0
        assert np.array_equal(L, np.tril(L)) and np.array_equal(U, np.triu(U))
        assert np.array_equal(np.diag(U), np.ones(n))
        assert np.array_equal(P @ P.T, np.eye(n)) and set(np.unique(P)) <= {0.0, 1.0}
        assert np.abs(L @ U @ P - A).max() <= 1e-13 * n * max(1, np.abs(A).max()), n
        LUo, ipo, _ = oracle.getrf(np.ascontiguousarray(A.T))
        assert np.abs(np.tril(LUo.T) - L).max() <= 1e-12 * n, n     # L = U_o^T
        permut = np.arange(n)
        for i, p in enumerate(ipo - 1):
            if p != i:
                permut[[i, p]] = permut[[p, i]]
        assert np.array_equal(np.argmax(P, axis=1), permut), n


def test_transpose():
    import torch
    rng = np.random.default_rng(48)
    for m, n, pad in ((1, 1, 0), (3, 70, 1), (129, 65, 3), (1000, 7, 0)):
        A = rng.uniform(-1, 1, (m, n + pad))
        dA = torch.from_numpy(A).cuda()
        dB = torch.full((n, m + pad), 7.0, dtype=torch.float64, device="cuda")
        ctx().call("bzg_dtranspose", m, n, C.c_void_p(dA.data_ptr()), n + pad, C.c_void_p(dB.data_ptr()), m + pad)
        B = dB.cpu().numpy()
        assert np.array_equal(B[:, :m], A[:, :n].T)
        assert (B[:, m:] == 7.0).all()
