"""CPU checks of the synthetic-input generators: the Haar-random orthogonal factors (Stewart's
method) and the condition-1e8 solve systems built from them."""
import ctypes as C

import numpy as np
import pytest

import blaze_amd as bz


def _gen():
    if not bz.GEN_PATH.exists():
        pytest.skip("generators not built")
    return bz.gen()


def _haar(n, stream, seed=99):
    Q = np.empty((n, n))
    _gen().bzg_gen_haar_orthogonal(C.c_uint64(seed), C.c_uint64(stream), C.c_int32(n), Q.ctypes.data_as(C.c_void_p))
    return Q


def test_haar_factor_is_orthogonal():
    for n in (1, 2, 3, 17, 64, 256):
        Q = _haar(n, 5)
        assert np.abs(Q @ Q.T - np.eye(n)).max() < 1e-13 * max(n, 4)


def test_haar_moments_match_the_uniform_measure():
    # Under the Haar measure on O(n): E[q_ij^2] = 1/n, E[q_ij^4] = 3/(n(n+2)), and det = +1 or -1
    # with probability 1/2 each.  A structured factor (e.g. a few reflectors applied to the
    # identity, which leaves most of the mass on the diagonal) fails these.
    n, trials = 12, 3000
    diag2, off2, off4, dets = [], [], [], []
    for t in range(trials):
        Q = _haar(n, 1000 + t)
        diag2.append(Q[3, 3] ** 2)
        off2.append(Q[0, 5] ** 2)
        off4.append(Q[7, 2] ** 4)
        dets.append(np.sign(np.linalg.det(Q)))
    se2 = np.sqrt((3.0 / (n * (n + 2)) - 1.0 / n ** 2) / trials)
    assert abs(np.mean(diag2) - 1.0 / n) < 5 * se2
    assert abs(np.mean(off2) - 1.0 / n) < 5 * se2
    m4 = 3.0 / (n * (n + 2))
    assert abs(np.mean(off4) - m4) < 0.15 * m4
    assert abs(np.mean(dets)) < 5 / np.sqrt(trials)


def test_condition_1e8_systems_have_the_stated_spectrum():
    g = _gen()
    ns = np.array([5, 40, 129, 256], np.int32)
    r = np.array([1, 3, 8, 2], np.int32)
    kind = np.ones(len(ns), np.int8)
    a_off = np.zeros(len(ns), np.int64); b_off = np.zeros(len(ns), np.int64)
    a_off[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    b_off[1:] = np.cumsum(ns.astype(np.int64) * r)[:-1]
    A = np.empty(int((ns.astype(np.int64) ** 2).sum())); B = np.empty(int((ns * r).sum()))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    g.bzg_gen_solve_batch(C.c_uint64(7), C.c_int64(len(ns)), p(ns), p(r), p(kind), p(a_off), p(b_off), p(A), p(B))
    for i, n in enumerate(ns):
        M = A[a_off[i]:a_off[i] + n * n].reshape(n, n)
        sv = np.linalg.svd(M, compute_uv=False)
        want = 10.0 ** (-8.0 * np.arange(n) / (n - 1))
        assert np.allclose(sv, want, rtol=1e-8, atol=0)
        # the factors are dense: about 1/n of the Frobenius mass sits on the diagonal (a diagonal
        # plus a low-rank change keeps most of it there)
        if n >= 40:
            assert np.sum(np.diag(M) ** 2) / np.sum(M ** 2) < 4.0 / n
