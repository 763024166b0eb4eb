"""GPU parity: batched inverses (bzg_dinv_batched), the reference's invert()/inv() by LU
(blaze/math/dense/Inversion.h:3337-3347: getrf, then getri).

Pins:
* numpy.linalg.inv (LAPACK dgesv against the identity) on well-conditioned random matrices
  from n = 1 to 512, within 1e-12 relative in max norm, and the residual |A X - I| within
  n u |A| |X| (Higham's bound for a backward-stable inverse, with slack);
* exact cases: a permutation matrix inverts to its transpose, a diagonal of powers of two to
  its reciprocals, bit for bit;
* a singular system (zero column) keeps bzg_dgetrf_batched's factors and info, bit for bit;
* n = 0 gives info 0; n > 512 gives info -1 and leaves A alone.
"""
import numpy as np
import pytest

from tests.gpu_util import have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def _ctx():
    from tests.gpu_util import ctx
    return ctx()


def test_inverse_against_lapack_and_exact_cases():
    from blaze_amd.solve import getrf_batched, inv_batched
    rng = np.random.default_rng(91)
    sizes = [1, 2, 3, 7, 8, 9, 16, 31, 64, 100, 200, 256, 257, 333, 512]
    As = [np.eye(n) * 2.0 + rng.uniform(-1, 1, (n, n)) / np.sqrt(n) for n in sizes]
    P = np.eye(50)[rng.permutation(50)]
    D = np.diag(2.0 ** rng.integers(-20, 21, 12).astype(np.float64))
    S = rng.uniform(-1, 1, (24, 24)); S[:, 10] = 0.0
    big = rng.uniform(-1, 1, (513, 513))
    batch = As + [P, D, S, np.zeros((0, 0)), big]
    Xs, infos = inv_batched(_ctx(), batch)
    for A, X, info in zip(As, Xs, infos):
        n = A.shape[0]
        assert info == 0
        ref = np.linalg.inv(A)
        assert np.abs(X - ref).max() <= 1e-12 * np.abs(ref).max(), n
        res = np.abs(A @ X - np.eye(n))
        assert np.all(res <= 4 * n * 2.0 ** -53 * (np.abs(A) @ np.abs(X)) + 1e-300), n
    k = len(As)
    assert infos[k] == 0 and np.array_equal(Xs[k], P.T)
    assert infos[k + 1] == 0 and np.array_equal(Xs[k + 1], np.diag(1.0 / np.diag(D)))
    LUs, ipivs, finfo = getrf_batched(_ctx(), [S])
    assert infos[k + 2] == finfo[0] == 11 and np.array_equal(Xs[k + 2], LUs[0])
    assert infos[k + 3] == 0 and Xs[k + 3].shape == (0, 0)
    assert infos[k + 4] == -1 and np.array_equal(Xs[k + 4], big)


def test_inverse_many_small_systems_deterministic():
    from blaze_amd.solve import inv_batched
    rng = np.random.default_rng(92)
    As = [np.eye(n) * 3.0 + rng.uniform(-1, 1, (n, n)) for n in rng.integers(1, 40, 300)]
    X1, i1 = inv_batched(_ctx(), As)
    X2, i2 = inv_batched(_ctx(), As, lpt=False)
    assert np.all(i1 == 0) and np.array_equal(i1, i2)
    for A, a, b in zip(As, X1, X2):
        assert np.array_equal(a, b)  # the order the workgroups take systems in does not matter
        n = A.shape[0]
        assert np.all(np.abs(A @ a - np.eye(n)) <= 4 * n * 2.0 ** -53 * (np.abs(A) @ np.abs(a)) + 1e-300)
