"""GPU test of bzg_dgesv_batched_noalloc (the batched FP64 solve whose scratch is a fixed device
array): two calls in a row, the second with a smaller batch, each bit-identical to
bzg_dgesv_batched on the same systems (solutions, factors, pivots, info), and within the
oracle's FP64 tolerance.  Collected last so that a failure here hides no other test."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import ctx, dev, have_gpu, host, p, torch

import blaze_amd as bz
import oracle

pytestmark = pytest.mark.gpu


def _batch(rng, sizes, nrhs):
    As = [rng.uniform(-1, 1, (s, s)) + (s if i % 3 else 0.5) * np.eye(s) for i, s in enumerate(sizes)]
    Bs = [rng.uniform(-1, 1, (s, r)) for s, r in zip(sizes, nrhs)]
    return As, Bs


def _run(fn, As, Bs, order_host):
    n = np.array([a.shape[0] for a in As], np.int32)
    r = np.array([b.shape[1] for b in Bs], np.int32)
    a_off = np.zeros(len(As), np.int64); b_off = np.zeros(len(As), np.int64); p_off = np.zeros(len(As), np.int64)
    a_off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    b_off[1:] = np.cumsum(n.astype(np.int64) * r)[:-1]
    p_off[1:] = np.cumsum(n.astype(np.int64))[:-1]
    dA = dev(np.concatenate([a.ravel() for a in As]))
    dB = dev(np.concatenate([b.ravel() for b in Bs]))
    dn, dr, da, db, dp = dev(n), dev(r), dev(a_off), dev(b_off), dev(p_off)
    dpiv = torch.zeros(int(n.sum()), dtype=torch.int32, device="cuda")
    dinfo = torch.full((len(As),), -99, dtype=torch.int32, device="cuda")
    dord = dev(order_host)
    ctx().call(fn, len(As), p(dn), p(dr), p(da), p(db), p(dp), p(dA), p(dB), p(dpiv), p(dinfo), p(dord))
    ctx().sync()
    return host(dA), host(dB), host(dpiv), host(dinfo), (n, r, b_off)


@pytest.mark.skipif(not have_gpu(), reason="needs a GPU")
def test_noalloc_matches_dgesv_batched_two_calls():
    rng = np.random.default_rng(20261203)
    # first call: orders across both kernel classes (<= 256 and 257..512), 1..8 RHS
    sizes1 = [1, 2, 5, 7, 16, 33, 64, 100, 255, 256, 257, 300, 512] + list(rng.integers(3, 200, 40))
    # second call: a smaller batch, orders <= 256 only, so stale list entries of the first call
    # (systems of order > 256) would show up as wrong results if the scratch were not reset
    sizes2 = list(rng.integers(2, 130, 9))
    for sizes in (sizes1, sizes2):
        nrhs = list(rng.integers(1, 9, len(sizes)))
        As, Bs = _batch(rng, [int(s) for s in sizes], [int(x) for x in nrhs])
        order = np.argsort(-np.array(sizes), kind="stable").astype(np.int32)
        print(f"noalloc case: {len(sizes)} systems, orders {min(sizes)}..{max(sizes)}", flush=True)
        ref = _run("bzg_dgesv_batched", As, Bs, order)
        got = _run("bzg_dgesv_batched_noalloc", As, Bs, order)
        for a, b in zip(ref[:4], got[:4]):
            assert np.array_equal(a, b)
        n, r, b_off = got[4]
        assert np.all(got[3] == 0)
        X = got[1]
        for i, (a, bb) in enumerate(zip(As, Bs)):
            xo, _, _, info_o = oracle.gesv(a, bb)
            assert info_o == 0
            xi = X[b_off[i]:b_off[i] + n[i] * r[i]].reshape(n[i], r[i])
            assert np.abs(xi - xo).max() <= 1e-10 * max(1.0, np.abs(xo).max())


def test_noalloc_rejects_oversized_batch_without_launching():
    if not bz.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = bz.lib()
    # argument checks run before any device work: a null context, then a batch beyond the limit
    assert lib.bzg_dgesv_batched_noalloc(None, 1, *([None] * 10)) == -1
    if have_gpu():
        st = lib.bzg_dgesv_batched_noalloc(ctx().handle, (1 << 20) + 1, *([None] * 10))
        assert st == -6 and b"BZG_NOALLOC_MAX_SYSTEMS" in lib.bzg_last_error()
