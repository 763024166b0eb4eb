"""GPU parity: y = trans(S) * x / a column-major sparse matrix times a vector, the reference's
TSMatDVecMultExpr (blaze/math/expressions/TSMatDVecMultExpr.h:329-383), against the golden
fixture written by the reference's own compiled kernel (tests/golden/spmv_trans.npz).

The device path: bzg_csc_to_csr_host packs S's CSR, read as the CSC of trans(S), into the CSR of
trans(S) (each row lists S's rows ascending), a plan with bzg_csr_plan_set_sum_start(plan, 1),
and bzg_dcsrmv / bzg_scsrmv.  Rows of trans(S) with up to 32 entries are summed sequentially
from +0 in the reference's order: bit-exact, including the sign of zero; longer rows use the
fixed tree order, within 2 gamma_len sum|v||x| of the reference.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

import blaze_amd as bz
from tests.gpu_util import ctx, dev, empty, full, have_gpu, host, p

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from inputs import load, uniform, zipf  # noqa: E402

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

FN = {np.dtype(np.float64): "bzg_dcsrmv", np.dtype(np.float32): "bzg_scsrmv"}


def pack_trans(rp, ci, v, n_cols):
    """CSR of trans(S) by the library's host pack."""
    n_rows = len(rp) - 1
    trp = np.zeros(n_cols + 1, np.int32)
    tci = np.zeros(max(len(v), 1), np.int32)
    tv = np.zeros(max(len(v), 1), v.dtype)
    q = lambda a: a.ctypes.data_as(C.c_void_p)
    bz.check("bzg_csc_to_csr_host", bz.lib().bzg_csc_to_csr_host(n_cols, n_rows, q(rp), q(ci), q(v), v.itemsize,
                                                                  q(trp), q(tci), q(tv)))
    return trp, tci[:len(v)], tv[:len(v)]


def gpu_trans_mv(rp, ci, v, x, n_cols):
    trp, tci, tv = pack_trans(rp, ci, v, n_cols)
    plan = C.c_void_p()
    bz.check("plan", bz.lib().bzg_csr_plan_create(ctx().handle, n_cols, len(rp) - 1,
                                                  trp.ctypes.data_as(C.c_void_p), C.byref(plan)))
    try:
        bz.check("sum start", bz.lib().bzg_csr_plan_set_sum_start(plan, 1))
        d = [dev(a) if a.size else empty((1,), a.dtype) for a in (trp, tci, tv, x)]
        dy = full((max(n_cols, 1),), v.dtype, -5)
        ctx().call(FN[v.dtype], plan, *[p(t) for t in d], p(dy))
        return host(dy)[:n_cols], np.diff(trp)
    finally:
        bz.lib().bzg_csr_plan_destroy(plan)


def _case(g, tag):
    seed, n, kmax = (int(v) for v in g[f"{tag}_params"])
    rp, ci, v = zipf(seed, n, float(g[f"{tag}_s"]), kmax)
    return rp, ci, v, uniform(seed, int(g["x_stream"]), n), n


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trans_spmv_vs_reference_golden(dt):
    g = load("spmv_trans.npz")
    suffix = "_y" if dt == np.float64 else "_y_f32"
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    for tag in ("short", "zipf"):
        rp, ci, v, x, n = _case(g, tag)
        v, x = v.astype(dt), x.astype(dt)
        got, lens = gpu_trans_mv(rp, ci, v, x, n)
        want = g[tag + suffix]
        short = lens <= 32
        if tag == "short":
            assert np.all(short)
        # bit-exact (same bits, so also the same sign of zero) on sequentially summed rows
        assert np.array_equal(got[short].view(np.uint8), want[short].view(np.uint8)), tag
        if not np.all(short):
            trp, tci, tv = pack_trans(rp, ci, v, n)
            absum = np.array([np.abs(tv[trp[j]:trp[j + 1]].astype(np.float64)) @
                              np.abs(x[tci[trp[j]:trp[j + 1]]].astype(np.float64)) for j in range(n)])
            bound = 2 * lens * u / (1 - lens * u) * absum
            assert np.all(np.abs(got.astype(np.float64) - want.astype(np.float64))[~short] <= bound[~short]), tag


def test_trans_spmv_signed_zero_and_empty():
    # entries whose products are all -0 come out +0 (the reference starts from a reset +0);
    # empty rows of trans(S) (columns of S without entries) are +0; an all-empty S too
    rp = np.array([0, 2, 3, 3, 5], np.int32)
    ci = np.array([0, 2, 2, 0, 2], np.int32)
    v = np.array([-1.0, 2.0, -3.0, -4.0, 5.0])
    x = np.array([0.0, 0.0, 7.0, 0.0])
    got, _ = gpu_trans_mv(rp, ci, v, x, 4)
    assert np.array_equal(got, [0.0, 0.0, 0.0, 0.0]) and not np.any(np.signbit(got))
    got2, _ = gpu_trans_mv(np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0), np.ones(3), 5)
    assert np.array_equal(got2, np.zeros(5)) and not np.any(np.signbit(got2))
