"""GPU parity of the sliced-ELL SpMV (bzg_csr_plan_pack_sell + bzg_dsellmv / bzg_ssellmv /
bzg_isellmv) for all-short-row matrices: every row must be the reference's sequential sum in
storage order (SMatDVecMultExpr.h:204-207 -> SVecDVecInnerExpr.h:127-135), bit for bit,
including the sign of zero sums and empty rows.

Checks: the reference's own y = S * x on the 16^3 Laplacian (tests/golden/spmv_laplace16.npz,
double and int32), the oracle's sequential loop on random short rows (0..32 nonzeros, empty
rows, row counts that leave the last 64-row slice partial, NaN and infinity in x), equality
with bzg_*csrmv on the same plan (also with sums started from +0 for column-major products),
repacking after the values change, and the error paths.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.gpu_util import ctx, dev, empty, full, have_gpu, host, p

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from inputs import ints, laplace3d, load, uniform  # noqa: E402

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

DTYPE = {np.float64: 0, np.float32: 1, np.int32: 2}
SELLMV = {np.float64: "bzg_dsellmv", np.float32: "bzg_ssellmv", np.int32: "bzg_isellmv"}
CSRMV = {np.float64: "bzg_dcsrmv", np.float32: "bzg_scsrmv", np.int32: "bzg_icsrmv"}


class Plan:
    def __init__(self, rowptr, n_cols):
        import blaze_amd as bz
        self.bz = bz
        self.h = C.c_void_p()
        bz.check("bzg_csr_plan_create", bz.lib().bzg_csr_plan_create(
            ctx().handle, len(rowptr) - 1, n_cols, rowptr.ctypes.data_as(C.c_void_p), C.byref(self.h)))

    def __del__(self):
        self.bz.lib().bzg_csr_plan_destroy(self.h)


def _dev(a):
    return dev(a) if a.size else empty((1,), a.dtype)


def sell_and_csr(rowptr, col, val, x, n_cols, plus_zero=False):
    """(y by sliced ELL, y by CSR) on one plan."""
    import blaze_amd as bz
    plan = Plan(rowptr, n_cols)
    if plus_zero:
        bz.check("sum start", bz.lib().bzg_csr_plan_set_sum_start(plan.h, 1))
    n = len(rowptr) - 1
    d = [_dev(a) for a in (rowptr, col, val, x)]
    bz.check("pack", bz.lib().bzg_csr_plan_pack_sell(ctx().handle, plan.h, p(d[0]), p(d[1]), p(d[2]),
                                                     DTYPE[val.dtype.type]))
    y1 = full((max(n, 1),), val.dtype, -5)
    y2 = full((max(n, 1),), val.dtype, -5)
    ctx().call(SELLMV[val.dtype.type], plan.h, p(d[3]), p(y1))
    ctx().call(CSRMV[val.dtype.type], plan.h, *[p(t) for t in d], p(y2))
    return host(y1)[:n], host(y2)[:n]


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def test_sell_laplace16_vs_reference():
    g = load("spmv_laplace16.npz")
    rp, ci, v = laplace3d(int(g["nx"]))
    n = len(rp) - 1
    x = uniform(int(g["x_seed"]), 0, n)
    y, yc = sell_and_csr(rp, ci, v, x, n)
    assert same_bits(y, g["y"]) and same_bits(y, yc)
    xi = ints(int(g["x_seed"]), 1, n, -100, 100)
    yi, _ = sell_and_csr(rp, ci, v.astype(np.int32), xi, n)
    assert np.array_equal(yi, g["y_int32"])


def _random_short(rng, n, ncols, maxlen=32, empty_frac=0.1):
    lens = rng.integers(0, maxlen + 1, n)
    lens[rng.random(n) < empty_frac] = 0
    lens = np.minimum(lens, ncols)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(ncols, int(c), replace=False)) for c in lens] or
                        [np.zeros(0, np.int64)]).astype(np.int32)
    return rp, ci


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int32])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 20011])
def test_sell_random_short_rows_bit_exact(dt, n):
    rng = np.random.default_rng(n * 7 + DTYPE[dt])
    ncols = max(40, n // 3)
    rp, ci = _random_short(rng, n, ncols)
    if dt == np.int32:
        va = rng.integers(-50, 50, int(rp[-1])).astype(np.int32)
        x = rng.integers(-50, 50, ncols).astype(np.int32)
    else:
        va = rng.standard_normal(int(rp[-1])).astype(dt)
        x = rng.uniform(-1, 1, ncols).astype(dt)
        va[::17] = -0.0  # signed-zero products
    y, yc = sell_and_csr(rp, ci, va, x, ncols)
    ref = oracle.csrmv(rp, ci, va, x)
    assert same_bits(y, ref), (dt, n)
    assert same_bits(y, yc), (dt, n)


def test_sell_nonfinite_x_and_plus_zero_start():
    rng = np.random.default_rng(9)
    n, ncols = 3000, 500
    rp, ci = _random_short(rng, n, ncols, maxlen=20)
    va = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, ncols)
    x[[3, 77]] = np.inf
    x[[11]] = np.nan
    x[[5]] = -0.0
    y, yc = sell_and_csr(rp, ci, va, x, ncols)
    assert same_bits(y, oracle.csrmv(rp, ci, va, x)) and same_bits(y, yc)
    # sums from +0 (the column-major product's convention): identical to the CSR kernels
    y0, yc0 = sell_and_csr(rp, ci, va, x, ncols, plus_zero=True)
    assert same_bits(y0, yc0)


def test_sell_repack_and_errors():
    import blaze_amd as bz
    L = bz.lib()
    rng = np.random.default_rng(4)
    n = 500
    rp, ci = _random_short(rng, n, n, maxlen=9)
    va = rng.standard_normal(int(rp[-1]))
    x = rng.uniform(-1, 1, n)
    plan = Plan(rp, n)
    d = [_dev(a) for a in (rp, ci, va, x)]
    y = full((n,), np.float64, 0)
    # not packed yet, and a dtype other than the packed one: argument errors
    assert L.bzg_dsellmv(ctx().handle, plan.h, p(d[3]), p(y)) != 0
    bz.check("pack", L.bzg_csr_plan_pack_sell(ctx().handle, plan.h, p(d[0]), p(d[1]), p(d[2]), 0))
    assert L.bzg_ssellmv(ctx().handle, plan.h, p(d[3]), p(y)) != 0
    ctx().call("bzg_dsellmv", plan.h, p(d[3]), p(y))
    assert same_bits(host(y), oracle.csrmv(rp, ci, va, x))
    # new values: the copy is stale until packed again
    va2 = va * 3.0 - 1.0
    d[2] = _dev(va2)
    bz.check("repack", L.bzg_csr_plan_pack_sell(ctx().handle, plan.h, p(d[0]), p(d[1]), p(d[2]), 0))
    ctx().call("bzg_dsellmv", plan.h, p(d[3]), p(y))
    assert same_bits(host(y), oracle.csrmv(rp, ci, va2, x))
    # a row longer than 32 nonzeros: the sliced layout is refused
    lens = np.full(70, 3)
    lens[10] = 40
    rp2 = np.zeros(71, np.int32); rp2[1:] = np.cumsum(lens)
    ci2 = np.concatenate([np.arange(c) for c in lens]).astype(np.int32)
    plan2 = Plan(rp2, 64)
    d2 = [_dev(a) for a in (rp2, ci2, np.ones(int(rp2[-1])))]
    assert L.bzg_csr_plan_pack_sell(ctx().handle, plan2.h, p(d2[0]), p(d2[1]), p(d2[2]), 0) != 0


def test_sell_laplacian_128_full_size_integer_exact():
    # the benchmark's 128^3 matrix: integer x makes every row sum exact (6 x_i minus neighbours),
    # and x = 1 gives each row's number of missing neighbours (0 inside, up to 3 at corners)
    rp, ci, va = laplace3d(128)
    n = len(rp) - 1
    x = np.random.default_rng(33).integers(-1000, 1000, n).astype(np.float64)
    y, yc = sell_and_csr(rp, ci, va, x, n)
    assert same_bits(y, oracle.csrmv(rp, ci, va, x)) and same_bits(y, yc)
    ones, _ = sell_and_csr(rp, ci, va, np.ones(n), n)
    assert ones.min() == 0 and ones.max() == 3
