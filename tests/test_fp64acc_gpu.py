"""GPU parity: FP64-accurate GEMM from BF16 / INT8 residue slices (Ozaki scheme II).

No reference function exists for this path; its oracle is the exact product, computed on the
CPU in double-double (oracle.gemm_dd_entries: TwoProduct/TwoSum, ~106 bits) and, for integer
inputs, exactly.  Bars:
  * integer-valued inputs: bit-exact (C' is exact, the final rounding is exact);
  * uniform[-1,1) inputs (sigma = 0): relative Frobenius error <= 1e-14 against the exact
    product (the configured target), on sampled entries;
  * graded exponents (sigma = 4, 10): every entry within the scheme's bound
    k * 2^-P * max|A_i,:| * max|B_:,j| + |C_ij| 2^-52 (P of the plan actually used), and the
    automatic moduli count keeps the relative Frobenius error <= 1e-14 for sigma in {0, 4, 10}.
"""
import ctypes as C

import numpy as np
import pytest

import blaze_amd as bz
import oracle
from tests.gpu_util import ctx, dev, full, have_gpu, host, p

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]

FNS = ["bzg_dgemm_fp64acc", "bzg_dgemm_fp64acc_i8"]


def gpu_fp64acc(fn, A, B, moduli=0):
    m, k = A.shape
    n = B.shape[1]
    dA, dB = dev(A), dev(B)
    dC = full((m, n), np.float64, float("nan"))
    ctx().call(fn, m, n, k, p(dA), k, p(dB), n, p(dC), n, moduli)
    return host(dC)


def graded(rng, shape, sigma):
    u = rng.uniform(-1, 1, shape)
    e = np.rint(sigma * rng.standard_normal(shape)) if sigma > 0 else 0
    return np.ldexp(u, np.asarray(e, dtype=np.int64))


@pytest.mark.parametrize("fn", FNS)
def test_integer_inputs_bit_exact(fn):
    rng = np.random.default_rng(50)
    for (m, n, k) in [(1, 1, 1), (7, 5, 3), (300, 200, 129), (256, 256, 256), (513, 257, 1000)]:
        A = rng.integers(-1000, 1001, (m, k)).astype(np.float64)
        B = rng.integers(-1000, 1001, (k, n)).astype(np.float64)
        assert np.array_equal(gpu_fp64acc(fn, A, B), A @ B), (fn, m, n, k)


@pytest.mark.parametrize("fn", FNS)
def test_uniform_relative_frobenius_1e14(fn):
    rng = np.random.default_rng(51)
    for (m, n, k) in [(512, 512, 512), (1024, 768, 2048), (300, 700, 4097)]:
        A = rng.uniform(-1, 1, (m, k)); B = rng.uniform(-1, 1, (k, n))
        Cg = gpu_fp64acc(fn, A, B)
        rows = rng.integers(0, m, 3000); cols = rng.integers(0, n, 3000)
        hi, lo = oracle.gemm_dd_entries(A, B, rows, cols)
        err = np.sqrt(np.sum((Cg[rows, cols] - hi - lo) ** 2) / np.sum(hi ** 2))
        assert err <= 1e-14, (fn, m, n, k, err)


@pytest.mark.parametrize("fn", FNS)
@pytest.mark.parametrize("sigma", [0, 4, 10])
def test_graded_exponents_within_bound(fn, sigma):
    rng = np.random.default_rng(52 + sigma)
    m, n, k = 384, 320, 1536
    A = graded(rng, (m, k), sigma); B = graded(rng, (k, n), sigma)
    Cg = gpu_fp64acc(fn, A, B)
    P = last_plan()[1]          # operand bits of the automatic choice
    assert P >= 40
    rows = rng.integers(0, m, 2000); cols = rng.integers(0, n, 2000)
    hi, lo = oracle.gemm_dd_entries(A, B, rows, cols)
    ra = np.abs(A).max(axis=1)[rows]; cb = np.abs(B).max(axis=0)[cols]
    bound = k * 2.0 ** -P * ra * cb + np.abs(hi) * 2.0 ** -52
    assert np.all(np.abs(Cg[rows, cols] - hi) <= bound + np.abs(lo)), (fn, sigma)


@pytest.mark.parametrize("fn", FNS)
def test_zero_rows_and_columns(fn):
    rng = np.random.default_rng(53)
    A = rng.uniform(-1, 1, (70, 90)); B = rng.uniform(-1, 1, (90, 40))
    A[5] = 0; B[:, 7] = 0
    Cg = gpu_fp64acc(fn, A, B)
    assert np.all(Cg[5] == 0) and np.all(Cg[:, 7] == 0)
    assert np.abs(Cg - A @ B).max() <= 1e-14 * np.abs(A @ B).max() * 10


def test_more_moduli_tighten_error_for_wide_exponents():
    rng = np.random.default_rng(54)
    m, n, k = 256, 256, 512
    A = graded(rng, (m, k), 10); B = graded(rng, (k, n), 10)
    rows = rng.integers(0, m, 2000); cols = rng.integers(0, n, 2000)
    hi, lo = oracle.gemm_dd_entries(A, B, rows, cols)
    errs = []
    for nm in (13, 20):
        Cg = gpu_fp64acc("bzg_dgemm_fp64acc", A, B, moduli=nm)
        errs.append(np.sqrt(np.sum((Cg[rows, cols] - hi) ** 2) / np.sum(hi ** 2)))
    assert errs[1] <= errs[0]
    assert errs[1] <= 1e-14, errs


@pytest.mark.parametrize("fn", FNS)
@pytest.mark.parametrize("sigma", [0, 4, 10])
def test_automatic_moduli_meet_1e14_per_sigma(fn, sigma):
    # entries u 2^e, e = round(N(0, sigma)): the automatic count (operand statistics -> error
    # estimate <= 8e-15) keeps the relative Frobenius error below 1e-14 for every sigma, and
    # the estimate is an upper bound of the measured error within a factor of a few
    rng = np.random.default_rng(60 + sigma)
    m, n, k = 512, 384, 4096
    A = graded(rng, (m, k), sigma); B = graded(rng, (k, n), sigma)
    Cg = gpu_fp64acc(fn, A, B)
    nm, P, est = last_plan()
    # the full Frobenius norm: for wide exponent spreads a few entries carry most of ||C||,
    # and a sampled ratio would weigh the small entries' absolute errors out of proportion
    rows, cols = (x.ravel() for x in np.meshgrid(np.arange(m), np.arange(n), indexing="ij"))
    hi, lo = oracle.gemm_dd_entries(A, B, rows, cols)
    err = np.sqrt(np.sum((Cg.ravel() - hi - lo) ** 2) / np.sum(hi ** 2))
    assert err <= 1e-14, (fn, sigma, nm, P, est, err)
    assert 0 < est <= 8e-15 and 2 <= nm <= 24
    if sigma > 0:  # single entries dominate their rows: the stricter target applies
        assert est <= 8e-15 / 2.5


def last_plan():
    import ctypes as C
    nm, bits, est = C.c_int32(), C.c_int32(), C.c_double()
    ctx().call("bzg_fp64acc_last_plan", C.byref(nm), C.byref(bits), C.byref(est))
    return nm.value, bits.value, est.value


def _dd_rel_frob(Cg, A, B, rows, cols):
    hi, lo = oracle.gemm_dd_entries(A, B, rows, cols)
    return float(np.sqrt(np.sum((Cg[rows, cols] - hi - lo) ** 2) / np.sum(hi ** 2)))


def test_rectangular_16384x1024x16384_int8():
    # the rectangular case of the FP64-accurate GEMM workload (M x N x K = 16384 x 1024 x 16384,
    # uniform[-1,1)): K = 16384 is the longest exact accumulation measured; sampled entries
    # against the double-double product
    rng = np.random.default_rng(70)
    m, n, k = 16384, 1024, 16384
    print(f"fp64acc_i8 {m}x{n}x{k}", flush=True)
    A = rng.uniform(-1, 1, (m, k)); B = rng.uniform(-1, 1, (k, n))
    Cg = gpu_fp64acc("bzg_dgemm_fp64acc_i8", A, B)
    nm, P, est = last_plan()
    rows = rng.integers(0, m, 1500); cols = rng.integers(0, n, 1500)
    err = _dd_rel_frob(Cg, A, B, rows, cols)
    assert err <= 1e-14, (nm, P, est, err)


@pytest.mark.parametrize("sigma", [0, 4, 10])
def test_k16384_full_frobenius_per_sigma(sigma):
    # K = 16384 with graded exponents u 2^e, e = round(N(0, sigma)): the automatic moduli count
    # keeps the full relative Frobenius error (every entry of a 128 x 128 product) below 1e-14
    rng = np.random.default_rng(80 + sigma)
    m, n, k = 128, 128, 16384
    A = graded(rng, (m, k), sigma); B = graded(rng, (k, n), sigma)
    for fn in FNS:
        print(f"{fn} {m}x{n}x{k} sigma={sigma}", flush=True)
        Cg = gpu_fp64acc(fn, A, B)
        nm, P, est = last_plan()
        rows, cols = (x.ravel() for x in np.meshgrid(np.arange(m), np.arange(n), indexing="ij"))
        err = _dd_rel_frob(Cg, A, B, rows, cols)
        assert err <= 1e-14, (fn, sigma, nm, P, est, err)


def _native(A, B):
    m, k = A.shape
    n = B.shape[1]
    dA, dB = dev(A), dev(B)
    dC = full((m, n), np.float64, 0.0)
    ctx().call("bzg_dgemm", m, n, k, 1.0, p(dA), k, p(dB), n, 0.0, p(dC), n)
    return host(dC)


@pytest.mark.parametrize("fn", FNS)
@pytest.mark.parametrize("moduli", [0, 16])
def test_nan_and_inf_operands_match_native_fp64(fn, moduli):
    # NaN and +-Inf entries propagate as in the plain FP64 GEMM (and the reference's IEEE
    # arithmetic): every entry in an affected row of A or column of B is NaN or +-Inf, with
    # the same pattern and signs; the other entries keep the split method's accuracy.
    # moduli = 0: the statistics pass sees the non-finite operand and the call runs the plain
    # FP64 GEMM; an explicit count finds the rows / columns on the device and recomputes them.
    rng = np.random.default_rng(90)
    m, n, k = 200, 150, 300
    A = rng.uniform(-1, 1, (m, k)); B = rng.uniform(-1, 1, (k, n))
    A[3, 5] = np.nan
    A[10, 7] = np.inf
    A[11, 7] = -np.inf
    A[12, 8] = np.inf; A[12, 9] = -np.inf      # +Inf and -Inf in one row: NaN wherever both survive
    B[2, 4] = np.nan
    B[20, 6] = -np.inf
    B[7, 30] = 0.0                             # Inf * 0 = NaN in rows 10 and 11, column 30
    Cg = gpu_fp64acc(fn, A, B, moduli)
    Cn = _native(A, B)
    assert np.array_equal(np.isnan(Cg), np.isnan(Cn))
    inf = np.isinf(Cn)
    assert np.array_equal(np.isinf(Cg), inf) and np.array_equal(Cg[inf], Cn[inf])
    bad_r = np.zeros(m, bool); bad_r[[3, 10, 11, 12]] = True
    bad_c = np.zeros(n, bool); bad_c[[4, 6]] = True
    assert np.all(~np.isfinite(Cg[bad_r])) and np.all(~np.isfinite(Cg[:, bad_c]))
    ok = ~bad_r[:, None] & ~bad_c[None, :]
    assert np.all(np.isfinite(Cg[ok]))
    ref = A[~bad_r] @ B[:, ~bad_c]
    assert np.abs(Cg[np.ix_(~bad_r, ~bad_c)] - ref).max() <= 1e-13 * np.abs(ref).max()
    # a clean product on the same context afterwards is unaffected
    A2 = rng.uniform(-1, 1, (64, 96)); B2 = rng.uniform(-1, 1, (96, 80))
    assert np.abs(gpu_fp64acc(fn, A2, B2, moduli) - A2 @ B2).max() <= 1e-14 * 10
