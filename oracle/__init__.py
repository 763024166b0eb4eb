"""CPU ORACLE -- test infrastructure only.

numpy-facing wrappers around oracle/liboracle.so (the C restatement of the reference
algorithms, oracle.c) and oracle/_ref/libblazeref.so (the reference's own kernels compiled
from its headers, when that build exists).  Only tests/, the smoke check and the benchmark's
CPU-baseline leg import this package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libblazeref.so"

i64, i32, f64, vp = C.c_int64, C.c_int32, C.c_double, C.c_void_p
_orc = None
_ref = None


def build() -> None:
    """Compile liboracle.so and, when the reference tree is present, _ref/libblazeref.so."""
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            build()
        L = C.CDLL(str(ORACLE_SO))
        L.orc_getrf.restype = i32
        L.orc_sgetrf.restype = i32
        L.orc_gesv.restype = i32
        L.orc_solve_closed.restype = i32
        L.orc_dsgesv.restype = i32
        L.orc_backward_error.restype = f64
        _orc = L
    return _orc


def have_ref() -> bool:
    return REF_SO.exists()


REF_LAPACK_SO = HERE / "_ref" / "libblazeref_lapack.so"
_ref_lapack = None


def have_ref_lapack() -> bool:
    return REF_LAPACK_SO.exists()


def ref_lapack() -> C.CDLL:
    """The reference's solve() (LAPACK-backed for N >= 7) compiled against the image's LAPACK;
    LAPACK runs sequentially, the systems are spread over OpenMP threads."""
    global _ref_lapack
    if _ref_lapack is None:
        if not REF_LAPACK_SO.exists():
            raise FileNotFoundError(f"{REF_LAPACK_SO} not built (needs the reference tree and a LAPACK)")
        import os
        os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
        os.environ.setdefault("MKL_NUM_THREADS", "1")
        L = C.CDLL(str(REF_LAPACK_SO))
        L.ref_time_dsolve_batch.restype = f64
        L.ref_time_dsolve_batch.argtypes = [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(i64)]
        _ref_lapack = L
    return _ref_lapack


def ref_solve_batch(n, nrhs, a_off, b_off, A, B, threads=0, reps=0):
    """(X, seconds, fails): the reference's solve(A_i, X_i, B_i) for every system of the pools."""
    X = np.zeros_like(B)
    fails = i64(0)
    arrs = [np.ascontiguousarray(n, np.int32), np.ascontiguousarray(nrhs, np.int32),
            np.ascontiguousarray(a_off, np.int64), np.ascontiguousarray(b_off, np.int64)]
    t = ref_lapack().ref_time_dsolve_batch(len(arrs[0]), *[_p(a) for a in arrs], _p(A), _p(B), _p(X),
                                           int(threads), int(reps), C.byref(fails))
    return X, t, fails.value


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise FileNotFoundError(f"{REF_SO} not built (needs the reference tree)")
        L = C.CDLL(str(REF_SO))
        L.ref_dsolve_small.restype = i32
        L.ref_get_threads.restype = i32
        for n in ("ref_time_dgemm", "ref_time_dcsrmv", "ref_time_dgemv", "ref_time_zgemm"):
            getattr(L, n).restype = f64
        _ref = L
    return _ref


def _p(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(vp)


# ---- oracle (restatement) ---------------------------------------------------------------
def op(M: np.ndarray, code: int) -> np.ndarray:
    """op(M) for the BLAS transpose codes 0 = N, 1 = T, 2 = C (conjugate transpose)."""
    return M if code == 0 else (M.T if code == 1 else M.conj().T)


def gemm_op(opa: int, opb: int, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """C = op(A) op(B) for complex operands (DMatDMatMultExpr with complex elements; the
    textbook product, evaluated in complex128)."""
    return (op(A, opa).astype(np.complex128) @ op(B, opb).astype(np.complex128))


def gemm_op_bound(opa: int, opb: int, A: np.ndarray, B: np.ndarray, eps: float) -> np.ndarray:
    """Entry-wise error bound of a computed complex product of inner dimension k, in any
    summation order: |c - c_exact| <= (k + 2) eps (|Re A| + |Im A|)(|Re B| + |Im B|), plus a
    few units of the smallest normal (the real 2 x 2 form rounds each of the 2k real products
    once and adds them in some order)."""
    a = op(np.abs(A.real) + np.abs(A.imag), opa)
    b = op(np.abs(B.real) + np.abs(B.imag), opb)
    k = a.shape[1]
    return 2.0 * (k + 2) * eps * (a @ b) + 1e-300


def gemm(A: np.ndarray, B: np.ndarray, fused: bool = True) -> np.ndarray:
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    if A.dtype == np.float64:
        Cm = np.empty((m, n), np.float64)
        orc().orc_gemm_d(i64(m), i64(n), i64(k), _p(A), i64(k), _p(B), i64(n), _p(Cm), i64(n), C.c_int(int(fused)))
    elif A.dtype == np.float32:
        Cm = np.empty((m, n), np.float32)
        orc().orc_gemm_s(i64(m), i64(n), i64(k), _p(A), i64(k), _p(B), i64(n), _p(Cm), i64(n))
    elif A.dtype == np.int32:
        Cm = np.empty((m, n), np.int32)
        orc().orc_gemm_i32(i64(m), i64(n), i64(k), _p(A), i64(k), _p(B), i64(n), _p(Cm), i64(n))
    else:
        raise TypeError(A.dtype)
    return Cm


def gemm_dd_entries(A: np.ndarray, B: np.ndarray, rows: np.ndarray, cols: np.ndarray):
    """Double-double C[rows, cols] of A @ B: returns (hi, lo)."""
    k = A.shape[1]
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    hi = np.empty(len(rows)); lo = np.empty(len(rows))
    orc().orc_gemm_dd_entries(i64(k), _p(A), i64(A.shape[1]), _p(B), i64(B.shape[1]), i64(len(rows)),
                              _p(rows), _p(cols), _p(hi), _p(lo))
    return hi, lo


def gemv(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    m, n = A.shape
    fn = {np.dtype(np.float64): "orc_gemv_d", np.dtype(np.float32): "orc_gemv_s",
          np.dtype(np.int32): "orc_gemv_i32"}[A.dtype]
    y = np.empty(m, A.dtype)
    getattr(orc(), fn)(i64(m), i64(n), _p(A), i64(n), _p(x), _p(y))
    return y


def csrmv(rowptr, col, val, x) -> np.ndarray:
    n = len(rowptr) - 1
    fn = {np.dtype(np.float64): "orc_csrmv_d", np.dtype(np.float32): "orc_csrmv_s",
          np.dtype(np.int32): "orc_csrmv_i32"}[val.dtype]
    y = np.empty(n, val.dtype)
    getattr(orc(), fn)(i64(n), _p(rowptr), _p(col), _p(val), _p(x), _p(y))
    return y


def tcsrmv(rowptr, col, val, x, n_cols: int) -> np.ndarray:
    """y = trans(S) * x for a row-major CSR S (oracle.c orc_tcsrmv_*: TSMatDVecMultExpr's loop)."""
    n = len(rowptr) - 1
    fn = {np.dtype(np.float64): "orc_tcsrmv_d", np.dtype(np.float32): "orc_tcsrmv_s"}[val.dtype]
    y = np.empty(n_cols, val.dtype)
    getattr(orc(), fn)(i64(n), i64(n_cols), _p(rowptr), _p(col), _p(val), _p(x), _p(y))
    return y


def getrf(A: np.ndarray):
    """LU with partial pivoting (dgetf2 rule) of a copy of A: returns (LU, ipiv 1-based, info)."""
    LU = np.array(A, np.float64, order="C", copy=True)
    n = LU.shape[0]
    ip = np.zeros(n, np.int32)
    info = orc().orc_getrf(i32(n), _p(LU), i32(n), _p(ip))
    return LU, ip, info


def cg(rowptr, col, val, b, x0, iterations: int):
    """The CG loop of blazemark/src/blaze/CG.cpp:114-135 (products by csrmv, the reference's
    sequential row sums; dot products by numpy).  Returns (x, r.r after the last iteration)."""
    x = np.array(x0, np.float64, copy=True)
    r = csrmv(rowptr, col, val, x) - b
    delta = float(r @ r)
    d = -r
    for _ in range(iterations):
        h = csrmv(rowptr, col, val, d)
        alpha = delta / float(d @ h)
        x += alpha * d
        r += alpha * h
        beta = float(r @ r)
        d = (beta / delta) * d - r
        delta = beta
    return x, delta


def laplace2d(N: int):
    """The 5-point matrix of blazemark/src/blaze/CG.cpp:86-103 (row-major CSR, diagonal 4,
    neighbours -1, rows in the order top, left, centre, right, bottom)."""
    rp, ci, va = [0], [], []
    for i in range(N):
        for j in range(N):
            if i > 0: ci.append((i - 1) * N + j); va.append(-1.0)
            if j > 0: ci.append(i * N + j - 1); va.append(-1.0)
            ci.append(i * N + j); va.append(4.0)
            if j < N - 1: ci.append(i * N + j + 1); va.append(-1.0)
            if i < N - 1: ci.append((i + 1) * N + j); va.append(-1.0)
            rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va)


def getrs(LU: np.ndarray, ipiv: np.ndarray, B: np.ndarray) -> np.ndarray:
    """dgetrs ('N'): X from the factors and 1-based pivots of getrf (B is n x nrhs)."""
    LU = np.ascontiguousarray(LU, np.float64)
    ip = np.ascontiguousarray(ipiv, np.int32)
    X = np.array(B, np.float64, order="C", copy=True)
    n, nrhs = X.shape
    orc().orc_getrs(i32(n), i32(nrhs), _p(LU), i32(n), _p(ip), _p(X), i32(nrhs))
    return X


def gesv(A: np.ndarray, B: np.ndarray):
    """(X, LU, ipiv, info) of A X = B (B is n x nrhs)."""
    LU = np.array(A, np.float64, order="C", copy=True)
    X = np.array(B, np.float64, order="C", copy=True)
    n = LU.shape[0]
    nrhs = X.shape[1]
    ip = np.zeros(n, np.int32)
    info = orc().orc_gesv(i32(n), i32(nrhs), _p(LU), i32(n), _p(ip), _p(X), i32(nrhs))
    return X, LU, ip, info


def solve_closed(A: np.ndarray, b: np.ndarray):
    n = A.shape[0]
    A = np.ascontiguousarray(A, np.float64); b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    rc = orc().orc_solve_closed(i32(n), _p(A), _p(b), _p(x))
    return x, rc


def dsgesv(A: np.ndarray, B: np.ndarray, eps: float = 0.0, max_iter: int = 30):
    """Mixed-precision solve: returns (X, iter, info)."""
    A = np.ascontiguousarray(A, np.float64); B = np.ascontiguousarray(B, np.float64)
    n, nrhs = B.shape
    X = np.zeros_like(B)
    info = C.c_int(0)
    it = orc().orc_dsgesv(i32(n), i32(nrhs), _p(A), _p(B), _p(X), f64(eps), i32(max_iter), C.byref(info))
    return X, it, info.value


def backward_error(A, B, X) -> float:
    A = np.ascontiguousarray(A, np.float64); B = np.ascontiguousarray(B, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    n, nrhs = B.shape
    return orc().orc_backward_error(i32(n), i32(nrhs), _p(A), _p(B), _p(X))


# ---- compiled reference ----------------------------------------------------------------
def ref_gemm(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    m, k = A.shape
    n = B.shape[1]
    fn = {np.dtype(np.float64): "ref_dgemm", np.dtype(np.float32): "ref_sgemm",
          np.dtype(np.int32): "ref_igemm"}[A.dtype]
    Cm = np.empty((m, n), A.dtype)
    getattr(ref(), fn)(i64(m), i64(n), i64(k), _p(A), _p(B), _p(Cm))
    return Cm


def ref_gemv(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    m, n = A.shape
    fn = {np.dtype(np.float64): "ref_dgemv", np.dtype(np.float32): "ref_sgemv",
          np.dtype(np.int32): "ref_igemv"}[A.dtype]
    y = np.empty(m, A.dtype)
    getattr(ref(), fn)(i64(m), i64(n), _p(A), _p(x), _p(y))
    return y


def ref_tgemv(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    """y = trans(A) * x by the reference's own TDMatDVecMultExpr (compiled reference)."""
    m, n = A.shape
    fn = {np.dtype(np.float64): "ref_dtgemv", np.dtype(np.float32): "ref_stgemv",
          np.dtype(np.int32): "ref_itgemv"}[A.dtype]
    y = np.empty(n, A.dtype)
    getattr(ref(), fn)(i64(m), i64(n), _p(A), _p(x), _p(y))
    return y


def ref_tcsrmv(rowptr, col, val, x, n_cols: int) -> np.ndarray:
    """trans(S) * x by the reference's own TSMatDVecMultExpr (compiled reference)."""
    n = len(rowptr) - 1
    fn = {np.dtype(np.float64): "ref_dtcsrmv", np.dtype(np.float32): "ref_stcsrmv"}[val.dtype]
    y = np.empty(n_cols, val.dtype)
    getattr(ref(), fn)(i64(n), i64(n_cols), _p(rowptr), _p(col), _p(val), _p(x), _p(y))
    return y


def ref_csrmv(rowptr, col, val, x, n_cols: int) -> np.ndarray:
    n = len(rowptr) - 1
    fn = {np.dtype(np.float64): "ref_dcsrmv", np.dtype(np.float32): "ref_scsrmv",
          np.dtype(np.int32): "ref_icsrmv"}[val.dtype]
    y = np.empty(n, val.dtype)
    getattr(ref(), fn)(i64(n), i64(n_cols), _p(rowptr), _p(col), _p(val), _p(x), _p(y))
    return y


def ref_gemm_op(opa: int, opb: int, A: np.ndarray, B: np.ndarray, m: int, n: int, k: int) -> np.ndarray:
    """C = op(A) * op(B) by the reference's own DMatDMatMultExpr for complex elements (op 0 =
    none, 1 = trans(), 2 = ctrans(); A stored m x k or k x m, B stored k x n or n x k)."""
    fn = {np.dtype(np.complex128): "ref_zgemm_op", np.dtype(np.complex64): "ref_cgemm_op"}[A.dtype]
    Cm = np.empty((m, n), A.dtype)
    getattr(ref(), fn)(C.c_int(opa), C.c_int(opb), i64(m), i64(n), i64(k), _p(np.ascontiguousarray(A)),
                       _p(np.ascontiguousarray(B)), _p(Cm))
    return Cm


def ref_gemv_op(op: int, A: np.ndarray, x: np.ndarray) -> np.ndarray:
    """y = op(A) * x by the reference's DMatDVecMultExpr for complex elements."""
    m, n = A.shape
    fn = {np.dtype(np.complex128): "ref_zgemv_op", np.dtype(np.complex64): "ref_cgemv_op"}[A.dtype]
    y = np.empty(n if op else m, A.dtype)
    getattr(ref(), fn)(C.c_int(op), i64(m), i64(n), _p(np.ascontiguousarray(A)), _p(np.ascontiguousarray(x)),
                       _p(y))
    return y


def ref_zsolve(A: np.ndarray, B: np.ndarray):
    """(X, threw): the reference's solve(A, X, B) for one complex<double> system (zgesv_ for
    n >= 7, through the image's LAPACK)."""
    n, nrhs = B.shape
    A = np.ascontiguousarray(A, np.complex128); B = np.ascontiguousarray(B, np.complex128)
    X = np.zeros((n, nrhs), np.complex128)
    L = ref_lapack()
    L.ref_zsolve.restype = i32
    rc = L.ref_zsolve(i64(n), i64(nrhs), _p(A), _p(B), _p(X))
    return X, rc


def ref_solve_small(A: np.ndarray, b: np.ndarray):
    n = A.shape[0]
    A = np.ascontiguousarray(A, np.float64); b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    rc = ref().ref_dsolve_small(i32(n), _p(A), _p(b), _p(x))
    return x, rc
