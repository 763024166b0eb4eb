"""Checks of the BLAS/LAPACK seam (blaze_amd/csrc/seam.cpp) shared by the CPU suite (seam over
a host stand-in of the ABI) and the GPU suite (seam over the device library): every storage
order, transpose, padded leading dimension and vector stride against numpy / LAPACK."""
import ctypes as C

import numpy as np

ROW, COL = 101, 102
NT, TR, CT = 111, 112, 113
ci = C.c_int
cd = C.c_double
vp = C.c_void_p


def bind(lib):
    lib.cblas_dgemm.argtypes = [ci, ci, ci, ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.cblas_dgemm.restype = None
    lib.cblas_dgemv.argtypes = [ci, ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.cblas_dgemv.restype = None
    for f in (lib.dgetrf_, lib.dgesv_, lib.dgetrs_):
        f.restype = None
    for f in (lib.cblas_zgemm, lib.cblas_cgemm):
        f.argtypes = [ci, ci, ci, ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci]
        f.restype = None
    for f in (lib.cblas_zgemv, lib.cblas_cgemv):
        f.argtypes = [ci, ci, ci, ci, vp, vp, ci, vp, ci, vp, vp, ci]
        f.restype = None
    return lib


def _p(a):
    return a.ctypes.data_as(vp)


def store(X, order, ld):
    """Flat storage of the math matrix X (r x c) in the given order with leading dimension ld."""
    r, c = X.shape
    if order == ROW:
        S = np.full((r, ld), 7.25)
        S[:, :c] = X
    else:
        S = np.full((c, ld), 7.25)
        S[:, :r] = X.T
    return S.ravel().copy()


def load(S, order, r, c, ld):
    S = S.reshape(-1, ld)
    return S[:r, :c].copy() if order == ROW else S[:c, :r].T.copy()


def check_gemm(lib, m, n, k, seed, tol):
    rng = np.random.default_rng(seed)
    for order in (ROW, COL):
        for ta in (NT, TR):
            for tb in (NT, TR):
                for beta in (0.0, -0.5):
                    A = rng.uniform(-1, 1, (k, m) if ta == TR else (m, k))
                    B = rng.uniform(-1, 1, (n, k) if tb == TR else (k, n))
                    C0 = rng.uniform(-1, 1, (m, n))
                    lda, ldb, ldc = A.shape[1 if order == ROW else 0] + 3, B.shape[1 if order == ROW else 0] + 2, \
                        (n if order == ROW else m) + 1
                    sA, sB, sC = store(A, order, lda), store(B, order, ldb), store(C0, order, ldc)
                    if beta == 0.0:
                        sC[:] = np.nan  # must not be read
                        keep = sC.copy()
                    lib.cblas_dgemm(order, ta, tb, m, n, k, 1.5, _p(sA), lda, _p(sB), ldb, beta, _p(sC), ldc)
                    opA = A.T if ta == TR else A
                    opB = B.T if tb == TR else B
                    want = 1.5 * opA @ opB + (beta * C0 if beta != 0.0 else 0.0)
                    got = load(sC, order, m, n, ldc)
                    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()), (order, ta, tb, beta)
                    if beta == 0.0:  # padding outside the m x n block is untouched
                        S = sC.reshape(-1, ldc)
                        pad = S[:, (n if order == ROW else m):]
                        assert np.isnan(pad).all()


def _strided(v, inc):
    n = len(v)
    buf = np.full(max(1, 1 + (n - 1) * abs(inc)), 3.5)
    for i in range(n):
        buf[(n - 1 - i if inc < 0 else i) * abs(inc)] = v[i]
    return buf


def _unstride(buf, n, inc):
    return np.array([buf[(n - 1 - i if inc < 0 else i) * abs(inc)] for i in range(n)])


def check_gemv(lib, m, n, seed, tol):
    rng = np.random.default_rng(seed)
    for order in (ROW, COL):
        for t in (NT, TR):
            for incx, incy in ((1, 1), (2, -1), (-3, 2)):
                A = rng.uniform(-1, 1, (m, n))
                lda = (n if order == ROW else m) + 2
                sA = store(A, order, lda)
                lx, ly = (m, n) if t == TR else (n, m)
                x = rng.uniform(-1, 1, lx)
                y0 = rng.uniform(-1, 1, ly)
                bx, by = _strided(x, incx), _strided(y0, incy)
                lib.cblas_dgemv(order, t, m, n, 0.75, _p(sA), lda, _p(bx), incx, 2.0, _p(by), incy)
                want = 0.75 * ((A.T if t == TR else A) @ x) + 2.0 * y0
                got = _unstride(by, ly, incy)
                assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()), (order, t, incx, incy)


def check_lapack(lib, n, nrhs, seed, tol):
    import scipy.linalg
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (n, n))
    lda, ldb = n + 3, n + 1
    sA = store(A, COL, lda)
    ipiv = np.zeros(n, np.int32)
    info = C.c_int(99)
    lib.dgetrf_(C.byref(ci(n)), C.byref(ci(n)), _p(sA), C.byref(ci(lda)), _p(ipiv), C.byref(info))
    lu_s, piv_s = scipy.linalg.lu_factor(A)
    assert info.value == 0
    assert np.array_equal(ipiv - 1, piv_s)
    LU = load(sA, COL, n, n, lda)
    assert np.abs(LU - lu_s).max() <= tol * np.abs(lu_s).max()
    # dgetrs_ ('N') from those factors
    B = rng.uniform(-1, 1, (n, nrhs))
    sB = store(B, COL, ldb)
    lib.dgetrs_(b"N", C.byref(ci(n)), C.byref(ci(nrhs)), _p(sA), C.byref(ci(lda)), _p(ipiv), _p(sB),
                C.byref(ci(ldb)), C.byref(info), C.c_size_t(1))
    X = load(sB, COL, n, nrhs, ldb)
    assert info.value == 0
    assert np.abs(A @ X - B).max() <= tol * 100 * max(1.0, np.abs(X).max())
    # dgesv_
    sA2, sB2 = store(A, COL, lda), store(B, COL, ldb)
    ipiv2 = np.zeros(n, np.int32)
    lib.dgesv_(C.byref(ci(n)), C.byref(ci(nrhs)), _p(sA2), C.byref(ci(lda)), _p(ipiv2), _p(sB2),
               C.byref(ci(ldb)), C.byref(info))
    assert info.value == 0 and np.array_equal(ipiv2, ipiv)
    assert np.abs(load(sB2, COL, n, nrhs, ldb) - X).max() <= tol * max(1.0, np.abs(X).max())
    # singular: info = first zero pivot, B unchanged
    As = A.copy()
    As[:, 2] = 0.0
    sA3, sB3 = store(As, COL, lda), store(B, COL, ldb)
    keep = sB3.copy()
    lib.dgesv_(C.byref(ci(n)), C.byref(ci(nrhs)), _p(sA3), C.byref(ci(lda)), _p(ipiv2), _p(sB3),
               C.byref(ci(ldb)), C.byref(info))
    assert info.value == 3 and np.array_equal(sB3, keep)
    # unsupported forms report an argument error instead of computing
    lib.dgetrf_(C.byref(ci(n)), C.byref(ci(n - 1)), _p(sA3), C.byref(ci(lda)), _p(ipiv2), C.byref(info))
    assert info.value == -2
    lib.dgetrs_(b"T", C.byref(ci(n)), C.byref(ci(nrhs)), _p(sA), C.byref(ci(lda)), _p(ipiv), _p(sB),
                C.byref(ci(ldb)), C.byref(info), C.c_size_t(1))
    assert info.value == -1


def _op(X, t):
    return X if t == NT else (X.T if t == TR else X.conj().T)


def cstore(X, order, ld, dtype):
    """Flat complex storage of X (r x c) in the given order with leading dimension ld."""
    r, c = X.shape
    S = np.full((r, ld) if order == ROW else (c, ld), 7.25 - 1.5j, dtype)
    if order == ROW:
        S[:, :c] = X
    else:
        S[:, :r] = X.T
    return S.ravel().copy()


def check_cgemm(lib, m, n, k, seed, tol, dtype=np.complex128):
    fn = lib.cblas_zgemm if dtype == np.complex128 else lib.cblas_cgemm
    rdt = np.float64 if dtype == np.complex128 else np.float32
    rng = np.random.default_rng(seed)
    z = lambda *sh: (rng.uniform(-1, 1, sh) + 1j * rng.uniform(-1, 1, sh)).astype(dtype)  # noqa: E731
    alpha = np.array([1.5, -0.5], rdt)
    for order in (ROW, COL):
        for ta in (NT, TR, CT):
            for tb in (NT, TR, CT):
                for beta in (0.0, -0.5 + 0.25j):
                    A = z(*((k, m) if ta != NT else (m, k)))
                    B = z(*((n, k) if tb != NT else (k, n)))
                    C0 = z(m, n)
                    lda = A.shape[1 if order == ROW else 0] + 3
                    ldb = B.shape[1 if order == ROW else 0] + 2
                    ldc = (n if order == ROW else m) + 1
                    sA, sB, sC = cstore(A, order, lda, dtype), cstore(B, order, ldb, dtype), cstore(C0, order, ldc, dtype)
                    if beta == 0.0:
                        sC[:] = np.nan
                    be = np.array([beta.real, beta.imag] if beta != 0.0 else [0.0, 0.0], rdt)
                    fn(order, ta, tb, m, n, k, _p(alpha), _p(sA), lda, _p(sB), ldb, _p(be), _p(sC), ldc)
                    want = (1.5 - 0.5j) * (_op(A, ta).astype(np.complex128) @ _op(B, tb)) + (
                        beta * C0 if beta != 0.0 else 0.0)
                    got = load(sC, order, m, n, ldc)
                    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()), (order, ta, tb, beta)
                    if beta == 0.0:
                        pad = sC.reshape(-1, ldc)[:, (n if order == ROW else m):]
                        assert np.isnan(pad).all()


def check_cgemv(lib, m, n, seed, tol, dtype=np.complex128):
    fn = lib.cblas_zgemv if dtype == np.complex128 else lib.cblas_cgemv
    rdt = np.float64 if dtype == np.complex128 else np.float32
    rng = np.random.default_rng(seed)
    z = lambda *sh: (rng.uniform(-1, 1, sh) + 1j * rng.uniform(-1, 1, sh)).astype(dtype)  # noqa: E731
    alpha = np.array([0.75, 0.5], rdt)
    for order in (ROW, COL):
        for t in (NT, TR, CT):
            for incx, incy, beta in ((1, 1, 2.0 - 1.0j), (2, -1, 0.0), (-3, 2, 0.5j)):
                A = z(m, n)
                lda = (n if order == ROW else m) + 2
                sA = cstore(A, order, lda, dtype)
                lx, ly = (n, m) if t == NT else (m, n)
                x, y0 = z(lx), z(ly)
                bx = np.full(max(1, 1 + (lx - 1) * abs(incx)), 3.5 + 0.5j, dtype)
                by = np.full(max(1, 1 + (ly - 1) * abs(incy)), 3.5 + 0.5j, dtype)
                for i in range(lx):
                    bx[(lx - 1 - i if incx < 0 else i) * abs(incx)] = x[i]
                for i in range(ly):
                    by[(ly - 1 - i if incy < 0 else i) * abs(incy)] = y0[i]
                be = np.array([beta.real, beta.imag] if beta != 0.0 else [0.0, 0.0], rdt)
                fn(order, t, m, n, _p(alpha), _p(sA), lda, _p(bx), incx, _p(be), _p(by), incy)
                want = (0.75 + 0.5j) * (_op(A, t).astype(np.complex128) @ x) + (beta * y0 if beta != 0.0 else 0.0)
                got = np.array([by[(ly - 1 - i if incy < 0 else i) * abs(incy)] for i in range(ly)])
                assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()), (order, t, incx, incy)
