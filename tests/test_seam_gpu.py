"""GPU: the BLAS/LAPACK seam library (blaze_amd/lib/libblaze_seam.so: cblas_dgemm, cblas_dgemv, the
complex cblas_zgemm / cblas_cgemm / cblas_zgemv / cblas_cgemv,
dgetrf_, dgetrs_, dgesv_ on host memory, computed by the device library) against numpy and
LAPACK (scipy) for every storage order, transpose, leading dimension and stride."""
import ctypes as C

import pytest

from tests import seam_cases as sc
from tests.gpu_util import have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


@pytest.fixture(scope="module")
def seam():
    import blaze_amd as bz
    path = bz.LIB_PATH.parent / "libblaze_seam.so"
    assert path.exists(), "libblaze_seam.so not built"
    return sc.bind(C.CDLL(str(path)))


def test_cblas_dgemm(seam):
    print("seam dgemm", flush=True)
    sc.check_gemm(seam, 130, 70, 90, seed=11, tol=1e-13)


def test_cblas_dgemv(seam):
    sc.check_gemv(seam, 300, 170, seed=12, tol=1e-13)


def test_lapack(seam):
    sc.check_lapack(seam, 700, 5, seed=13, tol=1e-10)


def test_cblas_complex_gemm(seam):
    import numpy as np
    sc.check_cgemm(seam, 130, 70, 90, seed=14, tol=1e-13)
    sc.check_cgemm(seam, 66, 40, 50, seed=15, tol=2e-5, dtype=np.complex64)


def test_cblas_complex_gemv(seam):
    import numpy as np
    sc.check_cgemv(seam, 300, 170, seed=16, tol=1e-13)
    sc.check_cgemv(seam, 120, 90, seed=17, tol=2e-5, dtype=np.complex64)
