"""CPU: the BLAS/LAPACK seam's argument mapping (blaze_amd/csrc/seam.cpp), linked against a host
stand-in of the ABI functions it calls (tests/seam_stub.cpp), against numpy and LAPACK
(scipy): CBLAS row/column-major with every transpose, padded leading dimensions, vector strides
(negative too), the complex forms with conjugate transposes, LAPACK pivots, factors, solves,
info codes."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

from tests import seam_cases as sc

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def stub_seam(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    out = tmp_path_factory.mktemp("seam") / "libseam_stub.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", f"-I{ROOT / 'include'}",
                    str(ROOT / "blaze_amd" / "csrc" / "seam.cpp"), str(ROOT / "tests" / "seam_stub.cpp"),
                    "-o", str(out)], check=True)
    return sc.bind(C.CDLL(str(out)))


def test_cblas_dgemm_mapping(stub_seam):
    sc.check_gemm(stub_seam, 7, 5, 4, seed=1, tol=1e-13)


def test_cblas_dgemv_mapping(stub_seam):
    sc.check_gemv(stub_seam, 6, 9, seed=2, tol=1e-13)


def test_lapack_mapping(stub_seam):
    sc.check_lapack(stub_seam, 9, 3, seed=3, tol=1e-12)


def test_cblas_complex_gemm_mapping(stub_seam):
    import numpy as np
    sc.check_cgemm(stub_seam, 7, 5, 4, seed=4, tol=1e-13)
    sc.check_cgemm(stub_seam, 6, 3, 5, seed=5, tol=1e-5, dtype=np.complex64)


def test_cblas_complex_gemv_mapping(stub_seam):
    import numpy as np
    sc.check_cgemv(stub_seam, 6, 9, seed=6, tol=1e-13)
    sc.check_cgemv(stub_seam, 5, 4, seed=7, tol=1e-5, dtype=np.complex64)
