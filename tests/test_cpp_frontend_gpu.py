"""GPU: the C++ front end (blaze::gpu containers, `C = A * B; y = A * x; y = S * x;
solve(A, x, b)`) -- runs tests/cpp/test_frontend.cpp, which mirrors the reference's
class/operation/LSE tests and exception messages, and requires every check to pass."""
import subprocess

import pytest

from blaze_amd import cxx
from tests.gpu_util import have_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_gpu(), reason="no GPU")]


def test_frontend_program():
    exe = cxx.BIN / "test_frontend"
    if not exe.exists():
        exe = cxx.build_examples()["test_frontend"]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "FAIL" not in r.stdout
