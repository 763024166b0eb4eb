"""CPU checks of the C ABI library: it loads and exports every function include/blaze_gpu.h declares."""
import re
from pathlib import Path

import pytest

import blaze_amd as bz

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "blaze_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bzg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "bzg_dgemm" in names and "bzg_dcsrmv" in names and "bzg_dgesv_batched" in names


def test_library_exports_every_declared_symbol():
    if not bz.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = bz.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_signature_table_matches_header():
    assert set(bz.SIGNATURES) == set(declared_functions())


def test_status_strings_without_gpu():
    if not bz.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = bz.lib()
    assert lib.bzg_status_string(0) == b"ok"
    assert b"singular" in lib.bzg_status_string(3)
    assert lib.bzg_status_string(-2) == b"sizes do not match"
