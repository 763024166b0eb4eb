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


def test_csc_to_csr_host_pack():
    # host-only entry point (no device call): the CSR of trans(S) against numpy's stable
    # column sort of S's entries; values move with their indices, rows list columns ascending
    if not bz.LIB_PATH.exists():
        pytest.skip("library not built")
    import ctypes as C

    import numpy as np
    lib = bz.lib()
    rng = np.random.default_rng(3)
    n_rows, n_cols = 300, 200
    dense = (rng.random((n_rows, n_cols)) < 0.05) * rng.uniform(-1, 1, (n_rows, n_cols))
    rows, cols = np.nonzero(dense)  # row-major order: S's CSR
    rp = np.zeros(n_rows + 1, np.int32); np.add.at(rp, rows + 1, 1); rp = np.cumsum(rp).astype(np.int32)
    ci = cols.astype(np.int32); v = dense[rows, cols]
    # S's CSR read as CSC is trans(S): n_rows(trans) = n_cols, n_cols(trans) = n_rows
    trp = np.zeros(n_cols + 1, np.int32); tci = np.zeros(len(v), np.int32); tv = np.zeros(len(v))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert lib.bzg_csc_to_csr_host(n_cols, n_rows, p(rp), p(ci), p(v), 8, p(trp), p(tci), p(tv)) == 0
    order = np.argsort(ci, kind="stable")
    want_rp = np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n_cols))]).astype(np.int32)
    assert np.array_equal(trp, want_rp)
    assert np.array_equal(tci, rows[order].astype(np.int32)) and np.array_equal(tv, v[order])
    for j in range(n_cols):
        assert np.all(np.diff(tci[trp[j]:trp[j + 1]]) > 0)
    # float values, and an out-of-range row index is rejected
    vf = v.astype(np.float32); tvf = np.zeros(len(v), np.float32)
    assert lib.bzg_csc_to_csr_host(n_cols, n_rows, p(rp), p(ci), p(vf), 4, p(trp), p(tci), p(tvf)) == 0
    assert np.array_equal(tvf, vf[order])
    bad = ci.copy(); bad[5] = n_cols
    assert lib.bzg_csc_to_csr_host(n_cols, n_rows, p(rp), p(bad), p(v), 8, p(trp), p(tci), p(tv)) != 0
