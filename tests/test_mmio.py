"""CPU checks of the Matrix Market reader (bzg_mm_read_csr, host code: no device call).

Pinned by an independent writer/reader: scipy.io.mmwrite files (general and symmetric, real
and integer) read back equal scipy's own CSR of the same matrix; hand-written files cover the
pattern and skew-symmetric forms, comments and blank lines, duplicate entries (summed in file
order) and every error path.
"""
import numpy as np
import pytest

import blaze_amd as bz

pytestmark = pytest.mark.skipif(not bz.LIB_PATH.exists(), reason="library not built")


def _dense(rp, ci, va, shape):
    D = np.zeros(shape)
    for i in range(shape[0]):
        D[i, ci[rp[i]:rp[i + 1]]] += va[rp[i]:rp[i + 1]]
    return D


def _sorted_rows(rp, ci):
    return all(np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0) for i in range(len(rp) - 1))


def test_scipy_written_files(tmp_path):
    import scipy.io
    import scipy.sparse as sp
    rng = np.random.default_rng(4)
    A = sp.random(120, 75, density=0.06, random_state=5, format="coo")
    scipy.io.mmwrite(str(tmp_path / "g.mtx"), A)
    rp, ci, va, shape = bz.read_matrix_market(tmp_path / "g.mtx")
    ref = A.tocsr(); ref.sort_indices()
    assert shape == (120, 75) and np.array_equal(rp, ref.indptr) and np.array_equal(ci, ref.indices)
    assert np.array_equal(va, ref.data) and _sorted_rows(rp, ci)
    S = sp.random(60, 60, density=0.08, random_state=6)
    S = (S + S.T).tocoo()
    scipy.io.mmwrite(str(tmp_path / "s.mtx"), S, symmetry="symmetric")
    rp, ci, va, shape = bz.read_matrix_market(tmp_path / "s.mtx")
    assert np.array_equal(_dense(rp, ci, va, shape), S.toarray()) and _sorted_rows(rp, ci)
    Ii = sp.coo_matrix(rng.integers(-5, 6, (30, 40)) * (rng.random((30, 40)) < 0.1))
    scipy.io.mmwrite(str(tmp_path / "i.mtx"), Ii, field="integer")
    rp, ci, va, shape = bz.read_matrix_market(tmp_path / "i.mtx")
    assert np.array_equal(_dense(rp, ci, va, shape), Ii.toarray())


def test_pattern_skew_duplicates_comments(tmp_path):
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n% a comment\n\n3 4 4\n"
                 "3 1\n1 4\n1 2\n% another\n2 3\n")
    rp, ci, va, shape = bz.read_matrix_market(f)
    assert shape == (3, 4) and list(rp) == [0, 2, 3, 4] and list(ci) == [1, 3, 2, 0] and np.all(va == 1.0)
    f.write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 1.5\n3 2 -2\n")
    rp, ci, va, shape = bz.read_matrix_market(f)
    assert np.array_equal(_dense(rp, ci, va, shape), [[0, -1.5, 0], [1.5, 0, 2], [0, -2, 0]])
    # duplicates are summed in file order: (0.1 + 0.2) + 0.3, not 0.1 + (0.2 + 0.3)
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 2 4\n1 1 0.1\n2 2 5\n1 1 0.2\n1 1 0.3\n")
    rp, ci, va, shape = bz.read_matrix_market(f)
    assert list(rp) == [0, 1, 2] and va[0] == (0.1 + 0.2) + 0.3 and va[1] == 5.0
    f.write_text("%%MatrixMarket matrix coordinate real general\n4 3 0\n")
    rp, ci, va, shape = bz.read_matrix_market(f)
    assert shape == (4, 3) and list(rp) == [0, 0, 0, 0, 0] and len(ci) == 0


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "UNSUPPORTED"),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "UNSUPPORTED"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n", "INVALID"),      # truncated
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", "INVALID"),      # out of range
    ("%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n1 2 1.0\n", "INVALID"),    # upper triangle
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 1.0\n", "INVALID"),
    ("not a header\n1 1 1\n1 1 1.0\n", "INVALID"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 x 1.0\n", "INVALID"),
])
def test_errors(tmp_path, text, code):
    f = tmp_path / "bad.mtx"
    f.write_text(text)
    with pytest.raises(bz.BlazeGpuError):
        bz.read_matrix_market(f)
    st = bz.lib().bzg_mm_read_csr(str(f).encode(), bz.C.byref(bz.CsrHost()))
    assert st == {"INVALID": -1, "UNSUPPORTED": -6}[code] or st < 0


def test_missing_file(tmp_path):
    with pytest.raises(bz.BlazeGpuError):
        bz.read_matrix_market(tmp_path / "absent.mtx")


def test_mutated_files_never_crash(tmp_path):
    # seeded mutations of valid files (truncation, byte garbage, token swaps, huge counts):
    # each read either returns a well-formed CSR or an error status, never a crash
    rng = np.random.default_rng(9)
    seeds = [
        "%%MatrixMarket matrix coordinate real general\n% c\n4 5 6\n1 1 1.5\n2 3 -2\n4 5 3\n3 3 7\n1 5 2\n4 1 -1\n",
        "%%MatrixMarket matrix coordinate pattern symmetric\n5 5 4\n2 1\n3 3\n5 2\n4 1\n",
        "%%MatrixMarket matrix coordinate integer skew-symmetric\n3 3 2\n2 1 4\n3 1 -5\n",
    ]
    tokens = ["0", "-1", "99999999999", "1e308", "nan", "x", "%", "\n", " ", "3", "2147483648"]
    f = tmp_path / "m.mtx"
    for it in range(300):
        s = list(seeds[it % len(seeds)])
        kind = it % 4
        if kind == 0:      # truncate
            s = s[:int(rng.integers(0, len(s)))]
        elif kind == 1:    # random bytes
            for _ in range(int(rng.integers(1, 6))):
                s[int(rng.integers(0, len(s)))] = chr(int(rng.integers(32, 127)))
        elif kind == 2:    # token insertion
            pos = int(rng.integers(0, len(s)))
            s[pos:pos] = list(" " + tokens[int(rng.integers(0, len(tokens)))] + " ")
        else:              # change the declared entry count / sizes
            text = "".join(s).split("\n")
            for li, line in enumerate(text):
                if line and not line.startswith("%"):
                    text[li] = " ".join(tokens[int(rng.integers(0, len(tokens)))] if rng.random() < 0.5 else w
                                        for w in line.split())
                    break
            s = list("\n".join(text))
        f.write_text("".join(s))
        try:
            rp, ci, va, (m, n) = bz.read_matrix_market(f)
        except bz.BlazeGpuError:
            continue
        assert len(rp) == m + 1 and rp[0] == 0 and np.all(np.diff(rp) >= 0) and rp[-1] == len(ci) == len(va)
        assert np.all((ci >= 0) & (ci < max(n, 1)))
        assert _sorted_rows(rp, ci)
