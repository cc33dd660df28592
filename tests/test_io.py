"""Host readers (no GPU): Matrix Market as mmio_allinone and edge lists as
CSR::read_from_edgelist.  The edge-list reader is compared with the compiled
reference on the reference's own data files; the MTX reader, whose reference
lives in a CUDA header, is pinned by hand-built files with known CSR."""
import ctypes
import os

import numpy as np
import pytest

import oracle_ffi as of

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REF_DATA = "/root/reference/preprocess/data"


def _smat():
    import smat
    return smat


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_mtx_general_file_order_and_padding(tmp_path):
    smat = _smat()
    p = _write(tmp_path, "a.mtx", "%%MatrixMarket matrix coordinate real general\n% comment\n3 5 4\n"
               "2 5 1.5\n1 3 -2\n2 1 0.25\n3 3 1e-9\n")
    c = smat.read_mtx(p, k=16)
    assert (c["rows"], c["cols"], c["nnz"]) == (16, 16, 4)
    assert c["rowptr"][:4].tolist() == [0, 1, 3, 4] and c["rowptr"][-1] == 4
    assert c["colidx"].tolist() == [2, 4, 0, 2]  # row 1 keeps file order: 5 then 1
    assert np.array_equal(c["values_f16"], np.float16([-2, 1.5, 0.25, 1e-9]).view(np.uint16))
    assert np.array_equal(c["values_f32"], np.float32([-2, 1.5, 0.25, 1e-9]))


def test_mtx_symmetric_mirror_order(tmp_path):
    smat = _smat()
    p = _write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate real symmetric\n4 4 4\n"
               "2 1 1\n3 3 2\n4 2 3\n4 1 4\n")
    c = smat.read_mtx(p, k=32)
    assert (c["rows"], c["cols"], c["nnz"]) == (16, 32, 7) and c["symmetric"]
    rp, ci = c["rowptr"], c["colidx"]
    rows = [ci[rp[r]:rp[r + 1]].tolist() for r in range(4)]
    # mmio_allinone places (i,j) then (j,i) entry by entry: row 0 gets 1 (from 2 1), then 3 (from 4 1)
    assert rows == [[1, 3], [0, 3], [2], [1, 0]]
    vals = c["values_f16"].view(np.float16)
    assert vals[rp[3]:rp[4]].tolist() == [3.0, 4.0]


def test_mtx_pattern_integer_complex(tmp_path):
    smat = _smat()
    p = _write(tmp_path, "p.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 1\n2 2\n")
    assert smat.read_mtx(p)["values_f32"].tolist() == [1.0, 1.0]
    p = _write(tmp_path, "i.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 2 -7\n")
    assert smat.read_mtx(p)["values_f32"].tolist() == [-7.0]
    p = _write(tmp_path, "c.mtx", "%%MatrixMarket matrix coordinate complex hermitian\n2 2 2\n1 1 2 0\n2 1 3 4\n")
    c = smat.read_mtx(p)
    assert c["nnz"] == 3 and c["values_f32"].tolist() == [2.0, 3.0, 3.0]


def test_mtx_double_to_half_single_rounding(tmp_path):
    smat = _smat()
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.standard_normal(3000) * 10.0 ** rng.integers(-9, 6, 3000),
                           [65504.0, 65519.99, 65520.0, 1e6, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                            1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 1 + 2.0 ** -11 + 2.0 ** -40, -0.0]])
    lines = "".join(f"{i + 1} 1 {float(v)!r}\n" for i, v in enumerate(vals))
    p = _write(tmp_path, "r.mtx", f"%%MatrixMarket matrix coordinate real general\n{vals.size} 1 {vals.size}\n{lines}")
    c = smat.read_mtx(p)
    want = vals.astype(np.float16).view(np.uint16)  # numpy: correctly rounded double -> half
    assert np.array_equal(c["values_f16"], want)


def test_mtx_rejects_bad_input(tmp_path):
    smat = _smat()
    for text in ["%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
                 "hello\n", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
                 "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1\n"]:
        with pytest.raises(smat.SmatError):
            smat.read_mtx(_write(tmp_path, "bad.mtx", text))


def test_edgelist_quirks(tmp_path):
    smat = _smat()
    # rows 0 and 2 (row 1 empty), a two-field weighted line takes its column as value
    p = _write(tmp_path, "q.el", "# head\n0 3 1.5\n0 1 2\n2 7\n")
    c = smat.read_edgelist(p)
    assert (c["rows"], c["cols"], c["nnz"]) == (3, 8, 3)
    assert c["rowptr"].tolist() == [0, 2, 2, 3] and c["colidx"].tolist() == [3, 1, 7]
    assert c["values_f32"].tolist() == [1.5, 2.0, 7.0]
    with pytest.raises(smat.SmatError):
        smat.read_edgelist(_write(tmp_path, "d.el", "1 2 1\n0 1 1\n"))  # descending rows


def _ref_read(path, pattern_only):
    rows, cols = ctypes.c_long(), ctypes.c_long()
    f = of.ref().ref_read_edgelist
    f.restype = ctypes.c_long
    f.argtypes = [ctypes.c_char_p, ctypes.c_int] + [ctypes.POINTER(ctypes.c_long)] * 2 + [ctypes.c_void_p] * 3 + \
        [ctypes.c_long] * 2
    cap = 10_000_000
    rp = np.zeros(cap, np.int64)
    ci = np.zeros(cap, np.int64)
    v = np.zeros(cap, np.float32)
    nnz = f(path.encode(), int(pattern_only), ctypes.byref(rows), ctypes.byref(cols), rp.ctypes.data, ci.ctypes.data,
            v.ctypes.data, cap, cap)
    assert nnz >= 0
    return rows.value, cols.value, rp[:rows.value + 1], ci[:nnz], v[:nnz]


def _fixtures():
    files = [os.path.join(GOLDEN, f) for f in ("TEST_matrix_weighted.el", "cs_department.el", "seventh_graders.el")]
    if os.path.isdir(REF_DATA):
        files += [os.path.join(REF_DATA, "minitest", f) for f in ("bcsstk18_r.el", "wiki-Vote_r.el")]
    return files


@pytest.mark.parametrize("path", _fixtures(), ids=os.path.basename)
@pytest.mark.parametrize("pattern_only", [False, True])
def test_edgelist_matches_compiled_reference(path, pattern_only):
    if not of.ref_available():
        pytest.skip("reference build (oracle/_ref) not available")
    smat = _smat()
    c = smat.read_edgelist(path, pattern_only=pattern_only)
    rows, cols, rp, ci, v = _ref_read(path, pattern_only)
    assert (c["rows"], c["cols"]) == (rows, cols)
    assert np.array_equal(c["rowptr"], rp) and np.array_equal(c["colidx"], ci)
    assert np.array_equal(c["values_f32"], v)


def test_edgelist_kat_matrix():
    """The reference's KAT input (TEST_results.txt: 9 rows, 9 cols, 13 nnz)."""
    smat = _smat()
    c = smat.read_edgelist(os.path.join(GOLDEN, "TEST_matrix_weighted.el"))
    assert (c["rows"], c["cols"], c["nnz"]) == (9, 9, 13)


def _ref_read_opts(path, pattern_only, mat_fmt, symmetrize):
    rows, cols = ctypes.c_long(), ctypes.c_long()
    f = of.ref().ref_read_edgelist_opts
    f.restype = ctypes.c_long
    f.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_long)] * 2 + \
        [ctypes.c_void_p] * 3 + [ctypes.c_long] * 2
    cap = 1_000_000
    rp = np.zeros(cap, np.int64)
    ci = np.zeros(cap, np.int64)
    v = np.zeros(cap, np.float32)
    nnz = f(path.encode(), int(pattern_only), int(mat_fmt), int(symmetrize), ctypes.byref(rows), ctypes.byref(cols),
            rp.ctypes.data, ci.ctypes.data, v.ctypes.data, cap, cap)
    if nnz < 0:
        return None
    return rows.value, cols.value, rp[:rows.value + 1], ci[:nnz], v[:nnz]


def _mtx_text(rng, n, m, nnz, weighted):
    """A Matrix Market coordinate file as SuiteSparse writes them: 1-based, sorted by column."""
    cells = rng.choice(n * m, size=nnz, replace=False)
    r, c = cells // m + 1, cells % m + 1
    order = np.lexsort((r, c))
    lines = ["%%MatrixMarket matrix coordinate " + ("real" if weighted else "pattern") + " general",
             "% written for a test", f"{n} {m} {nnz}"]
    for i in order:
        lines.append(f"{r[i]} {c[i]} {rng.uniform(-4, 4):.6g}" if weighted else f"{r[i]} {c[i]}")
    return "\n".join(lines) + "\n"


def _upper_el(rng, n, nnz):
    """Pattern edge list with every entry on or above the diagonal, the last row holding (n-1, n-1)."""
    cells = set()
    while len(cells) < nnz:
        i, j = sorted(rng.integers(0, n, 2).tolist())
        cells.add((i, j))
    cells.add((n - 1, n - 1))
    return "".join(f"{i} {j}\n" for i, j in sorted(cells))


CASES_OPTS = [("mtx_weighted", 1, 0, False), ("mtx_pattern", 1, 0, True), ("el_sym", 0, 1, True),
              ("mtx_sym", 1, 1, True), ("el_sym_weighted", 0, 1, False), ("el_nonsym_sym", 0, 1, True)]


@pytest.mark.parametrize("name,mat_fmt,symmetrize,pattern_only", CASES_OPTS)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_edgelist_options_match_compiled_reference(tmp_path, name, mat_fmt, symmetrize, pattern_only, seed):
    """read_from_edgelist's MTX mode (size line skipped, index fields swapped)
    and symmetrize (csr.cpp:184, 199-202, 241-260) against the compiled
    reference on generated files; where the reference rejects a file
    (descending rows, symmetrize of weighted input), so does the reader."""
    if not of.ref_available():
        pytest.skip("reference build (oracle/_ref) not available")
    smat = _smat()
    rng = np.random.default_rng(seed)
    if name.startswith("mtx_sym"):
        # lower triangle in MTX: after the field swap every entry is on or above the diagonal
        cells = sorted({tuple(sorted(rng.integers(1, 40, 2).tolist(), reverse=True)) for _ in range(120)} | {(40, 40)},
                       key=lambda t: (t[1], t[0]))
        text = "%%MatrixMarket matrix coordinate pattern symmetric\n40 40 %d\n" % len(cells) + \
            "".join(f"{i} {j}\n" for i, j in cells)
    elif name.startswith("mtx"):
        text = _mtx_text(rng, 50, 37, 200, weighted=not pattern_only)
    elif name == "el_nonsym_sym":
        text = "0 1\n0 3\n1 0\n2 2\n3 1\n"
    elif name == "el_sym_weighted":
        text = "".join(f"{line} 1.5\n" for line in _upper_el(rng, 30, 60).splitlines())
    else:
        text = _upper_el(rng, 30, 60)
    p = _write(tmp_path, "f.txt", text)
    ref = _ref_read_opts(p, pattern_only, mat_fmt, symmetrize)
    if ref is None:
        with pytest.raises(smat.SmatError):
            smat.read_edgelist(p, pattern_only=pattern_only, mat_fmt=mat_fmt, symmetrize=bool(symmetrize))
        return
    rows, cols, rp, ci, v = ref
    for c in (smat.read_edgelist(p, pattern_only=pattern_only, mat_fmt=mat_fmt, symmetrize=bool(symmetrize)),
              smat.parse_edgelist(text, pattern_only=pattern_only, mat_fmt=mat_fmt, symmetrize=bool(symmetrize))):
        assert (c["rows"], c["cols"]) == (rows, cols)
        assert np.array_equal(c["rowptr"], rp) and np.array_equal(c["colidx"], ci)
        assert np.array_equal(c["values_f32"], v)
    if symmetrize and name in ("el_sym", "mtx_sym"):  # the mirrored pattern is symmetric
        dense = np.zeros((rows, rows), bool)
        dense[np.repeat(np.arange(rows), np.diff(rp)), ci] = True
        assert np.array_equal(dense, dense.T)


def test_symmetrize_column_without_row_is_refused():
    """An entry whose column lies past the last row: the reference indexes
    past its row list (csr.cpp:248-249); the reader refuses the file."""
    smat = _smat()
    with pytest.raises(smat.SmatError) as e:
        smat.parse_edgelist("0 1\n0 5\n1 2\n", pattern_only=True, symmetrize=True)
    assert "no row" in str(e.value)
