"""Matrix Market reader (host side) against scipy.io.mmread.

The reference reads .mtx files in src/library/io/mm-reader.cpp; its tests
take matrices from a directory given on the command line and ship none, so
scipy's independent reader is the check here, on files written by the test.
"""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import clsparse_amd as cls


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment
%another comment

4 5 7
3 2 1.5
1 1 -2.25e0
4 5 3.0
1 4 0.0
2 2 7
3 1 -1e-3
1 2 0.5
"""

SYMMETRIC = """%%MatrixMarket matrix coordinate real symmetric
3 3 5
1 1 4
2 1 -1
2 2 4
3 2 -2.5
3 3 9
"""


def _scipy_csr(path, drop_zeros=True):
    m = sp.csr_matrix(scipy.io.mmread(path))
    m.sort_indices()
    if drop_zeros:
        m.eliminate_zeros()
    return m


@pytest.mark.parametrize("text,name", [(GENERAL, "g.mtx"), (SYMMETRIC, "s.mtx")])
def test_reader_matches_scipy(tmp_path, text, name):
    path = _write(tmp_path, name, text)
    rp, ci, v, ncols = cls.read_mtx_host(path)
    ref = _scipy_csr(path)
    assert ncols == ref.shape[1] and rp.size - 1 == ref.shape[0]
    np.testing.assert_array_equal(rp, ref.indptr)
    np.testing.assert_array_equal(ci, ref.indices)
    np.testing.assert_array_equal(v, ref.data)


def test_explicit_zeroes_kept_on_request(tmp_path):
    path = _write(tmp_path, "g.mtx", GENERAL)
    rp, ci, v, _ = cls.read_mtx_host(path, read_explicit_zeroes=True)
    assert rp[-1] == 7 and 0.0 in v
    ref = _scipy_csr(path, drop_zeros=False)
    np.testing.assert_array_equal(ci, ref.indices)


def test_header_doubles_symmetric_nnz_and_checks_format(tmp_path):
    assert cls.mtx_header(_write(tmp_path, "s.mtx", SYMMETRIC)) == (10, 3, 3)
    assert cls.mtx_header(_write(tmp_path, "g.mtx", GENERAL)) == (7, 4, 5)
    lib = cls.load()
    import ctypes
    a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    args = (ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    # extension must be .mtx (mm-reader.cpp L447-454)
    assert lib.clsparseHeaderfromFile(*args, _write(tmp_path, "g.txt", GENERAL).encode()) == \
        cls.clsparseInvalidFileFormat
    assert lib.clsparseHeaderfromFile(*args, str(tmp_path / "missing.mtx").encode()) == \
        cls.clsparseInvalidFile
    cplx = GENERAL.replace("real", "complex")
    assert lib.clsparseHeaderfromFile(*args, _write(tmp_path, "c.mtx", cplx).encode()) == \
        cls.clsparseInvalidFile


def test_bad_files_are_rejected(tmp_path):
    lib = cls.load()
    m = cls.clsparseAmdHostCsr()
    import ctypes
    for name, text in [("oob.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n"),
                       ("short.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n"),
                       ("array.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n"),
                       ("nobanner.mtx", "2 2 1\n1 1 1.0\n")]:
        st = lib.clsparseAmdCsrMatrixfromFileHost(_write(tmp_path, name, text).encode(), 0,
                                                  ctypes.byref(m))
        assert st == cls.clsparseInvalidFile, name


def test_header_claiming_billions_of_entries_is_rejected_without_reserving_them(tmp_path):
    # Four billion entries announced, two present: the reader reserves what
    # the text can hold (24 bytes an entry would be 96 GB here) and reports a
    # short file, in the small-file and the threaded (>= 1 MiB) parse alike.
    import ctypes
    import resource
    lib = cls.load()
    m = cls.clsparseAmdHostCsr()
    small = "%%MatrixMarket matrix coordinate real general\n4 4 4000000000\n1 1 1.0\n2 2 2.0\n"
    big = small + "%" + "x" * (1 << 21) + "\n"
    before = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    for name, text in [("claims.mtx", small), ("claims_big.mtx", big)]:
        st = lib.clsparseAmdCsrMatrixfromFileHost(_write(tmp_path, name, text).encode(), 0, ctypes.byref(m))
        assert st == cls.clsparseInvalidFile, name
    grown_kib = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss - before
    assert grown_kib < 256 * 1024, grown_kib


def test_integer_and_pattern_files(tmp_path):
    ints = "%%MatrixMarket matrix coordinate integer general\n3 3 3\n3 3 -4\n1 2 7\n2 2 0\n"
    rp, ci, v, _ = cls.read_mtx_host(_write(tmp_path, "i.mtx", ints))
    np.testing.assert_array_equal(rp, [0, 1, 1, 2])
    np.testing.assert_array_equal(ci, [1, 2])
    np.testing.assert_array_equal(v, [7.0, -4.0])
    pat = "%%MatrixMarket matrix coordinate pattern symmetric\n4 4 3\n2 1\n4 4\n3 2\n"
    path = _write(tmp_path, "p.mtx", pat)
    rp, ci, v, _ = cls.read_mtx_host(path)
    ref = _scipy_csr(path)
    np.testing.assert_array_equal(rp, ref.indptr)
    np.testing.assert_array_equal(ci, ref.indices)
    # pattern values: 5 * rand() / (RAND_MAX + 1), mirrored pairs share a value
    assert np.all((v >= 0) & (v < 5.0))
    dense = sp.csr_matrix((v, ci, rp), shape=(4, 4)).toarray()
    np.testing.assert_array_equal(dense, dense.T)


def test_generated_matrix_roundtrip(tmp_path):
    rp, ci, v, m = cls.gen_powerlaw(3000, 6.0, 1.0, 200, 1, 500, seed=4)
    a = sp.csr_matrix((v, ci, rp), shape=(3000, m))
    path = str(tmp_path / "pl.mtx")
    scipy.io.mmwrite(path, a, precision=17)
    rp2, ci2, v2, m2 = cls.read_mtx_host(path)
    assert m2 == m
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(ci2, ci)
    np.testing.assert_array_equal(v2, v)


def _big_text(rng, n=40000, per_row=6, header_nnz=None, symmetric=False, shuffle=True):
    """A file over 1 MB (the reader splits it over threads): random entries,
    optionally shuffled so the (row, column) sort has work to do."""
    rows = np.repeat(np.arange(1, n + 1), per_row)
    cols = rng.integers(1, n + 1, rows.size)
    if symmetric:
        keep = cols <= rows
        rows, cols = rows[keep], cols[keep]
    vals = rng.uniform(-1, 1, rows.size)
    vals[rng.random(rows.size) < 0.01] = 0.0  # explicit zeroes to drop
    order = rng.permutation(rows.size) if shuffle else np.arange(rows.size)
    rows, cols, vals = rows[order], cols[order], vals[order]
    kind = "symmetric" if symmetric else "general"
    body = "\n".join(f"{r} {c} {x:.17g}" for r, c, x in zip(rows, cols, vals))
    nnz = rows.size if header_nnz is None else header_nnz
    return f"%%MatrixMarket matrix coordinate real {kind}\n{n} {n} {nnz}\n{body}\n", rows.size


def _expected_csr(text):
    """The reader's rules restated with numpy: zeroes dropped, a symmetric
    file's off-diagonal entry followed by its mirror, then a stable sort by
    (row, column) -- duplicates stay, in file order."""
    lines = text.splitlines()
    kind = lines[0].split()[-1]
    n, _, nnz = (int(t) for t in lines[1].split())
    arr = np.array([ln.split() for ln in lines[2:2 + nnz]], dtype=np.float64)
    r, c, v = arr[:, 0].astype(np.int64) - 1, arr[:, 1].astype(np.int64) - 1, arr[:, 2]
    keep = v != 0
    r, c, v = r[keep], c[keep], v[keep]
    if kind == "symmetric":
        off = r != c
        idx = np.repeat(np.arange(r.size), 1 + off)
        first = np.ones(idx.size, bool)
        first[1:] = idx[1:] != idx[:-1]
        rr = np.where(first, r[idx], c[idx])
        cc = np.where(first, c[idx], r[idx])
        r, c, v = rr, cc, v[idx]
    order = np.lexsort((c, r))
    rp = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))])
    return rp, c[order], v[order]


@pytest.mark.parametrize("symmetric", [False, True])
def test_large_file_parallel_parse_matches_scipy(tmp_path, symmetric):
    """Over 1 MB the body is parsed in newline-aligned chunks on several
    threads; duplicates, zeroes and symmetric expansion as in the small case
    (scipy sums duplicates, so the expected CSR is formed with numpy)."""
    text, _ = _big_text(np.random.default_rng(7 + symmetric), symmetric=symmetric)
    assert len(text) > (1 << 20)
    path = _write(tmp_path, "big.mtx", text)
    rp, ci, v, ncols = cls.read_mtx_host(path)
    erp, eci, ev = _expected_csr(text)
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(ci, eci)
    np.testing.assert_array_equal(v, ev)


def test_large_file_entries_across_lines_and_trailing_text(tmp_path):
    """Entries whose tokens span lines (accepted by the reference's fscanf)
    make the chunked parse fall back to the sequential one; text after the
    announced entries is ignored; a bad token among them is an error."""
    rng = np.random.default_rng(3)
    text, count = _big_text(rng, shuffle=False)
    head, body = text.split("\n", 2)[:2], text.split("\n", 2)[2]
    lines = body.splitlines()
    mid = len(lines) // 2
    r, c, x = lines[mid].split()
    lines[mid] = f"{r}\n{c}\n{x}"  # one entry over three lines, in the middle of the file
    spread = "\n".join(head) + "\n" + "\n".join(lines) + "\n"
    ref_path = _write(tmp_path, "ref.mtx", text)
    p1 = _write(tmp_path, "spread.mtx", spread)
    ref = cls.read_mtx_host(ref_path)
    got = cls.read_mtx_host(p1)
    for a, b in zip(got[:3], ref[:3]):
        np.testing.assert_array_equal(a, b)
    trailing = text + "this is not an entry\n" * 50000
    got = cls.read_mtx_host(_write(tmp_path, "trail.mtx", trailing))
    for a, b in zip(got[:3], ref[:3]):
        np.testing.assert_array_equal(a, b)
    lines = body.splitlines()
    lines[len(lines) // 3] = "12 x 1.0"
    bad = "\n".join(head) + "\n" + "\n".join(lines) + "\n"
    lib = cls.load()
    import ctypes
    m = cls.clsparseAmdHostCsr()
    assert lib.clsparseAmdCsrMatrixfromFileHost(_write(tmp_path, "bad.mtx", bad).encode(), 0,
                                                ctypes.byref(m)) == cls.clsparseInvalidFile
    short, _ = _big_text(rng, header_nnz=count + 5)
    assert lib.clsparseAmdCsrMatrixfromFileHost(_write(tmp_path, "short.mtx", short).encode(), 0,
                                                ctypes.byref(m)) == cls.clsparseInvalidFile
