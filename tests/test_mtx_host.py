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
