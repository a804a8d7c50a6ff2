"""Matrix Market readers into device memory (clsparse[S/D][Csr/Coo]MatrixfromFile)."""
import ctypes

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
import torch

import clsparse_amd as cls
import oracle

from gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


@pytest.fixture(scope="module")
def mtx(tmp_path_factory):
    rp, ci, v, m = cls.gen_powerlaw(5000, 8.0, 1.0, 400, 2, 2000, seed=9)
    a = sp.csr_matrix((v, ci, rp), shape=(5000, m))
    path = str(tmp_path_factory.mktemp("mm") / "pl.mtx")
    scipy.io.mmwrite(path, a, precision=17)
    return path, a


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_from_file_matches_host_reader_and_spmv(ctl, mtx, dtype):
    path, a = mtx
    lib = cls.load()
    nnz, rows, cols = cls.mtx_header(path)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    vals = torch.zeros(nnz, dtype=tdt, device="cuda")
    colt = torch.zeros(nnz, dtype=torch.int32, device="cuda")
    rpt = torch.zeros(rows + 1, dtype=torch.int32, device="cuda")
    A = cls.clsparseCsrMatrix()
    lib.clsparseInitCsrMatrix(ctypes.byref(A))
    A.values, A.col_indices, A.row_pointer = vals.data_ptr(), colt.data_ptr(), rpt.data_ptr()
    fn = lib.clsparseSCsrMatrixfromFile if dtype == np.float32 else lib.clsparseDCsrMatrixfromFile
    say(f"mtx: csr from file {np.dtype(dtype).name}")
    cls.check("CsrMatrixfromFile", fn(ctypes.byref(A), path.encode(), ctl.handle, 0))
    assert (A.num_rows, A.num_cols, A.num_nonzeros) == (5000, a.shape[1], a.nnz)
    rp_h, ci_h, v_h, _ = cls.read_mtx_host(path)
    np.testing.assert_array_equal(rpt.cpu().numpy().view(np.uint32), rp_h)
    np.testing.assert_array_equal(colt[:a.nnz].cpu().numpy().view(np.uint32), ci_h)
    np.testing.assert_array_equal(vals[:a.nnz].cpu().numpy(), v_h.astype(dtype))
    # SpMV on the matrix read from the file
    cls.check("MetaCreate", lib.clsparseCsrMetaCreate(ctypes.byref(A), ctl.handle))
    x = cls.gen_uniform(a.shape[1], -1, 1, seed=3).astype(dtype)
    xd = cls.DenseVector(torch.tensor(x, device="cuda"))
    yd = cls.DenseVector(torch.zeros(5000, dtype=tdt, device="cuda"))
    one, zero = cls.Scalar(1.0, tdt), cls.Scalar(0.0, tdt)
    spmv = lib.clsparseScsrmv if dtype == np.float32 else lib.clsparseDcsrmv
    cls.check("csrmv", spmv(ctypes.byref(one.s), ctypes.byref(A), ctypes.byref(xd.s),
                            ctypes.byref(zero.s), ctypes.byref(yd.s), ctl.handle))
    y = yd.t.cpu().numpy().astype(np.float64)
    yref, scale = oracle.csrmv(rp_h, ci_h, v_h.astype(dtype), x)
    tol = 1e-6 if dtype == np.float32 else 1e-13
    assert np.max(np.abs(y - yref) / np.maximum(scale, 1e-300)) <= tol
    lib.clsparseCsrMetaDelete(ctypes.byref(A))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_coo_from_file_sorted_by_row_then_column(ctl, mtx, dtype):
    path, a = mtx
    lib = cls.load()
    nnz, rows, cols = cls.mtx_header(path)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    vals = torch.zeros(nnz, dtype=tdt, device="cuda")
    ct = torch.zeros(nnz, dtype=torch.int32, device="cuda")
    rt = torch.zeros(nnz, dtype=torch.int32, device="cuda")
    C = cls.clsparseCooMatrix()
    lib.clsparseInitCooMatrix(ctypes.byref(C))
    C.values, C.col_indices, C.row_indices = vals.data_ptr(), ct.data_ptr(), rt.data_ptr()
    fn = lib.clsparseSCooMatrixfromFile if dtype == np.float32 else lib.clsparseDCooMatrixfromFile
    say(f"mtx: coo from file {np.dtype(dtype).name}")
    cls.check("CooMatrixfromFile", fn(ctypes.byref(C), path.encode(), ctl.handle, 0))
    k = C.num_nonzeros
    assert k == a.nnz
    coo = a.tocoo()
    order = np.lexsort((coo.col, coo.row))
    np.testing.assert_array_equal(rt[:k].cpu().numpy(), coo.row[order])
    np.testing.assert_array_equal(ct[:k].cpu().numpy(), coo.col[order])
    np.testing.assert_array_equal(vals[:k].cpu().numpy(), coo.data[order].astype(dtype))
