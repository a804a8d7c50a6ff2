"""The CG solvers' diagonal-storage SpMV, row by row, against the oracle and
against clsparse[S/D]csrmv (CSR-Adaptive) on the same matrix and x.

The solvers move a stencil matrix's values to diagonal storage and walk the
rows in line order when the largest offset is long (a 3-D grid's n_x n_y).
y starts as NaN, so a row the partition misses fails the comparison. Cases:
a line longer than the grid's 262,144 threads (600 x 600 x 10: line 360,000),
a line shorter than the grid (several stretches per line), a principal
submatrix whose size is not a multiple of the line, and a nonsymmetric stencil
(full storage instead of the symmetric half). Bars: per-row error
|y - y*| / sum |a_ij x_j| <= 1e-13 (FP64) and 1e-6 (FP32) against the oracle's
compensated long-double sums.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import clsparse_amd as cls
import oracle
from gpu_helpers import per_row_error, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


def _grid(name):
    if name == "600x600x10":
        return cls.gen_diffusion3d(600, 600, 10, sigma=1.0, ax=1.0, ay=1.0, az=10.0, seed=41)
    if name == "64x64x16":
        return cls.gen_diffusion3d(64, 64, 16, sigma=1.0, ax=1.0, ay=1.0, az=10.0, seed=42)
    if name == "sub-70x50x9":
        # A principal submatrix of a 70 x 50 x 9 grid: n = 29,999 is not a
        # multiple of the line (3,500), and the last line is short.
        rp, ci, v, m = cls.gen_diffusion3d(70, 50, 9, sigma=0.5, seed=43)
        A = sp.csr_matrix((v, ci, rp), shape=(m, m))[:29999, :29999].tocsr()
        A.sort_indices()
        return A.indptr.astype(np.uint32), A.indices.astype(np.uint32), A.data, 29999
    if name == "nonsym-64x64x16":
        rp, ci, v, m = cls.gen_diffusion3d(64, 64, 16, sigma=1.0, ax=1.0, ay=1.0, az=10.0, seed=44)
        rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(rp.astype(np.int64)))
        v = np.where(ci.astype(np.int64) > rows, v * 1.25, v)
        return rp, ci, v, m
    raise ValueError(name)


@pytest.mark.parametrize("name,ndiag,sym", [("600x600x10", 7, True), ("64x64x16", 7, True),
                                            ("sub-70x50x9", 7, True), ("nonsym-64x64x16", 7, False)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("fused", [0, 1], ids=["spmv", "fused-p"])
def test_diagonal_storage_spmv_every_row(ctl, name, ndiag, sym, dtype, fused):
    say(f"dia spmv {name} {np.dtype(dtype).name} fused={fused}")
    rp, ci, v, m = _grid(name)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    x = cls.gen_uniform(m, -1.0, 1.0, seed=7)
    A = cls.CsrMatrix(rp, ci, v, m, dtype=dtype)
    A.meta_create(ctl)
    xd = cls.DenseVector(torch.tensor(x.astype(dtype), device="cuda"))
    yd = cls.DenseVector(torch.full((m,), float("nan"), dtype=tdt, device="cuda"))
    yc = cls.DenseVector(torch.full((m,), float("nan"), dtype=tdt, device="cuda"))
    lib = cls.load()
    fn = lib.clsparseAmdScsrmvDiag if dtype == np.float32 else lib.clsparseAmdDcsrmvDiag
    nd = ctypes.c_int32(-1)
    cls.check("csrmvDiag", fn(ctypes.byref(A.s), ctypes.byref(xd.s), ctypes.byref(yd.s), fused,
                              ctypes.byref(nd), ctl.handle))
    assert nd.value == ndiag
    cls.check("csrmv", A.spmv(cls.Scalar(1.0, tdt), xd, cls.Scalar(0.0, tdt), yc, ctl))
    y_dia = yd.t.cpu().numpy().astype(np.float64)
    y_csr = yc.t.cpu().numpy().astype(np.float64)
    A.meta_delete()
    assert np.all(np.isfinite(y_dia)), int(np.argmin(np.isfinite(y_dia)))
    yref, scale = oracle.csrmv(rp, ci, np.asarray(v, dtype=dtype), x.astype(dtype))
    bar = 1e-13 if dtype == np.float64 else 1e-6
    err = per_row_error(y_dia, yref, scale)
    assert err.max() <= bar, (float(err.max()), int(np.argmax(err)))
    # Products rounded on their own and added in ascending column order, as
    # the CSR-Adaptive row-wise path adds them.
    d = np.abs(y_dia - y_csr) / np.maximum(scale, 1e-300)
    assert d.max() <= 2 * bar, (float(d.max()), int(np.argmax(d)))


def test_not_a_stencil_reports_zero_diagonals(ctl):
    say("dia spmv: power-law matrix is not a stencil")
    rp, ci, v, m = cls.gen_powerlaw(4096, 16.0, 1.5, 2000, 0, 0, seed=3)
    A = cls.CsrMatrix(rp, ci, v, m)
    xd = cls.DenseVector(torch.zeros(m, dtype=torch.float64, device="cuda"))
    yd = cls.DenseVector(torch.full((m,), 7.0, dtype=torch.float64, device="cuda"))
    nd = ctypes.c_int32(-1)
    cls.check("csrmvDiag", cls.load().clsparseAmdDcsrmvDiag(ctypes.byref(A.s), ctypes.byref(xd.s),
                                                           ctypes.byref(yd.s), 0, ctypes.byref(nd),
                                                           ctl.handle))
    assert nd.value == 0
    assert torch.all(yd.t == 7.0)
