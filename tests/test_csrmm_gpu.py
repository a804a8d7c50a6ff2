"""Sparse x dense (clsparse[S/D]csrmm) against the oracle's long-double product."""
import ctypes

import numpy as np
import pytest
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


def _dense(t, rows, cols, ld, off=0):
    d = cls.cldenseMatrix()
    cls.load().cldenseInitMatrix(ctypes.byref(d))
    d.num_rows, d.num_cols, d.lead_dim, d.major = rows, cols, ld, 1
    d.values, d.off_values = t.data_ptr(), off
    return d


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ncols", [1, 3, 8, 33, 64, 100])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.75, 1.0), (2.0, 0.5), (0.0, -1.5)])
def test_csrmm_matches_oracle(ctl, dtype, ncols, alpha, beta):
    rp, ci, v, m = cls.gen_powerlaw(3000, 10.0, 1.2, 2000, 2, 1500, seed=ncols)
    n = rp.size - 1
    rng = np.random.default_rng(ncols)
    ldb, ldc = ncols + 3, ncols + 5  # padded rows
    B = rng.uniform(-1, 1, (m, ldb)).astype(dtype)
    C0 = rng.uniform(-1, 1, (n, ldc)).astype(dtype)
    if beta == 0.0:
        C0[:, :ncols] = np.nan  # beta == 0 must not read C
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    A = cls.CsrMatrix(rp, ci, v.astype(dtype), m, dtype=dtype)
    Bt = torch.tensor(B.ravel(), device="cuda")
    Ct = torch.tensor(C0.ravel(), device="cuda")
    Bd = _dense(Bt, m, ncols, ldb)
    Cd = _dense(Ct, n, ncols, ldc)
    a, b = cls.Scalar(alpha, tdt), cls.Scalar(beta, tdt)
    lib = cls.load()
    fn = lib.clsparseScsrmm if dtype == np.float32 else lib.clsparseDcsrmm
    say(f"csrmm {np.dtype(dtype).name} ncols={ncols} alpha={alpha} beta={beta}")
    cls.check("csrmm", fn(ctypes.byref(a.s), ctypes.byref(A.s), ctypes.byref(Bd), ctypes.byref(b.s),
                          ctypes.byref(Cd), ctl.handle))
    C = Ct.cpu().numpy().reshape(n, ldc).astype(np.float64)
    Cin = np.where(np.isnan(C0), 0, C0)[:, :ncols].astype(np.float64)
    ref, scale = oracle.csrmm(rp, ci, v.astype(dtype).astype(np.float64),
                              B[:, :ncols].astype(np.float64), Cin, alpha, beta)
    got = C[:, :ncols]
    zero = scale == 0
    assert np.all(got[zero] == 0)
    err = np.abs(got - ref)[~zero] / scale[~zero]
    lens = np.diff(rp.astype(np.int64))
    bound = (1e-6 if dtype == np.float32 else 1e-13)
    # long rows accumulate more rounding: allow len * eps where it exceeds the bar
    eps = np.finfo(dtype).eps
    bound_rows = np.maximum(bound, (lens + 4) * eps)[:, None].repeat(ncols, 1)[~zero]
    assert np.all(err <= bound_rows)
    # padding columns untouched
    np.testing.assert_array_equal(C[:, ncols:], C0[:, ncols:].astype(np.float64))


def test_csrmm_size_checks(ctl):
    a = np.array([0, 1, 2], dtype=np.uint32)
    A = cls.CsrMatrix(a, np.array([0, 1], np.uint32), np.ones(2), 2)
    t = torch.zeros(16, dtype=torch.float64, device="cuda")
    s = cls.Scalar(1.0)
    lib = cls.load()
    Bd, Cd = _dense(t, 3, 2, 2), _dense(t, 2, 2, 2)  # B rows != A cols
    assert lib.clsparseDcsrmm(ctypes.byref(s.s), ctypes.byref(A.s), ctypes.byref(Bd), ctypes.byref(s.s),
                              ctypes.byref(Cd), ctl.handle) == cls.clsparseInvalidSize
    Bd, Cd = _dense(t, 2, 2, 1), _dense(t, 2, 2, 2)  # lead_dim < columns
    assert lib.clsparseDcsrmm(ctypes.byref(s.s), ctypes.byref(A.s), ctypes.byref(Bd), ctypes.byref(s.s),
                              ctypes.byref(Cd), ctl.handle) == cls.clsparseInvalidSize


@pytest.mark.parametrize("ncols", [1, 4, 32, 64])
def test_csrmm_many_row_steps_matches_oracle(ctl, ncols):
    """600k rows: every wave runs many pipelined row steps (the next step's
    nonzeros and the bounds of the step after are loaded ahead), including
    rows longer than the column tile, against the oracle."""
    rp, ci, v, m = cls.gen_powerlaw(600_000, 6.0, 1.4, 3000, 4, 2500, seed=40 + ncols)
    n = rp.size - 1
    rng = np.random.default_rng(ncols)
    B = rng.uniform(-1, 1, (m, ncols))
    C0 = rng.uniform(-1, 1, (n, ncols))
    A = cls.CsrMatrix(rp, ci, v, m)
    Bt = torch.tensor(B.ravel(), device="cuda")
    Ct = torch.tensor(C0.ravel(), device="cuda")
    Bd, Cd = _dense(Bt, m, ncols, ncols), _dense(Ct, n, ncols, ncols)
    a, b = cls.Scalar(1.5), cls.Scalar(0.5)
    say(f"csrmm 600k rows ncols={ncols}")
    cls.check("csrmm", cls.load().clsparseDcsrmm(ctypes.byref(a.s), ctypes.byref(A.s), ctypes.byref(Bd),
                                                  ctypes.byref(b.s), ctypes.byref(Cd), ctl.handle))
    got = Ct.cpu().numpy().reshape(n, ncols)
    ref, scale = oracle.csrmm(rp, ci, v, B, C0, 1.5, 0.5)
    lens = np.diff(rp.astype(np.int64))
    bound = np.maximum(1e-13, (lens + 4) * np.finfo(np.float64).eps)[:, None]
    assert np.all(np.abs(got - ref) <= bound * scale)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ncols,pad,off", [(4, 0, 0), (8, 4, 4), (12, 0, 4), (64, 0, 0), (96, 4, 0),
                                           (300, 4, 4), (8, 0, 1), (64, 4, 3)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.75, 1.0), (2.0, 0.5), (0.0, -1.5)])
def test_csrmm_column_groups_match_oracle(ctl, dtype, ncols, pad, off, alpha, beta):
    """Shapes whose columns, leading dimensions and base offsets are multiples
    of 16 bytes take 16-byte column groups per lane (several column tiles for
    300 columns); an offset of 1 or 3 elements keeps one column per lane. Both
    against the oracle, with every alpha / beta case."""
    rp, ci, v, m = cls.gen_powerlaw(2500, 9.0, 1.2, 1800, 2, 1200, seed=ncols + off)
    n = rp.size - 1
    rng = np.random.default_rng(ncols * 7 + off)
    ldb, ldc = ncols + pad, ncols + pad
    B = rng.uniform(-1, 1, (m, ldb)).astype(dtype)
    C0 = rng.uniform(-1, 1, (n, ldc)).astype(dtype)
    if beta == 0.0:
        C0[:, :ncols] = np.nan  # beta == 0 must not read C
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    A = cls.CsrMatrix(rp, ci, v.astype(dtype), m, dtype=dtype)
    lead = np.zeros(off, dtype=dtype)
    Bt = torch.tensor(np.concatenate([lead, B.ravel()]), device="cuda")
    Ct = torch.tensor(np.concatenate([lead, C0.ravel()]), device="cuda")
    Bd = _dense(Bt, m, ncols, ldb, off)
    Cd = _dense(Ct, n, ncols, ldc, off)
    a, b = cls.Scalar(alpha, tdt), cls.Scalar(beta, tdt)
    lib = cls.load()
    fn = lib.clsparseScsrmm if dtype == np.float32 else lib.clsparseDcsrmm
    say(f"csrmm groups {np.dtype(dtype).name} ncols={ncols} pad={pad} off={off} alpha={alpha} beta={beta}")
    cls.check("csrmm", fn(ctypes.byref(a.s), ctypes.byref(A.s), ctypes.byref(Bd), ctypes.byref(b.s),
                          ctypes.byref(Cd), ctl.handle))
    C = Ct.cpu().numpy()[off:].reshape(n, ldc).astype(np.float64)
    Cin = np.where(np.isnan(C0), 0, C0)[:, :ncols].astype(np.float64)
    ref, scale = oracle.csrmm(rp, ci, v.astype(dtype).astype(np.float64),
                              B[:, :ncols].astype(np.float64), Cin, alpha, beta)
    got = C[:, :ncols]
    zero = scale == 0
    assert np.all(got[zero] == 0)
    lens = np.diff(rp.astype(np.int64))
    eps = np.finfo(dtype).eps
    bound = np.maximum(1e-6 if dtype == np.float32 else 1e-13, (lens + 4) * eps)[:, None].repeat(ncols, 1)
    assert np.all(np.abs(got - ref)[~zero] <= bound[~zero] * scale[~zero])
    np.testing.assert_array_equal(C[:, ncols:], C0[:, ncols:].astype(np.float64))
