"""clsparseAmdDcsrSumAxNoAlloc (no device allocation) against clsparseAmdDcsrSumAx.

Both run the same two kernels on the same grid, so the results must be
bit-identical; each is also held to the oracle's FP64 sum of y within
1e-13 * sum |a_ij x_j|. Two calls in a row, the second on a smaller matrix
(fewer partials than the first left in the shared device array), then the
larger one again.
"""
import ctypes

import numpy as np
import pytest
import torch

import clsparse_amd as cls
import oracle

from gpu_helpers import say

pytestmark = pytest.mark.gpu


def _case(rp, ci, v, m, seed):
    x = cls.gen_uniform(m, -1, 1, seed=seed)
    A = cls.CsrMatrix(rp, ci, v, m, dtype=np.float64)
    xd = cls.DenseVector(torch.tensor(x, dtype=torch.float64, device="cuda"))
    y, _ = oracle.csrmv(rp, ci, v, x)
    _, scale = oracle.csrmv(rp, ci, np.abs(v), np.abs(x))
    return A, xd, float(np.sum(y.astype(np.longdouble))), 1e-13 * float(np.sum(scale))


def test_no_alloc_sum_ax_matches_the_allocating_entry():
    lib = cls.load()
    with cls.Control() as ctl:
        big = cls.gen_powerlaw(200000, 16.0, 1.5, 200000, 2, 50000, seed=31)
        small = cls.gen_poisson2d(40, 30)
        cases = [("power-law 200000 rows", _case(*big, seed=5)),
                 ("Poisson 40x30", _case(*small, seed=6)),
                 ("power-law 200000 rows again", _case(*big, seed=5))]
        for name, (A, xd, ref, bound) in cases:
            say(f"no-alloc sum_ax: {name}, nnz {A.s.num_nonzeros}")
            a = cls.Scalar(0.0, torch.float64)
            b = cls.Scalar(7.0, torch.float64)
            assert lib.clsparseAmdDcsrSumAx(ctypes.byref(A.s), ctypes.byref(xd.s), ctypes.byref(a.s),
                                            ctl.handle) == 0
            assert lib.clsparseAmdDcsrSumAxNoAlloc(ctypes.byref(A.s), ctypes.byref(xd.s),
                                                   ctypes.byref(b.s), ctl.handle) == 0
            assert b.item() == a.item(), (name, b.item(), a.item())
            assert abs(b.item() - ref) <= bound, (name, b.item(), ref)


def test_no_alloc_sum_ax_argument_checks():
    lib = cls.load()
    with cls.Control() as ctl:
        empty = cls.CsrMatrix(np.zeros(4, np.uint32), np.zeros(0, np.uint32), np.zeros(0), 3,
                              dtype=np.float64)
        xd = cls.DenseVector(torch.ones(3, dtype=torch.float64, device="cuda"))
        out = cls.Scalar(5.0, torch.float64)
        say("no-alloc sum_ax: empty matrix")
        assert lib.clsparseAmdDcsrSumAxNoAlloc(ctypes.byref(empty.s), ctypes.byref(xd.s),
                                               ctypes.byref(out.s), ctl.handle) == 0
        assert out.item() == 0.0
        short = cls.DenseVector(torch.ones(2, dtype=torch.float64, device="cuda"))
        assert lib.clsparseAmdDcsrSumAxNoAlloc(ctypes.byref(empty.s), ctypes.byref(short.s),
                                               ctypes.byref(out.s), ctl.handle) == cls.clsparseInvalidSize
        assert lib.clsparseAmdDcsrSumAxNoAlloc(ctypes.byref(empty.s), ctypes.byref(xd.s),
                                               ctypes.byref(out.s), None) == cls.clsparseInvalidControlObject
