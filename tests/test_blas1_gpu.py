"""Dense BLAS-1 on the GPU (cldense*) against the oracle and closed forms.

Reductions accumulate in FP64 for both precisions; FP64 results must match the
compensated oracle to 1e-14 relative, FP32 results to one FP32 rounding of it.
"""
import ctypes

import numpy as np
import pytest
import torch

import clsparse_amd as cls
import oracle
from gpu_helpers import say

pytestmark = pytest.mark.gpu
B = ctypes.byref


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


def _vec(a, dt):
    return cls.DenseVector(torch.tensor(np.asarray(a), dtype=dt, device="cuda"))


@pytest.mark.parametrize("n", [0, 1, 1000, 65536, 1 << 21])
@pytest.mark.parametrize("prec", ["S", "D"])
def test_reductions_vs_oracle(ctl, prec, n):
    say(f"reductions {prec} n={n}")
    lib = cls.load()
    dt = torch.float32 if prec == "S" else torch.float64
    x = cls.gen_uniform(n, -1, 1, seed=1)
    y = cls.gen_uniform(n, -1, 1, seed=2)
    if prec == "S":
        x, y = x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)
    xd, yd = _vec(x, dt), _vec(y, dt)
    out = cls.Scalar(0.0, dt)
    rtol = 1e-14 if prec == "D" else 2.0 ** -23
    checks = [("dot", lambda: getattr(lib, f"cldense{prec}dot")(B(out.s), B(xd.s), B(yd.s), ctl.handle),
               oracle.dot(x, y), np.abs(x * y).sum()),
              ("nrm1", lambda: getattr(lib, f"cldense{prec}nrm1")(B(out.s), B(xd.s), ctl.handle),
               oracle.nrm1(x), np.abs(x).sum()),
              ("nrm2", lambda: getattr(lib, f"cldense{prec}nrm2")(B(out.s), B(xd.s), ctl.handle),
               oracle.nrm2(x), oracle.nrm2(x)),
              ("reduce", lambda: getattr(lib, f"cldense{prec}reduce")(B(out.s), B(xd.s), ctl.handle),
               float(np.sum(x, dtype=np.longdouble)), np.abs(x).sum())]
    for name, call, ref, mag in checks:
        assert call() == 0, name
        got = out.item()
        assert abs(got - ref) <= rtol * max(mag, 1e-300) + 1e-300, (name, got, ref)


@pytest.mark.parametrize("prec", ["S", "D"])
def test_constant_vectors_closed_form(ctl, prec):
    """x = 4, y = 2 (src/tests/test-blas1.cpp L61-62): dot = 8N, nrm1 = 4N, nrm2 = 4 sqrt(N)."""
    lib = cls.load()
    dt = torch.float32 if prec == "S" else torch.float64
    n = 1024
    xd = cls.DenseVector(torch.full((n,), 4.0, dtype=dt, device="cuda"))
    yd = cls.DenseVector(torch.full((n,), 2.0, dtype=dt, device="cuda"))
    out = cls.Scalar(0.0, dt)
    assert getattr(lib, f"cldense{prec}dot")(B(out.s), B(xd.s), B(yd.s), ctl.handle) == 0
    assert out.item() == 8.0 * n
    assert getattr(lib, f"cldense{prec}nrm1")(B(out.s), B(xd.s), ctl.handle) == 0
    assert out.item() == 4.0 * n
    assert getattr(lib, f"cldense{prec}nrm2")(B(out.s), B(xd.s), ctl.handle) == 0
    assert out.item() == 4.0 * 32.0


def test_integer_reduce_wraps(ctl):
    lib = cls.load()
    n = 1003
    x = torch.full((n,), 2 ** 30, dtype=torch.int32, device="cuda")
    xd = cls.DenseVector(x)
    out = cls.Scalar(torch.zeros(1, dtype=torch.int32, device="cuda"))
    assert lib.cldenseIreduce(B(out.s), B(xd.s), ctl.handle) == 0
    assert out.item() == ((n * 2 ** 30 + 2 ** 31) % 2 ** 32) - 2 ** 31  # two's-complement wrap


@pytest.mark.parametrize("prec", ["S", "D"])
def test_axpy_axpby_scale_special_cases(ctl, prec):
    say(f"axpy/axpby/scale {prec}")
    lib = cls.load()
    dt = torch.float32 if prec == "S" else torch.float64
    npdt = np.float32 if prec == "S" else np.float64
    n = 5000
    x = cls.gen_uniform(n, -1, 1, seed=3).astype(npdt)
    y = cls.gen_uniform(n, -1, 1, seed=4).astype(npdt)
    nan = np.full(n, np.nan, dtype=npdt)
    r = _vec(np.zeros(n), dt)

    keep = []  # device objects must outlive the call that uses their pointers

    def a(v):
        keep.append(cls.Scalar(v, dt))
        return B(keep[-1].s)

    def vec(arr):
        keep.append(_vec(arr, dt))
        return B(keep[-1].s)

    def run(fn, *args):
        assert getattr(lib, f"cldense{prec}{fn}")(*args, ctl.handle) == 0
        return r.t.cpu().numpy()

    X, Y, NAN = vec(x), vec(y), vec(nan)
    # axpy: r = alpha x + y; alpha == 0 copies y even where x is NaN.
    got = run("axpy", B(r.s), a(0.5), X, Y)
    mag = np.abs(0.5 * x.astype(np.float64)) + np.abs(y.astype(np.float64))
    assert np.all(np.abs(got.astype(np.float64) - (npdt(0.5) * x + y)) <=
                  4 * np.finfo(npdt).eps * mag)
    got = run("axpy", B(r.s), a(0.0), NAN, Y)
    np.testing.assert_array_equal(got, y)
    # scale: alpha == 0 writes zeros even for NaN input.
    got = run("scale", B(r.s), a(-2.0), X)
    np.testing.assert_array_equal(got, x * npdt(-2.0))
    got = run("scale", B(r.s), a(0.0), NAN)
    np.testing.assert_array_equal(got, np.zeros(n))
    # axpby special cases (cldense-axpby.cpp L27-84).
    cases = [(0.0, 0.0, np.zeros(n)), (0.0, 3.0, y * npdt(3.0)), (2.0, 0.0, x * npdt(2.0)),
             (1.0, -0.25, x + npdt(-0.25) * y), (-1.5, 1.0, npdt(-1.5) * x + y),
             (0.7, -0.3, npdt(0.7) * x + npdt(-0.3) * y)]
    for al, be, want in cases:
        got = run("axpby", B(r.s), a(al), X, a(be), Y)
        # FMA contraction may round once instead of twice: bound by the terms' size.
        mag = np.abs(npdt(al) * x.astype(np.float64)) + np.abs(npdt(be) * y.astype(np.float64))
        assert np.all(np.abs(got.astype(np.float64) - want) <= 4 * np.finfo(npdt).eps * mag)
    got = run("axpby", B(r.s), a(0.0), NAN, a(0.0), NAN)
    np.testing.assert_array_equal(got, np.zeros(n))


@pytest.mark.parametrize("prec", ["S", "D"])
def test_elementwise_exact(ctl, prec):
    lib = cls.load()
    dt = torch.float32 if prec == "S" else torch.float64
    npdt = np.float32 if prec == "S" else np.float64
    n = 3000
    x = cls.gen_uniform(n, 0.5, 2, seed=5).astype(npdt)
    y = cls.gen_uniform(n, 0.5, 2, seed=6).astype(npdt)
    r = _vec(np.zeros(n), dt)
    X, Y = _vec(x, dt), _vec(y, dt)
    for op, want in (("add", x + y), ("sub", x - y), ("mul", x * y), ("div", x / y)):
        assert getattr(lib, f"cldense{prec}{op}")(B(r.s), B(X.s), B(Y.s), ctl.handle) == 0
        np.testing.assert_array_equal(r.t.cpu().numpy(), want)


def test_reduction_deterministic_and_async(ctl):
    lib = cls.load()
    x = _vec(cls.gen_uniform(3_000_000, -1, 1, seed=8), torch.float64)
    outs = []
    with cls.Control(async_mode=True) as ac:
        for _ in range(3):
            o = cls.Scalar(0.0)
            assert lib.cldenseDnrm2(B(o.s), B(x.s), ac.handle) == 0
            ev = lib.clsparseGetEvent(ac.handle)
            assert ev.status == 0 and ev.event
            torch.cuda.synchronize()
            outs.append(o.item())
    assert outs[0] == outs[1] == outs[2]
