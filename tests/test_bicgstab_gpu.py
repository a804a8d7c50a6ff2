"""BiCGStab (clsparse[S/D]csrbicgStab) against the CPU restatement of
src/library/solvers/biconjugate-gradients-stabilized.hpp L41-268."""
import ctypes

import numpy as np
import pytest
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


def _convection_diffusion(nx, ny, c=0.4):
    """5-point Laplacian plus an upwind convection term in x: nonsymmetric."""
    n = nx * ny
    lap = cls.gen_poisson2d(nx, ny)
    a = sp.csr_matrix((lap[2], lap[1], lap[0]), shape=(n, n)).tolil()
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            a[i, i] += c
            if x > 0:
                a[i, i - 1] -= c
    a = sp.csr_matrix(a)
    a.sort_indices()
    return a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, n


def _order_window(rp, ci, v, b, precond, tol, dtype):
    """The iteration counts of the same BiCGStab recurrence (right
    preconditioning, the two L1 checks of biconjugate-gradients-stabilized.hpp)
    in numpy with the dot products summed in three orders (sequential,
    numpy's pairwise, reversed) -- the spread that summation order alone gives
    on this system. Vectors in `dtype`, dots and norms in FP64, as the library
    forms them; in FP32 also with the SpMV rows summed in FP32 and in FP64
    (the CSR kernels sum in FP64, the diagonal-storage ones in FP32)."""
    n = rp.size - 1
    a = sp.csr_matrix((np.asarray(v, dtype=dtype), ci, rp), shape=(n, n))
    a64 = sp.csr_matrix((np.asarray(v, dtype=dtype).astype(np.float64), ci, rp), shape=(n, n))
    inv = (1 / a.diagonal()).astype(dtype) if precond == cls.DIAGONAL else np.ones(n, dtype)
    bb = np.asarray(b, dtype=dtype)
    nb = np.abs(bb.astype(np.float64)).sum()

    def run(dot, wide=False):
        mv = (lambda t: (a64 @ t.astype(np.float64)).astype(dtype)) if wide else (lambda t: a @ t)
        r = bb.copy()
        p, rs = r.copy(), r.copy()
        rr = dot(rs, r)
        for it in range(5000):
            mp = p * inv
            amp = mv(mp)
            alpha = dtype(rr / dot(rs, amp))
            s_ = r - alpha * amp
            if np.abs(s_.astype(np.float64)).sum() / nb <= tol:
                return it
            ms = s_ * inv
            ams = mv(ms)
            omega = dtype(dot(ams, s_) / dot(ams, ams))
            r = s_ - omega * ams
            if np.abs(r.astype(np.float64)).sum() / nb <= tol:
                return it
            rrn = dot(rs, r)
            beta = dtype(rrn / rr * float(alpha) / float(omega))
            rr = rrn
            p = beta * p + r - beta * omega * amp
        return 5000

    def seq(u, w):
        acc = 0.0
        for t in (u.astype(np.float64) * w.astype(np.float64)).tolist():
            acc += t
        return acc

    dots = [seq, lambda u, w: float(np.dot(u.astype(np.float64), w.astype(np.float64))),
            lambda u, w: seq(u[::-1], w[::-1])]
    counts = [run(d) for d in dots]
    if dtype == np.float32:
        counts += [run(d, wide=True) for d in dots]
    return min(counts), max(counts)


def _solve(ctl, rp, ci, v, n, b, dtype, precond, max_iters, rel, absol, n_prior=None):
    lib = cls.load()
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    A = cls.CsrMatrix(rp, ci, np.asarray(v, dtype=dtype), n, dtype=dtype)
    A.meta_create(ctl)
    x = torch.zeros(n, dtype=tdt, device="cuda")
    xd, bd = cls.DenseVector(x), cls.DenseVector(torch.tensor(np.asarray(b, dtype=dtype), device="cuda"))
    sc = cls.SolverControl(precond, max_iters, rel, absol)
    try:
        fn = lib.clsparseScsrbicgStab if dtype == np.float32 else lib.clsparseDcsrbicgStab
        st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle, ctl.handle)
        return st, sc.stats(), x.cpu().numpy()
    finally:
        sc.release()


@pytest.mark.parametrize("precond", [cls.DIAGONAL, cls.NOPRECOND])
def test_nonsymmetric_fp64_matches_oracle(ctl, precond):
    rp, ci, v, n = _convection_diffusion(48, 40)
    b = cls.gen_uniform(n, -1, 1, seed=21)
    say(f"bicgstab fp64 precond={precond}")
    st, s, x = _solve(ctl, rp, ci, v, n, b, np.float64, precond, 2000, 1e-10, 0.0)
    assert st == 0
    xo, ro = oracle.bicgstab(rp, ci, v, b, np.zeros(n), precond, 2000, 1e-10, 0.0)
    # BiCGStab's residual history is erratic: the iteration count moves with
    # the order in which the dot products are summed alone, so the bar is the
    # spread of the same recurrence over three summation orders on this
    # system (and the oracle's own count), one iteration either side.
    lo, hi = _order_window(rp, ci, v, b, precond, 1e-10, np.float64)
    lo, hi = min(lo, ro.n_iters), max(hi, ro.n_iters)
    say(f"  iterations gpu {s.n_iters}, oracle {ro.n_iters}, summation-order window [{lo}, {hi}]")
    assert lo - 1 <= s.n_iters <= hi + 1, (s.n_iters, lo, hi)
    r = b - oracle.csrmv(rp, ci, v, x)[0]
    assert np.abs(r).sum() / np.abs(b).sum() <= 1e-10 * 1.5
    assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()


def test_fp32_converges(ctl):
    rp, ci, v, n = _convection_diffusion(32, 32)
    b = cls.gen_uniform(n, -1, 1, seed=5)
    st, s, x = _solve(ctl, rp, ci, v, n, b, np.float32, cls.DIAGONAL, 1000, 1e-5, 0.0)
    assert st == 0 and s.status_detail == 0
    _, ro = oracle.bicgstab(rp, ci, v.astype(np.float32), b.astype(np.float32), np.zeros(n, np.float32),
                            1, 1000, 1e-5, 0.0)
    lo, hi = _order_window(rp, ci, v, b, cls.DIAGONAL, 1e-5, np.float32)
    lo, hi = min(lo, ro.n_iters), max(hi, ro.n_iters)
    say(f"  fp32 iterations gpu {s.n_iters}, oracle {ro.n_iters}, summation-order window [{lo}, {hi}]")
    assert lo - 1 <= s.n_iters <= hi + 1, (s.n_iters, lo, hi)
    r = b - oracle.csrmv(rp, ci, v, x.astype(np.float64))[0]
    assert np.abs(r).sum() / np.abs(b).sum() <= 2e-5


def test_diagonal_matrix_stops_at_first_check_with_exact_solution(ctl):
    d = np.array([2.0, 5.0, 0.25, 9.0, 3.0])
    a = sp.diags(d).tocsr()
    b = np.array([1.0, -2.0, 3.0, 0.5, 7.0])
    st, s, x = _solve(ctl, a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, 5, b,
                      np.float64, cls.DIAGONAL, 100, 1e-12, 0.0)
    assert st == 0 and s.n_iters == 0
    np.testing.assert_allclose(x, b / d, rtol=1e-15)


def test_zero_rhs_and_max_iters(ctl):
    rp, ci, v, n = _convection_diffusion(20, 20)
    st, s, x = _solve(ctl, rp, ci, v, n, np.zeros(n), np.float64, cls.DIAGONAL, 100, 1e-10, 0.0)
    assert st == 0 and s.n_iters == 0 and np.all(x == 0)
    assert s.relative_tolerance == 0.0 and s.absolute_tolerance == 0.0
    b = cls.gen_uniform(n, -1, 1, seed=8)
    st, s, x = _solve(ctl, rp, ci, v, n, b, np.float64, cls.DIAGONAL, 5, 1e-14, 0.0)
    xo, ro = oracle.bicgstab(rp, ci, v, b, np.zeros(n), 1, 5, 1e-14, 0.0)
    assert st == 0 and s.n_iters == ro.n_iters == 5 and s.status_detail == 1
    np.testing.assert_allclose(x, xo, rtol=1e-9, atol=1e-12)


def test_error_codes(ctl):
    lib = cls.load()
    x = cls.DenseVector(torch.zeros(4, dtype=torch.float64, device="cuda"))
    a = sp.identity(3, format="csr")
    A = cls.CsrMatrix(a.indptr, a.indices, a.data, 3)
    sc = cls.SolverControl(cls.DIAGONAL, 10, 1e-6, 0.0)
    try:
        assert lib.clsparseDcsrbicgStab(ctypes.byref(x.s), ctypes.byref(A.s), ctypes.byref(x.s),
                                        sc.handle, None) == cls.clsparseInvalidControlObject
        assert lib.clsparseDcsrbicgStab(ctypes.byref(x.s), ctypes.byref(A.s), ctypes.byref(x.s),
                                        None, ctl.handle) == cls.clsparseInvalidSolverControlObject
        assert lib.clsparseDcsrbicgStab(ctypes.byref(x.s), ctypes.byref(A.s), ctypes.byref(x.s),
                                        sc.handle, ctl.handle) == cls.clsparseInvalidSystemSize
    finally:
        sc.release()


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_diagonal_storage_matches_csr(ctl, monkeypatch, dtype):
    """The nonsymmetric 5-point matrix runs from full diagonal storage (its
    mirrored values differ); the solve with CLSPARSE_AMD_DIA=0 (CSR-Adaptive)
    stops within a few iterations of it and both meet the tolerance."""
    rp, ci, v, n = _convection_diffusion(72, 64)
    b = cls.gen_uniform(n, -1, 1, seed=33)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    say(f"bicgstab diagonal storage vs CSR {np.dtype(dtype).name}")
    st_d, s_d, x_d = _solve(ctl, rp, ci, v, n, b, dtype, cls.DIAGONAL, 5000, tol, 0.0)
    monkeypatch.setenv("CLSPARSE_AMD_DIA", "0")
    st_c, s_c, x_c = _solve(ctl, rp, ci, v, n, b, dtype, cls.DIAGONAL, 5000, tol, 0.0)
    assert st_d == 0 and st_c == 0 and s_d.status_detail == 0 and s_c.status_detail == 0
    assert s_d.a_diagonals == 5 and s_c.a_diagonals == 0
    assert abs(s_d.n_iters - s_c.n_iters) <= max(3, 0.1 * s_c.n_iters)
    for x in (x_d, x_c):
        r = b - oracle.csrmv(rp, ci, v, x.astype(np.float64))[0]
        assert np.abs(r).sum() / np.abs(b).sum() <= tol * (1.5 if dtype == np.float64 else 3.0)
