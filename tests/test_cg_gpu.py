"""clsparseScsrcg / clsparseDcsrcg on the GPU against the oracle's restatement of
conjugate-gradients.hpp L47-229 and the closed forms it must reproduce.

Parity bar: FP64 -- nIters equal to the oracle's within 1 and the final L1
residual below the tolerance; FP32 -- nIters within 3 (reduction order differs).
"""
import ctypes

import numpy as np
import pytest
import torch

import clsparse_amd as cls
import oracle
from gpu_helpers import cg_call, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


def _solve(ctl, rp, ci, v, m, b, x0, dtype, precond, max_iters, rel, abst, meta=True):
    dt = torch.float32 if dtype == np.float32 else torch.float64
    A = cls.CsrMatrix(rp, ci, v, m, dtype=dtype)
    if meta:
        A.meta_create(ctl)
    x = torch.tensor(np.asarray(x0, dtype=dtype), device="cuda")
    bt = torch.tensor(np.asarray(b, dtype=dtype), device="cuda")
    sc = cls.SolverControl(precond, max_iters, rel, abst)
    st = cg_call(A, x, bt, sc, ctl)
    stats = sc.stats()
    sc.release()
    return st, x.cpu().numpy(), stats


@pytest.mark.parametrize("dia", ["1", "0"], ids=["diagonal-storage", "csr"])
@pytest.mark.parametrize("meta", [True, False], ids=["adaptive", "vector"])
@pytest.mark.parametrize("precond", [cls.DIAGONAL, cls.NOPRECOND], ids=["jacobi", "none"])
def test_poisson_fp64_matches_oracle(ctl, monkeypatch, precond, meta, dia):
    """The 5-point Poisson matrix runs the iterations from diagonal storage by
    default; CLSPARSE_AMD_DIA=0 keeps the CSR SpMV (CSR-Adaptive with
    metadata, CSR-Vector without)."""
    monkeypatch.setenv("CLSPARSE_AMD_DIA", dia)
    say(f"cg poisson fp64 precond={precond} meta={meta} dia={dia}")
    rp, ci, v, m = cls.gen_poisson2d(64, 64)
    b = cls.gen_uniform(m, -1, 1, seed=3)
    st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float64, precond, 2000, 1e-10, 0.0,
                      meta)
    assert st == 0
    assert s.a_diagonals == (5 if dia == "1" else 0)
    xr, r = oracle.cg(rp, ci, v, b, np.zeros(m), precond, 2000, 1e-10, 0.0)
    assert abs(s.n_iters - r.n_iters) <= 1, (s.n_iters, r.n_iters)
    assert s.current_residual <= 1e-10 and s.iterations == s.n_iters + 1
    assert s.initial_residual == pytest.approx(r.initial_residual, rel=1e-13)
    np.testing.assert_allclose(x, xr, atol=1e-8 * np.abs(xr).max())


def test_poisson_fp32_matches_oracle(ctl):
    say("cg poisson fp32")
    rp, ci, v, m = cls.gen_poisson2d(48, 48)
    b = np.ones(m)
    st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float32, cls.DIAGONAL, 1000, 1e-5,
                      0.0)
    assert st == 0
    xr, r = oracle.cg(rp, ci, v.astype(np.float32), b.astype(np.float32),
                      np.zeros(m, np.float32), 1, 1000, 1e-5, 0.0)
    assert abs(s.n_iters - r.n_iters) <= 3, (s.n_iters, r.n_iters)
    assert s.current_residual <= 1e-5
    np.testing.assert_allclose(x, xr, rtol=0, atol=2e-4 * np.abs(xr).max())


def test_reference_defaults(ctl):
    """test-solvers.cpp defaults: relTol 1e-2, absTol 1e-5, maxIters 100, Diagonal, x0 = 0, b = 1."""
    rp, ci, v, m = cls.gen_poisson2d(32, 32)
    for dtype in (np.float32, np.float64):
        st, x, s = _solve(ctl, rp, ci, v, m, np.ones(m), np.zeros(m), dtype, cls.DIAGONAL, 100,
                          1e-2, 1e-5)
        _, r = oracle.cg(rp, ci, v.astype(dtype), np.ones(m, dtype), np.zeros(m, dtype), 1, 100,
                         1e-2, 1e-5)
        assert st == 0 and abs(s.n_iters - r.n_iters) <= 1


def test_diagonal_matrix_one_iteration(ctl):
    n = 1000
    d = np.linspace(0.5, 20, n)
    rp, ci = np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32)
    b = cls.gen_uniform(n, -1, 1, seed=1)
    st, x, s = _solve(ctl, rp, ci, d, n, b, np.zeros(n), np.float64, cls.DIAGONAL, 50, 1e-12, 0.0)
    assert st == 0 and s.n_iters == 0 and s.iterations == 1
    np.testing.assert_allclose(x, b / d, rtol=1e-15)


def test_three_distinct_eigenvalues(ctl):
    n = 48
    d = np.array([1.0, 3.0, 7.0] * 16)
    q = np.linalg.qr(np.random.default_rng(0).normal(size=(n, n)))[0]
    a = (q * d) @ q.T
    a = 0.5 * (a + a.T)
    rp = np.arange(0, n * n + 1, n, dtype=np.uint32)
    ci = np.tile(np.arange(n, dtype=np.uint32), n)
    b = cls.gen_uniform(n, -1, 1, seed=2)
    st, x, s = _solve(ctl, rp, ci, a.reshape(-1), n, b, np.zeros(n), np.float64, cls.NOPRECOND,
                      50, 1e-12, 0.0)
    assert st == 0 and s.iterations <= 3
    np.testing.assert_allclose(a @ x, b, atol=1e-10)


def test_zero_rhs_returns_b(ctl):
    rp, ci, v, m = cls.gen_poisson2d(16, 16)
    st, x, s = _solve(ctl, rp, ci, v, m, np.zeros(m), np.full(m, 5.0), np.float64, cls.DIAGONAL,
                      10, 1e-2, 1e-5)
    assert st == 0 and np.all(x == 0) and s.n_iters == 0
    assert s.relative_tolerance == 0 and s.absolute_tolerance == 0


def test_max_iters_and_stored_niters_quirks(ctl):
    rp, ci, v, m = cls.gen_poisson2d(32, 32)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    b = torch.ones(m, dtype=torch.float64, device="cuda")
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    sc = cls.SolverControl(cls.DIAGONAL, 3, 0.0, 0.0)
    assert cg_call(A, x, b, sc, ctl) == 0
    s = sc.stats()
    assert s.n_iters == 3 and s.iterations == 4  # maxIters + 1 iterations run
    xr, r = oracle.cg(rp, ci, v, np.ones(m), np.zeros(m), 1, 3, 0.0, 0.0)
    np.testing.assert_allclose(x.cpu().numpy(), xr, rtol=1e-12)
    # Same solver control again: stored nIters (3) >= maxIters returns at once.
    x2 = torch.zeros(m, dtype=torch.float64, device="cuda")
    assert cg_call(A, x2, b, sc, ctl) == 0
    assert sc.stats().n_iters == 0 and torch.count_nonzero(x2).item() == 0
    sc.release()


def test_error_codes(ctl):
    lib = cls.load()
    rp, ci, v, m = cls.gen_poisson2d(8, 8)
    A = cls.CsrMatrix(rp, ci, v, m)
    x = cls.DenseVector(torch.zeros(m - 1, dtype=torch.float64, device="cuda"))
    b = cls.DenseVector(torch.ones(m, dtype=torch.float64, device="cuda"))
    sc = cls.SolverControl()
    B = ctypes.byref
    assert lib.clsparseDcsrcg(B(x.s), B(A.s), B(b.s), None, ctl.handle) == \
        cls.clsparseInvalidSolverControlObject
    assert lib.clsparseDcsrcg(B(x.s), B(A.s), B(b.s), sc.handle, None) == \
        cls.clsparseInvalidControlObject
    assert lib.clsparseDcsrcg(B(x.s), B(A.s), B(b.s), sc.handle, ctl.handle) == \
        cls.clsparseInvalidSystemSize
    sc.release()


def test_verbose_prints_every_iteration(ctl, capfd):
    rp, ci, v, m = cls.gen_poisson2d(16, 16)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    b = torch.ones(m, dtype=torch.float64, device="cuda")
    sc = cls.SolverControl(cls.DIAGONAL, 500, 1e-8, 0.0, print_mode=cls.VERBOSE)
    assert cg_call(A, x, b, sc, ctl) == 0
    out = capfd.readouterr().out
    n_iters = sc.stats().n_iters
    lines = [ln for ln in out.splitlines() if ln.startswith("Iteration: ")]
    assert len(lines) == n_iters + 1
    assert lines[-1].startswith(f"Iteration: {n_iters} ")
    assert "Solver finished calculations with status 0" in out
    sc.release()


def test_graph_and_stream_paths_identical(ctl, monkeypatch):
    rp, ci, v, m = cls.gen_diffusion3d(24, 24, 24, sigma=1.0, seed=4)
    b = cls.gen_uniform(m, -1, 1, seed=5)
    res = []
    for flag in ("0", "1"):
        monkeypatch.setenv("CLSPARSE_AMD_NO_GRAPH", flag)
        st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float64, cls.DIAGONAL, 3000,
                          1e-9, 0.0)
        assert st == 0
        res.append((x, s.n_iters))
    assert res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-11), (np.float32, 1e-4)])
def test_line_order_matches_grid_stride_and_csr(ctl, monkeypatch, dtype, rtol):
    """A 64 x 64 x 16 grid (largest offset 4096, 16 lines): the diagonal-storage
    SpMV walks rows in line order (the symmetric FP64 and FP32 kernels carry
    the far operands in registers, the fused FP32 Jacobi SpMV forms p' per
    row). Row sums are the same as in grid-stride order and in the CSR SpMV;
    only the <y, p> partials are added in another order, so six iterations
    agree to rounding and the full solves take the same iterations."""
    say(f"cg line order 64x64x16 {np.dtype(dtype).name}")
    rp, ci, v, m = cls.gen_diffusion3d(64, 64, 16, sigma=1.0, ax=1.0, ay=1.0, az=10.0, seed=23)
    b = cls.gen_uniform(m, -1, 1, seed=24)
    short, full = [], []
    for dia, lines in (("1", "1"), ("1", "0"), ("0", "1")):
        monkeypatch.setenv("CLSPARSE_AMD_DIA", dia)
        monkeypatch.setenv("CLSPARSE_AMD_DIA_LINES", lines)
        st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), dtype, cls.DIAGONAL, 6, 1e-30, 0.0)
        assert 0 < s.n_iters <= 6
        short.append((s.n_iters, np.asarray(x, dtype=np.float64)))
        tol = 1e-9 if dtype == np.float64 else 1e-5
        st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), dtype, cls.DIAGONAL, 5000, tol, 0.0)
        assert st == 0
        full.append(s.n_iters)
    assert short[0][0] == short[1][0] == short[2][0]
    scale = np.abs(short[1][1]).max()
    for _, x in (short[0], short[2]):
        assert np.allclose(x, short[1][1], rtol=rtol, atol=rtol * scale)
    assert max(full) - min(full) <= 2, full


def test_diffusion_fp64_true_residual(ctl):
    say("cg diffusion 32^3")
    rp, ci, v, m = cls.gen_diffusion3d(32, 32, 32, sigma=1.0, seed=6)
    xs = cls.gen_uniform(m, -1, 1, seed=7)
    b, _ = oracle.csrmv(rp, ci, v, xs)
    st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float64, cls.DIAGONAL, 5000, 1e-11,
                      0.0)
    assert st == 0
    r = b - oracle.csrmv(rp, ci, v, x)[0]
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-9
    _, ro = oracle.cg(rp, ci, v, b, np.zeros(m), 1, 5000, 1e-11, 0.0)
    assert abs(s.n_iters - ro.n_iters) <= 2, (s.n_iters, ro.n_iters)


def test_mixed_precision_refinement(ctl):
    """FP32 inner Jacobi-PCG + FP64 outer residual reaches an FP64 true residual of 1e-10."""
    say("mixed precision IR")
    rp, ci, v, m = cls.gen_diffusion3d(32, 32, 32, sigma=1.0, seed=8)
    b = cls.gen_uniform(m, -1, 1, seed=9)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    bt = torch.tensor(b, device="cuda")
    sc = cls.SolverControl(cls.DIAGONAL, 20000, 1e-10, 0.0)
    lib = cls.load()
    xd, bd = cls.DenseVector(x), cls.DenseVector(bt)
    assert lib.clsparseAmdDcsrcgMixed(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s),
                                      sc.handle, ctl.handle) == 0
    s = sc.stats()
    xh = x.cpu().numpy()
    r = b - oracle.csrmv(rp, ci, v, xh)[0]
    true_rel = np.linalg.norm(r) / np.linalg.norm(b)
    assert true_rel <= 1e-10 and s.true_residual <= 1e-10
    assert s.true_residual == pytest.approx(true_rel, rel=1e-3)
    assert s.outer_iterations >= 2
    sc.release()


def test_mixed_reliable_updates_track_fp64_iterations(ctl):
    """Residual replacement keeps the FP32 Krylov run going: the mixed solver
    needs about as many iterations as FP64 PCG for the same FP64 true residual."""
    say("mixed vs FP64 iteration counts, diffusion 40^3")
    rp, ci, v, m = cls.gen_diffusion3d(40, 40, 40, sigma=1.0, ax=1.0, ay=1.0, az=30.0, seed=11)
    b = cls.gen_uniform(m, -1, 1, seed=12)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    lib = cls.load()
    bt = torch.tensor(b, device="cuda")
    res = {}
    for name, fn in (("mixed", lib.clsparseAmdDcsrcgMixed), ("fp64", lib.clsparseAmdDcsrcgRefined)):
        x = torch.zeros(m, dtype=torch.float64, device="cuda")
        sc = cls.SolverControl(cls.DIAGONAL, 20000, 1e-9, 0.0)
        xd, bd = cls.DenseVector(x), cls.DenseVector(bt)
        assert fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle,
                  ctl.handle) == 0
        s = sc.stats()
        r = b - oracle.csrmv(rp, ci, v, x.cpu().numpy())[0]
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-9
        assert s.status_detail == 0 and s.preconditioner == 1
        res[name] = s.iterations
        sc.release()
    say(f"iterations mixed {res['mixed']} fp64 {res['fp64']}")
    assert res["mixed"] <= 1.15 * res["fp64"] + 32


def _mixed_solve(ctl, rp, ci, v, m, b, tol=1e-9):
    lib = cls.load()
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    sc = cls.SolverControl(cls.DIAGONAL, 20000, tol, 0.0)
    xd, bd = cls.DenseVector(x), cls.DenseVector(torch.tensor(b, device="cuda"))
    try:
        assert lib.clsparseAmdDcsrcgMixed(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s),
                                          sc.handle, ctl.handle) == 0
        return sc.stats(), x.cpu().numpy()
    finally:
        sc.release()
        A.meta_delete()


@pytest.mark.parametrize("case", ["diffusion3d", "diffusion3d_pow2", "poisson2d", "q1_2d"])
def test_mixed_diagonal_storage_matches_csr(ctl, monkeypatch, case):
    """Stencil matrices go to diagonal storage in the FP32 inner solver (7, 5
    and 9 diagonals). The row sums are formed in the same order as the CSR
    SpMV, so the solve takes the same iterations as with CSR. (32 x 32 x 16:
    rows sampled at a regular stride n / 64 would all sit on the x = 0
    boundary and miss the -1 diagonal.)"""
    if case == "diffusion3d":
        rp, ci, v, m = cls.gen_diffusion3d(28, 26, 24, sigma=1.0, ax=1.0, ay=1.0, az=20.0, seed=17)
        nd = 7
    elif case == "diffusion3d_pow2":
        rp, ci, v, m = cls.gen_diffusion3d(32, 32, 16, sigma=1.0, ax=1.0, ay=1.0, az=20.0, seed=18)
        nd = 7
    elif case == "poisson2d":
        rp, ci, v, m = cls.gen_poisson2d(120, 110)
        nd = 5
    else:
        rp, ci, v, m = cls.gen_aniso2d(90, 80, corr_len=6.0, sigma=1.0, ratio=30.0, theta=0.4, seed=2)
        nd = 9
    b = cls.gen_uniform(m, -1, 1, seed=23)
    say(f"mixed CG diagonal storage vs CSR: {case}")
    st_d, x_d = _mixed_solve(ctl, rp, ci, v, m, b)
    monkeypatch.setenv("CLSPARSE_AMD_DIA", "0")
    st_c, x_c = _mixed_solve(ctl, rp, ci, v, m, b)
    assert st_d.a_diagonals == nd and st_c.a_diagonals == 0
    assert st_d.status_detail == 0 and st_c.status_detail == 0
    assert abs(st_d.iterations - st_c.iterations) <= max(2, st_c.iterations // 50)
    for x in (x_d, x_c):
        r = b - oracle.csrmv(rp, ci, v, x)[0]
        assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-9


def test_mixed_keeps_csr_for_entries_off_the_diagonals(ctl):
    """One symmetric pair of entries far from the stencil's diagonals: the
    fill kernel finds it and the solve runs on the CSR SpMV."""
    import scipy.sparse as sp
    rp, ci, v, m = cls.gen_poisson2d(60, 50)
    a = sp.csr_matrix((v, ci, rp), shape=(m, m)).tolil()
    a[3, m - 7] = -0.05
    a[m - 7, 3] = -0.05
    a = sp.csr_matrix(a)
    a.sort_indices()
    rp2, ci2 = a.indptr.astype(np.uint32), a.indices.astype(np.uint32)
    b = cls.gen_uniform(m, -1, 1, seed=24)
    say("mixed CG: stencil with a far entry -> CSR")
    st, x = _mixed_solve(ctl, rp2, ci2, a.data, m, b)
    assert st.a_diagonals == 0 and st.status_detail == 0
    r = b - oracle.csrmv(rp2, ci2, a.data, x)[0]
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-9


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_reference_cg_diagonal_storage_matches_csr(ctl, monkeypatch, dtype):
    """clsparse[S/D]csrcg on a 3-D stencil: the iterations from diagonal storage
    and from the CSR-Adaptive SpMV (same row sums) stop at the same iteration
    count with the same solution to rounding."""
    rp, ci, v, m = cls.gen_diffusion3d(30, 28, 26, sigma=1.0, ax=1.0, ay=1.0, az=10.0, seed=19)
    b = cls.gen_uniform(m, -1, 1, seed=29)
    tol = 1e-5 if dtype == np.float32 else 1e-10
    say(f"reference CG diagonal storage vs CSR {dtype.__name__}")
    st_d, x_d, s_d = _solve(ctl, rp, ci, v, m, b, np.zeros(m), dtype, cls.DIAGONAL, 5000, tol, 0.0)
    monkeypatch.setenv("CLSPARSE_AMD_DIA", "0")
    st_c, x_c, s_c = _solve(ctl, rp, ci, v, m, b, np.zeros(m), dtype, cls.DIAGONAL, 5000, tol, 0.0)
    assert st_d == 0 and st_c == 0
    assert s_d.a_diagonals == 7 and s_c.a_diagonals == 0
    assert abs(s_d.n_iters - s_c.n_iters) <= (2 if dtype == np.float32 else 1)
    scale = np.abs(x_c).max()
    assert np.abs(x_d - x_c).max() <= (1e-3 if dtype == np.float32 else 1e-8) * scale


def test_reference_cg_extended_precision_keeps_csr(ctl):
    """Extended precision needs the compensated CSR sums: no diagonal storage."""
    rp, ci, v, m = cls.gen_poisson2d(40, 40)
    b = cls.gen_uniform(m, -1, 1, seed=31)
    ext = cls.Control(extended_precision=True)
    try:
        st, x, s = _solve(ext, rp, ci, v, m, b, np.zeros(m), np.float64, cls.DIAGONAL, 2000, 1e-10, 0.0)
        assert st == 0 and s.a_diagonals == 0
        _, r = oracle.cg(rp, ci, v, b, np.zeros(m), 1, 2000, 1e-10, 0.0)
        assert abs(s.n_iters - r.n_iters) <= 1
    finally:
        ext.release()


def test_symmetric_half_storage_is_bitwise_the_full_one(ctl, monkeypatch):
    """A symmetric stencil reads only its lower diagonals (the upper values come
    from the mirrored positions, checked equal): the solve is bit for bit the
    one from full diagonal storage."""
    rp, ci, v, m = cls.gen_diffusion3d(26, 24, 22, sigma=1.0, ax=1.0, ay=1.0, az=5.0, seed=37)
    b = cls.gen_uniform(m, -1, 1, seed=41)
    say("reference CG: symmetric half storage vs full")
    st_h, x_h, s_h = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float64, cls.DIAGONAL, 5000, 1e-10, 0.0)
    monkeypatch.setenv("CLSPARSE_AMD_DIA_SYM", "0")
    st_f, x_f, s_f = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float64, cls.DIAGONAL, 5000, 1e-10, 0.0)
    assert st_h == 0 and st_f == 0 and s_h.a_diagonals == 7 and s_f.a_diagonals == 7
    assert s_h.n_iters == s_f.n_iters
    assert np.array_equal(x_h, x_f)


def test_nonsymmetric_stencil_uses_full_storage(ctl, monkeypatch):
    """Upwind convection makes the 5-point matrix nonsymmetric: the mirrored
    values differ, the full diagonal storage is used, and 20 CG iterations
    give the CSR result to rounding."""
    import scipy.sparse as sp
    nx = ny = 40
    lap = cls.gen_poisson2d(nx, ny)
    n = nx * ny
    a = sp.csr_matrix((lap[2], lap[1], lap[0]), shape=(n, n)).tolil()
    for i in range(n):
        a[i, i] += 0.3
        if i % nx > 0:
            a[i, i - 1] -= 0.3
    a = sp.csr_matrix(a)
    a.sort_indices()
    rp, ci, v = a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data
    b = cls.gen_uniform(n, -1, 1, seed=43)
    say("reference CG: nonsymmetric stencil, 20 iterations, diagonal vs CSR")
    _, x_d, s_d = _solve(ctl, rp, ci, v, n, b, np.zeros(n), np.float64, cls.DIAGONAL, 20, 1e-30, 0.0)
    monkeypatch.setenv("CLSPARSE_AMD_DIA", "0")
    _, x_c, s_c = _solve(ctl, rp, ci, v, n, b, np.zeros(n), np.float64, cls.DIAGONAL, 20, 1e-30, 0.0)
    assert s_d.a_diagonals == 5 and s_c.a_diagonals == 0
    assert np.abs(x_d - x_c).max() <= 1e-10 * np.abs(x_c).max()


@pytest.mark.parametrize("solver", ["cg", "bicgstab", "mixed"])
def test_offsets_with_diagonal_storage(ctl, solver):
    """Matrix arrays, x and b with element offsets (garbage in front): the
    diagonal-storage detection and fill read through the offsets, and the
    solve equals the one on offset-free arrays."""
    import ctypes
    rp, ci, v, m = cls.gen_poisson2d(48, 40)
    ov, oc, orp, ox, ob = 5, 9, 4, 7, 3
    dev = "cuda"
    vals = torch.cat([torch.full((ov,), 99.0, dtype=torch.float64), torch.tensor(v)]).to(dev)
    cols = torch.cat([torch.full((oc,), 12345, dtype=torch.int32), torch.tensor(ci.view(np.int32))]).to(dev)
    rptr = torch.cat([torch.full((orp,), 77, dtype=torch.int32), torch.tensor(rp.view(np.int32))]).to(dev)
    A = cls.clsparseCsrMatrix()
    lib = cls.load()
    lib.clsparseInitCsrMatrix(ctypes.byref(A))
    A.num_rows, A.num_cols, A.num_nonzeros = m, m, v.size
    A.values, A.col_indices, A.row_pointer = vals.data_ptr(), cols.data_ptr(), rptr.data_ptr()
    A.off_values, A.off_col_indices, A.off_row_pointer = ov, oc, orp
    assert lib.clsparseCsrMetaCreate(ctypes.byref(A), ctl.handle) == 0
    b = cls.gen_uniform(m, -1, 1, seed=51)
    say(f"offsets with diagonal storage: {solver}")

    def run(Amat, x_off, b_off):
        xt = torch.cat([torch.full((x_off,), 1e30, dtype=torch.float64), torch.zeros(m, dtype=torch.float64)]).to(dev)
        bt = torch.cat([torch.full((b_off,), -7.0, dtype=torch.float64), torch.tensor(b)]).to(dev)
        xd = cls.cldenseVector(m, xt.data_ptr(), x_off)
        bd = cls.cldenseVector(m, bt.data_ptr(), b_off)
        sc = cls.SolverControl(cls.DIAGONAL, 3000, 1e-10, 0.0)
        fn = {"cg": lib.clsparseDcsrcg, "bicgstab": lib.clsparseDcsrbicgStab,
              "mixed": lib.clsparseAmdDcsrcgMixed}[solver]
        assert fn(ctypes.byref(xd), ctypes.byref(Amat), ctypes.byref(bd), sc.handle, ctl.handle) == 0
        s = sc.stats()
        sc.release()
        out = xt.cpu().numpy()
        assert np.all(out[:x_off] == 1e30)
        return out[x_off:], s
    x1, s1 = run(A, ox, ob)
    B = cls.CsrMatrix(rp, ci, v, m)
    B.meta_create(ctl)
    x2, s2 = run(B.s, 0, 0)
    assert s1.a_diagonals == 5 and s2.a_diagonals == 5
    assert s1.iterations == s2.iterations
    assert np.array_equal(x1, x2)
    lib.clsparseCsrMetaDelete(ctypes.byref(A))


@pytest.mark.parametrize("solver", ["reference_fp32", "mixed"])
def test_fused_p_update_matches_separate_kernels(ctl, monkeypatch, solver):
    """The FP32 Jacobi iteration with A in diagonal storage folds the p update
    into the next SpMV (p' = z + beta p formed at every row it reads, stored by
    the x / r update; the chunk's last convergence test in a kernel of its
    own; p brought up to date before a residual replacement). It must give
    the iterates of the separate p update kernel bit for bit."""
    rp, ci, v, m = cls.gen_diffusion3d(28, 24, 20, sigma=1.0, ax=1.0, ay=1.0, az=20.0, seed=17)
    b = cls.gen_uniform(m, -1, 1, seed=18)
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("CLSPARSE_AMD_CG_FUSE", fuse)
        say(f"cg fused p update={fuse} solver={solver}")
        if solver == "reference_fp32":
            st, x, s = _solve(ctl, rp, ci, v, m, b, np.zeros(m), np.float32, cls.DIAGONAL, 5000, 1e-6, 0.0)
            assert st == 0 and s.a_diagonals == 7
            out[fuse] = (x, s.n_iters, s.l1_residual)
        else:
            A = cls.CsrMatrix(rp, ci, v, m)
            A.meta_create(ctl)
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            xd, bd = cls.DenseVector(x), cls.DenseVector(torch.tensor(b, device="cuda"))
            sc = cls.SolverControl(cls.DIAGONAL, 20000, 1e-10, 0.0)
            assert cls.load().clsparseAmdDcsrcgMixed(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s),
                                                     sc.handle, ctl.handle) == 0
            s = sc.stats()
            sc.release()
            assert s.status_detail == 0 and s.outer_iterations >= 2 and s.a_diagonals == 7
            out[fuse] = (x.cpu().numpy(), s.iterations, s.true_residual)
    np.testing.assert_array_equal(out["1"][0], out["0"][0])
    assert out["1"][1:] == out["0"][1:]
