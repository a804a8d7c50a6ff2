"""Pins the CPU oracle before anything is checked against it (CPU only).

What fixes each oracle function:
  orc_row_blocks  -- 17 hand-derived known answers (tests/golden/row_blocks_known.json)
                     covering every branch of csr-meta.hpp L93-269 (entering / leaving a
                     long-row region, back-off, exact fill, single-row vector, long-row
                     segments, the tail off-by-one, the +2 size-only pass), plus structural
                     invariants on random row-length mixes that a wrong index, a dropped OR or
                     a wrong thread-count formula breaks.
  orc_csrmv_*     -- brute-force dense GEMV with math.fsum on tiny matrices; closed forms
                     (Poisson A*1 = row sums; alpha/beta linearity; beta == 0 ignores a NaN y).
  orc_cg_*        -- closed forms: diagonal A converges in one iteration to the exact
                     solution; A with k distinct eigenvalues needs <= k iterations; b = 0 returns
                     x = b with zeroed tolerances; the maxIters and stored-nIters quirks of
                     solver-control.hpp L47-64; a manufactured solution solved to tolerance.
  orc_dot/nrm*    -- constant vectors (test-blas1.cpp L61-62: x = 4, y = 2).
The reference path itself is unbuildable here (its header chain needs the CMake-generated
clsparse_export.h), so no oracle/_ref comparison exists.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "row_blocks_known.json")


def _rp(lengths):
    return np.concatenate([[0], np.cumsum(np.asarray(lengths, dtype=np.int64))]).astype(np.uint32)


def _known():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _known(), ids=lambda c: c["name"])
def test_oracle_row_blocks_known_answers(case):
    rp = _rp(case["row_lengths"])
    entries, size = oracle.row_blocks(rp)
    assert [int(e) for e in entries] == case["entries"]
    assert size == case["row_block_size"]
    assert oracle.row_blocks(rp, count_only=True) == case["size_pass"]


def nt(r):
    """Reduction threads for a block of r rows: largest power of two <= 256 / r rounded
    the way a 256-lane group splits (256 >> ceil(log2 r)); 0 above 256 rows."""
    return 256 >> math.ceil(math.log2(r)) if r <= 256 else 0


def check_row_block_invariants(rp, entries):
    """Properties any correct binning of csr-meta.hpp must have."""
    n = rp.size - 1
    e = [int(v) for v in entries]
    rows = [v >> 32 for v in e]
    low = [v & 0xFFFFFF for v in e]
    assert e[0] >> 32 == 0 and rows[-1] == n
    assert all(b >= a for a, b in zip(rows, rows[1:]))
    assert all((v >> 24) & 1 == 0 for v in e)  # flag bit clear as created
    k = 0
    while k < len(e) - 1:
        r0, r1 = rows[k], rows[k + 1]
        if r1 == r0:
            # long row: segments k, k+1, ... with ids 1.. until the row ends
            length = int(rp[r0 + 1] - rp[r0])
            segs = math.ceil(length / 3072)
            assert low[k] == 0
            for w in range(1, segs):
                assert rows[k + w] == r0 and low[k + w] == w
            assert rows[k + segs] == r0 + 1
            k += segs
            continue
        nrows = r1 - r0
        nnz = int(rp[r1] - rp[r0])
        if nrows == 1:
            assert low[k] == 0
            length = nnz
            # one row alone: either it fits a block, or it is a <= 3072 vector row
            assert length <= 3072
        else:
            assert nnz <= 1024
            # The block closed by the tail step (the last one, unless an exact 1024-nnz
            # fill ended it) counts one row too many.
            tail = k == len(e) - 2 and nnz != 1024
            expect = nt(nrows + 1) if tail else nt(nrows)
            assert low[k] == expect, (k, nrows, low[k], expect)
        k += 1


@pytest.mark.parametrize("seed", range(6))
def test_oracle_row_blocks_invariants(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    kind = seed % 3
    if kind == 0:
        lens = rng.integers(0, 40, n)
    elif kind == 1:
        lens = np.clip(np.round(np.exp(rng.normal(1.6, 1.5, n))), 0, 20000).astype(np.int64)
    else:
        lens = rng.choice([0, 1, 5, 31, 32, 129, 600, 1024, 1025, 3072, 3073, 7000], n)
    rp = _rp(lens)
    entries, size = oracle.row_blocks(rp)
    assert size == 2 * len(entries)
    check_row_block_invariants(rp, entries)
    sp = oracle.row_blocks(rp, count_only=True)
    assert sp in (size, size + 2)


def _dense(rp, ci, v, n, m):
    a = np.zeros((n, m))
    for r in range(n):
        for j in range(rp[r], rp[r + 1]):
            a[r, ci[j]] += v[j]
    return a


def _random_csr(rng, n, m, maxlen):
    lens = rng.integers(0, maxlen + 1, n)
    rp = _rp(lens)
    ci = np.concatenate([np.sort(rng.choice(m, size=k, replace=False)) for k in lens]).astype(np.uint32)
    v = rng.uniform(-1, 1, rp[-1])
    return rp, ci, v


@pytest.mark.parametrize("seed", range(4))
def test_oracle_csrmv_matches_dense_fsum(seed):
    rng = np.random.default_rng(100 + seed)
    n, m = 40, 37
    rp, ci, v = _random_csr(rng, n, m, 30)
    x = rng.uniform(-1, 1, m)
    y0 = rng.uniform(-1, 1, n)
    alpha, beta = 1.5, -0.75
    y, scale = oracle.csrmv(rp, ci, v, x, alpha, beta, y0)
    for r in range(n):
        terms = [alpha * v[j] * x[ci[j]] for j in range(rp[r], rp[r + 1])] + [beta * y0[r]]
        assert abs(y[r] - math.fsum(terms)) <= 4e-16 * max(1.0, sum(abs(t) for t in terms))
        assert scale[r] == pytest.approx(sum(abs(t) for t in terms), rel=1e-12)
    a = _dense(rp, ci, v, n, m)
    np.testing.assert_allclose(y, alpha * a @ x + beta * y0, rtol=0, atol=1e-13)


def test_oracle_csrmv_poisson_row_sums_and_beta_zero():
    import clsparse_amd as cls
    rp, ci, v, m = cls.gen_poisson2d(9, 7)
    y, _ = oracle.csrmv(rp, ci, v, np.ones(m), 1.0, 0.0, np.full(m, np.nan))
    grid = y.reshape(7, 9)
    expect = np.zeros((7, 9))
    expect[0, :] += 1
    expect[-1, :] += 1
    expect[:, 0] += 1
    expect[:, -1] += 1
    np.testing.assert_array_equal(grid, expect)  # exact: integers


def test_oracle_csrmv_float32_inputs():
    rng = np.random.default_rng(7)
    rp, ci, v = _random_csr(rng, 25, 25, 12)
    v32 = v.astype(np.float32)
    x32 = rng.uniform(-1, 1, 25).astype(np.float32)
    y, _ = oracle.csrmv(rp, ci, v32, x32, 1.0, 0.0)
    a = _dense(rp, ci, v32.astype(np.float64), 25, 25)
    np.testing.assert_allclose(y, a @ x32.astype(np.float64), rtol=0, atol=1e-14)


def test_oracle_blas1_constants():
    n = 1024
    x = np.full(n, 4.0)
    y = np.full(n, 2.0)
    assert oracle.dot(x, y) == 8.0 * n
    assert oracle.nrm1(x) == 4.0 * n
    assert oracle.nrm2(x) == pytest.approx(4.0 * math.sqrt(n), rel=1e-15)


def _diag_csr(d):
    n = d.size
    return np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), d


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_cg_diagonal_one_iteration(dtype):
    d = np.linspace(1.0, 50.0, 64)
    rp, ci, v = _diag_csr(d)
    b = np.linspace(-1, 2, 64)
    x, r = oracle.cg(rp, ci, v.astype(dtype), b.astype(dtype), np.zeros(64), precond=1,
                     max_iters=100, rel_tol=1e-6, abs_tol=0.0)
    assert r.n_iters == 0 and r.executed == 1 and not r.early_exit
    tol = 1e-6 if dtype == np.float32 else 1e-14
    np.testing.assert_allclose(x, b / d, rtol=tol)


def test_oracle_cg_distinct_eigenvalues():
    # Block-diagonal SPD with eigenvalues {1, 2, 5}: CG (no preconditioner) needs <= 3 steps.
    n = 30
    d = np.array([1.0, 2.0, 5.0] * 10)
    q = np.linalg.qr(np.random.default_rng(3).normal(size=(n, n)))[0]
    a = (q * d) @ q.T
    a = 0.5 * (a + a.T)
    rp = np.arange(0, n * n + 1, n, dtype=np.uint32)
    ci = np.tile(np.arange(n, dtype=np.uint32), n)
    b = np.random.default_rng(4).uniform(-1, 1, n)
    x, r = oracle.cg(rp, ci, a.reshape(-1), b, np.zeros(n), precond=0, max_iters=50,
                     rel_tol=1e-12, abs_tol=0.0)
    assert r.executed <= 3
    np.testing.assert_allclose(a @ x, b, atol=1e-10)


def test_oracle_cg_zero_rhs_and_quirks():
    rp, ci, v, m = __import__("clsparse_amd").gen_poisson2d(8, 8)
    x0 = np.full(m, 3.0)
    x, r = oracle.cg(rp, ci, v, np.zeros(m), x0, 1, 10, 1e-2, 1e-5)
    assert r.early_exit and r.n_iters == 0 and r.rel_tol_out == 0 and r.abs_tol_out == 0
    np.testing.assert_array_equal(x, np.zeros(m))
    b = np.ones(m)
    # maxIters = 3, unreachable tolerance: 4 iterations run, nIters reports 3.
    x, r = oracle.cg(rp, ci, v, b, np.zeros(m), 1, 3, 0.0, 0.0)
    assert r.n_iters == 3 and r.executed == 4
    # A stored nIters >= maxIters makes the next call return before the loop.
    x, r = oracle.cg(rp, ci, v, b, np.zeros(m), 1, 3, 0.0, 0.0, n_iters_prior=3)
    assert r.early_exit and r.n_iters == 0
    np.testing.assert_array_equal(x, np.zeros(m))


def test_oracle_cg_manufactured_solution():
    import clsparse_amd as cls
    rp, ci, v, m = cls.gen_poisson2d(24, 24)
    xs = np.random.default_rng(9).uniform(-1, 1, m)
    b, _ = oracle.csrmv(rp, ci, v, xs)
    x, r = oracle.cg(rp, ci, v, b, np.zeros(m), 1, 1000, 1e-12, 0.0)
    assert r.current_residual <= 1e-12
    np.testing.assert_allclose(x, xs, atol=1e-9)


# ---- BiCGStab restatement (biconjugate-gradients-stabilized.hpp L41-268) ------

def _nonsym(n=300, seed=0):
    import scipy.sparse as sp
    main = 4 * np.ones(n)
    a = sp.diags([-1.3 * np.ones(n - 1), main, -0.7 * np.ones(n - 1)], [-1, 0, 1]).tocsr()
    return a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, a


def test_bicgstab_diagonal_jacobi_first_check_exact():
    import scipy.sparse as sp
    d = np.array([2.0, 5.0, 0.25, 9.0])
    a = sp.diags(d).tocsr()
    b = np.array([1.0, -2.0, 3.0, 0.5])
    x, r = oracle.bicgstab(a.indptr, a.indices, a.data, b, np.zeros(4), 1, 100, 1e-12, 0.0)
    assert r.n_iters == 0 and r.executed == 1
    np.testing.assert_array_equal(x, b / d)


def test_bicgstab_manufactured_nonsymmetric_solution():
    rp, ci, v, a = _nonsym()
    xs = np.random.default_rng(1).uniform(-1, 1, a.shape[0])
    b = a @ xs
    for precond in (0, 1):
        x, r = oracle.bicgstab(rp, ci, v, b, np.zeros(a.shape[0]), precond, 1000, 1e-13, 0.0)
        assert np.abs(x - xs).max() < 1e-11
        assert r.current_residual <= 1e-13


def test_bicgstab_zero_rhs_and_max_iters_quirks():
    rp, ci, v, a = _nonsym(50)
    x, r = oracle.bicgstab(rp, ci, v, np.zeros(50), np.ones(50), 1, 100, 1e-6, 1e-3)
    assert r.early_exit == 1 and r.n_iters == 0 and r.rel_tol_out == 0 and r.abs_tol_out == 0
    assert np.all(x == 0)
    b = np.random.default_rng(2).uniform(-1, 1, 50)
    x, r = oracle.bicgstab(rp, ci, v, b, np.zeros(50), 1, 3, 1e-15, 0.0)
    assert r.n_iters == 3 and r.executed == 4
    # a solver control that already holds nIters >= maxIters stops before the loop
    x, r = oracle.bicgstab(rp, ci, v, b, np.zeros(50), 1, 3, 1e-15, 0.0, n_iters_prior=5)
    assert r.early_exit == 1 and r.n_iters == 0


def test_csrmm_oracle_matches_scipy_and_special_cases():
    import scipy.sparse as sp
    m = 350
    a = sp.random(400, m, density=0.03, random_state=7, format="csr")
    a.sort_indices()
    rp, ci, v = a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data
    rng = np.random.default_rng(3)
    B = rng.uniform(-1, 1, (m, 5))
    C = rng.uniform(-1, 1, (400, 5))
    out, scale = oracle.csrmm(rp, ci, v, B, C, alpha=1.5, beta=-0.5)
    np.testing.assert_allclose(out, 1.5 * (a @ B) - 0.5 * C, rtol=0, atol=1e-13)
    assert np.all(scale > 0)
    # alpha = 0: beta * C only; beta = 0: C is never read (NaN stays out)
    out, _ = oracle.csrmm(rp, ci, v, B, C, alpha=0.0, beta=2.0)
    np.testing.assert_array_equal(out, 2.0 * C)
    out, _ = oracle.csrmm(rp, ci, v, B, np.full((400, 5), np.nan), alpha=1.0, beta=0.0)
    assert np.all(np.isfinite(out))


def test_bicgstab_iteration_count_depends_on_summation_order():
    """Why the GPU parity test compares BiCGStab iteration counts within 10 %:
    a numpy restatement with sequential dots matches the oracle exactly, and
    only reordering the dot-product sums moves the count."""
    import scipy.sparse as sp
    nx, ny, c = 48, 40, 0.4
    n = nx * ny
    lap = sp.diags([-np.ones(n - nx), -np.ones(n - 1), 4 * np.ones(n), -np.ones(n - 1),
                    -np.ones(n - nx)], [-nx, -1, 0, 1, nx], format="lil")
    for i in range(1, n):
        if i % nx == 0:
            lap[i, i - 1] = 0.0
            lap[i - 1, i] = 0.0
    a = sp.csr_matrix(lap)
    conv = sp.diags([c * np.ones(n), -c * np.ones(n - 1)], [0, -1], format="lil")
    for i in range(1, n):
        if i % nx == 0:
            conv[i, i - 1] = 0.0
    a = sp.csr_matrix(a + conv)
    a.eliminate_zeros()
    a.sort_indices()
    b = np.random.default_rng(21).uniform(-1, 1, n)

    def run(dot):
        inv = 1 / a.diagonal()
        x = np.zeros(n)
        r = b.copy()
        nb = np.abs(b).sum()
        p, rs = r.copy(), r.copy()
        rr = dot(rs, r)
        for it in range(2000):
            mp = p * inv
            amp = a @ mp
            alpha = rr / dot(rs, amp)
            s = r - alpha * amp
            if np.abs(s).sum() / nb <= 1e-10:
                return it
            ms = s * inv
            ams = a @ ms
            omega = dot(ams, s) / dot(ams, ams)
            x = x + alpha * mp + omega * ms
            r = s - omega * ams
            if np.abs(r).sum() / nb <= 1e-10:
                return it
            rrn = dot(rs, r)
            beta = rrn / rr * alpha / omega
            rr = rrn
            p = beta * p + r - beta * omega * amp
        return 2000

    def seq(u, w):
        acc = 0.0
        for t in (u * w).tolist():
            acc += t
        return acc

    _, ro = oracle.bicgstab(a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, b,
                            np.zeros(n), 1, 2000, 1e-10, 0.0)
    counts = {run(seq), run(lambda u, w: float(np.sum(u * w))), run(lambda u, w: seq(u[::-1], w[::-1]))}
    assert ro.n_iters in counts or min(abs(ro.n_iters - k) for k in counts) <= 2
    assert max(counts) - min(counts) <= 0.15 * max(counts)


def test_oracle_cg_omp_baseline_runs_the_reference_recurrence():
    """The all-core CPU baseline (orc_cg_omp_d) runs the same Jacobi-PCG
    recurrence as orc_cg_d: on the reference's L1 criterion it stops at the
    same iteration (reductions differ only in order), on any thread count; on
    the L2 criterion its solution's true residual meets the tolerance; the
    diagonal closed form converges in one iteration; FP32 SpMV baseline =
    FP32 products of the FP64 oracle to rounding."""
    import clsparse_amd as cls
    rp, ci, v, m = cls.gen_diffusion3d(18, 16, 14, sigma=1.0, seed=5)
    b = cls.gen_uniform(m, -1, 1, seed=6)
    _, ref = oracle.cg(rp, ci, v, b, np.zeros(m), 1, 5000, 1e-8, 0.0)
    for th in (1, 3, 8):
        x, it, res = oracle.cg_omp(rp, ci, v, b, 5000, 1e-8, crit=0, threads=th)
        assert abs(it - (ref.n_iters + 1)) <= 1, (th, it, ref.n_iters)
        assert res <= 1e-8
        x, it, res = oracle.cg_omp(rp, ci, v, b, 5000, 1e-10, crit=1, threads=th)
        ax, _ = oracle.csrmv(rp, ci, v, x)
        assert np.linalg.norm(b - ax) / np.linalg.norm(b) <= 1.01e-10
    d = np.array([2.0, 5.0, 0.5, 4.0])
    rpd = np.arange(5, dtype=np.uint32)
    x, it, _ = oracle.cg_omp(rpd, np.arange(4, dtype=np.uint32), d, np.ones(4), 10, 1e-14, 1, 2)
    assert it == 1 and np.allclose(x, 1 / d, rtol=1e-15)
    xs = cls.gen_uniform(m, -1, 1, seed=7).astype(np.float32)
    y32 = np.zeros(m, np.float32)
    oracle.spmv_omp_s(rp, ci, v.astype(np.float32), xs, y32, 4)
    yref, scale = oracle.csrmv(rp, ci, v.astype(np.float32), xs)
    assert np.all(np.abs(y32 - yref) <= 8 * 2.0 ** -24 * scale)
