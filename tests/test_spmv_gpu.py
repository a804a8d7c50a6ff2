"""CSR SpMV on the GPU (clsparseScsrmv / clsparseDcsrmv) against the FP64 oracle.

Per-row relative error = |y - y*| / (|alpha| sum_j |a_ij x_j| + |beta y_i|), with
y* from the oracle (long-double compensated). Bars: 1e-6 FP32 and 1e-13 FP64 on
every row, whatever its length (the FP32 kernels sum each row in FP64).
"""
import numpy as np
import pytest
import torch

import clsparse_amd as cls
import oracle
from gpu_helpers import U32, U64, random_csr, row_tolerance, rp_from_lengths, say, spmv_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctl():
    assert torch.cuda.is_available()
    c = cls.Control()
    yield c
    c.release()


@pytest.mark.parametrize("adaptive", [True, False], ids=["adaptive", "vector"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_poisson_256(ctl, dtype, adaptive):
    """2-D 5-point Poisson on 256x256, x ~ U[-1,1] (the FP32 CPU-path workload)."""
    say(f"poisson256 {dtype.__name__} adaptive={adaptive}")
    rp, ci, v, m = cls.gen_poisson2d(256, 256)
    tol = 1e-6 if dtype == np.float32 else 1e-13
    err, y = spmv_check(rp, ci, v, m, dtype, seed=42, adaptive=adaptive, control=ctl, tol=tol)
    assert not np.isnan(y).any()  # beta = 0 never reads the NaN-filled y


@pytest.mark.parametrize("adaptive", [True, False], ids=["adaptive", "vector"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_powerlaw_small_alpha_beta(ctl, dtype, adaptive):
    say(f"powerlaw small {dtype.__name__} adaptive={adaptive}")
    rp, ci, v, m = cls.gen_powerlaw(40000, 16.0, 1.5, 200000, 3, 25000, seed=7)
    spmv_check(rp, ci, v, m, dtype, alpha=1.5, beta=-0.5, seed=3, adaptive=adaptive,
               control=ctl, tol=row_tolerance(rp, dtype))


EDGE_CASES = {
    "one_row_5000": [5000],
    "one_row_3072": [3072],
    "one_row_1025": [1025],
    "one_row_100k": [100000],
    "rows_1024_exact": [1024, 1024, 512, 512, 1, 1],
    "all_empty": [0] * 300,
    "mostly_empty_one_long": [0] * 100 + [900] + [0] * 100,
    "long_rows_middle_and_tail": [3] * 500 + [7000, 3073, 6144, 6145] + [2] * 300 + [9000],
    "medium_rows": [200, 50, 50, 10] * 100,
    "many_short_rows": [1] * 20000,
    "mixed_lengths": [0, 1, 31, 32, 33, 127, 128, 129, 1023, 1024, 1025, 3071, 3072, 3073] * 20,
}


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_edge_cases(ctl, dtype, name):
    say(f"edge {name} {dtype.__name__}")
    lens = EDGE_CASES[name]
    rng = np.random.default_rng(abs(hash(name)) % 1000)
    ncols = max(120000, max(lens) + 1)
    rp, ci, v = random_csr(rng, lens, ncols)
    for adaptive in (True, False):
        for beta in (0.0, 2.0):
            spmv_check(rp, ci, v, ncols, dtype, alpha=-1.25, beta=beta, seed=11,
                       adaptive=adaptive, control=ctl, tol=row_tolerance(rp, dtype))


MORE_CASES = {"single_entry": [1], "two_rows": [0, 1], "long_run_of_empty_rows": [0] * 5000 + [3],
              "row_of_40k_between_singles": [1, 40000, 1],
              "rows_at_block_edges": [511, 1, 512, 0, 0, 1023, 2, 4096]}


@pytest.mark.parametrize("name", sorted(MORE_CASES))
def test_more_edge_cases(ctl, name):
    """FP32 and FP64, beta 0 and 2: a single entry, runs of empty rows, a
    40k-entry row between single entries, rows ending on block edges."""
    say(f"edge {name}")
    lens = MORE_CASES[name]
    rng = np.random.default_rng(abs(hash(name)) % 1000 + 7)
    ncols = max(60000, max(lens) + 1)
    rp, ci, v = random_csr(rng, lens, ncols)
    for dtype in (np.float32, np.float64):
        for beta in (0.0, 2.0):
            spmv_check(rp, ci, v, ncols, dtype, alpha=0.75, beta=beta, seed=5, control=ctl,
                       tol=row_tolerance(rp, dtype))


def test_row_pointer_not_starting_at_zero(ctl):
    """row_pointer[0] = 5: rows index values and columns from entry 5 on."""
    say("row pointer base 5")
    rng = np.random.default_rng(3)
    rp, ci, v = random_csr(rng, rng.integers(0, 60, 3000), 4000)
    rp5 = (rp + 5).astype(np.uint32)
    ci5 = np.concatenate([np.full(5, 3999, np.uint32), ci])
    v5 = np.concatenate([np.full(5, 1e30), v])
    x = cls.gen_uniform(4000, -1, 1, seed=4)
    A = cls.CsrMatrix(rp5, ci5, v5, 4000)
    A.meta_create(ctl)
    xd = cls.DenseVector(torch.tensor(x, device="cuda"))
    yd = cls.DenseVector(torch.zeros(3000, dtype=torch.float64, device="cuda"))
    cls.check("csrmv", A.spmv(cls.Scalar(1.0), xd, cls.Scalar(0.0), yd, ctl))
    yref, scale = oracle.csrmv(rp, ci, v, x)
    err = np.abs(yd.t.cpu().numpy() - yref) / np.maximum(scale, 1e-300)
    assert err.max() <= 1e-13
    A.meta_delete()


def test_rectangular_and_empty_nnz(ctl):
    rng = np.random.default_rng(5)
    rp, ci, v = random_csr(rng, rng.integers(0, 40, 777), 5000)
    spmv_check(rp, ci, v, 5000, np.float64, alpha=0.5, beta=1.0, control=ctl,
               tol=row_tolerance(rp, np.float64))
    rp0 = np.zeros(11, dtype=np.uint32)
    err, y = spmv_check(rp0, np.zeros(0, np.uint32), np.zeros(0), 10, np.float64, beta=0.0,
                        control=ctl, tol=0.0)
    assert np.all(y == 0)


def test_offsets_are_honoured(ctl):
    """x, y, alpha, beta and all three matrix arrays with element offsets."""
    say("offsets")
    rp, ci, v, m = cls.gen_poisson2d(40, 30)
    n = rp.size - 1
    ox, oy, oa, ov, oc, orp = 7, 3, 2, 5, 9, 4
    dev = "cuda"
    vals = torch.cat([torch.full((ov,), 99.0), torch.tensor(v)]).to(dev)
    cols = torch.cat([torch.full((oc,), 12345, dtype=torch.int32),
                      torch.tensor(ci.view(np.int32))]).to(dev)
    rptr = torch.cat([torch.full((orp,), 77, dtype=torch.int32),
                      torch.tensor(rp.view(np.int32))]).to(dev)
    x = cls.gen_uniform(m, -1, 1, seed=2)
    y0 = cls.gen_uniform(n, -1, 1, seed=3)
    xt = torch.cat([torch.full((ox,), 1e30), torch.tensor(x)]).to(dev)
    yt = torch.cat([torch.full((oy,), -5.0), torch.tensor(y0)]).to(dev)
    sc = torch.tensor([0.0, 0.0, 0.75, -0.5], dtype=torch.float64, device=dev)
    A = cls.clsparseCsrMatrix()
    A.num_rows, A.num_cols, A.num_nonzeros = n, m, v.size
    A.values, A.col_indices, A.row_pointer = vals.data_ptr(), cols.data_ptr(), rptr.data_ptr()
    A.off_values, A.off_col_indices, A.off_row_pointer = ov, oc, orp
    import ctypes
    lib = cls.load()
    xd = cls.cldenseVector(m, xt.data_ptr(), ox)
    yd = cls.cldenseVector(n, yt.data_ptr(), oy)
    al = cls.clsparseScalar(sc.data_ptr(), oa)
    be = cls.clsparseScalar(sc.data_ptr(), oa + 1)
    yref, scale = oracle.csrmv(rp, ci, v, x, 0.75, -0.5, y0)
    for with_meta in (False, True):
        yt[oy:] = torch.tensor(y0, device=dev)
        if with_meta:
            assert lib.clsparseCsrMetaCreate(ctypes.byref(A), ctl.handle) == 0
        assert lib.clsparseDcsrmv(ctypes.byref(al), ctypes.byref(A), ctypes.byref(xd),
                                  ctypes.byref(be), ctypes.byref(yd), ctl.handle) == 0
        out = yt.cpu().numpy()
        assert np.all(out[:oy] == -5.0)
        err = np.abs(out[oy:] - yref) / scale
        assert err.max() <= 1e-13
    lib.clsparseCsrMetaDelete(ctypes.byref(A))


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_extended_precision_on_cancelling_rows(dtype):
    """Rows whose terms cancel: compensated mode is accurate to ~1 ulp of |y|."""
    say(f"extended precision {dtype.__name__}")
    rng = np.random.default_rng(21)
    lens = [8, 40, 700, 1000, 2000, 5000, 20000] * 30
    ncols = 50000
    rp, ci, v = random_csr(rng, lens, ncols)
    x = rng.uniform(-1, 1, ncols)
    # Make each row nearly cancel: append the negated sum of its terms to the last entry.
    for r in range(rp.size - 1):
        s, e = rp[r], rp[r + 1]
        prod = v[s:e - 1] * x[ci[s:e - 1]]
        v[e - 1] = -(prod.sum() + 1e-7) / x[ci[e - 1]]
    u = U32 if dtype == np.float32 else U64
    with cls.Control(extended_precision=True) as c:
        for adaptive in (True, False):
            err_ext, y = spmv_check(rp, ci, v, ncols, dtype, x=x, adaptive=adaptive, control=c)
            yref, scale = oracle.csrmv(rp, ci, v.astype(dtype), x.astype(dtype))
            lens_a = np.diff(rp.astype(np.int64))
            bound = 4 * u * np.abs(yref) + (lens_a * u) ** 2 * scale * 4
            assert np.all(np.abs(y.astype(np.float64) - yref) <= bound)


def test_deterministic_bitwise(ctl):
    say("determinism")
    rp, ci, v, m = cls.gen_powerlaw(60000, 16.0, 1.5, 200000, 2, 60000, seed=9)
    _, y1 = spmv_check(rp, ci, v, m, np.float64, seed=4, control=ctl)
    _, y2 = spmv_check(rp, ci, v, m, np.float64, seed=4, control=ctl)
    assert np.array_equal(y1, y2)


def test_powerlaw_full_size_fp64(ctl):
    """Power-law FP64 workload at full size: n = 2^21, lognormal lengths (mean 16,
    sigma 1.5) plus 8 rows of 100k nnz, against the oracle on every row."""
    say("powerlaw full size")
    rp, ci, v, m = cls.gen_powerlaw(1 << 21, 16.0, 1.5, 200000, 8, 100000, seed=1)
    err, _ = spmv_check(rp, ci, v, m, np.float64, seed=2, control=ctl, tol=1e-13)
    assert err.max() <= 1e-13


def test_linearity_fp64(ctl):
    """A(a x1 + b x2) = a A x1 + b A x2 to rounding, on a power-law matrix."""
    say("linearity")
    rp, ci, v, m = cls.gen_powerlaw(200000, 16.0, 1.5, 200000, 2, 50000, seed=13)
    x1 = cls.gen_uniform(m, -1, 1, seed=1)
    x2 = cls.gen_uniform(m, -1, 1, seed=2)
    _, y1 = spmv_check(rp, ci, v, m, np.float64, x=x1, control=ctl)
    _, y2 = spmv_check(rp, ci, v, m, np.float64, x=x2, control=ctl)
    _, y3 = spmv_check(rp, ci, v, m, np.float64, x=0.5 * x1 - 2.0 * x2, control=ctl)
    _, scale = oracle.csrmv(rp, ci, v, np.abs(0.5 * x1) + np.abs(2.0 * x2))
    diff = np.abs(y3 - (0.5 * y1 - 2.0 * y2))
    assert np.all(diff <= 1e-13 * np.maximum(scale, 1e-300))


def test_meta_shape_mismatch_is_rejected(ctl):
    import ctypes
    rp, ci, v, m = cls.gen_poisson2d(20, 20)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    rp2, ci2, v2, m2 = cls.gen_poisson2d(21, 20)
    B = cls.CsrMatrix(rp2, ci2, v2, m2)
    B.s.meta = A.s.meta  # metadata of another pattern
    x = cls.DenseVector(torch.zeros(m2, dtype=torch.float64, device="cuda"))
    y = cls.DenseVector(torch.zeros(m2, dtype=torch.float64, device="cuda"))
    st = B.spmv(cls.Scalar(1.0), x, cls.Scalar(0.0), y, ctl)
    B.s.meta = None
    assert st == cls.clsparseStructInvalid
    assert cls.load().clsparseDcsrmv(None, ctypes.byref(A.s), None, None, None, ctl.handle) == \
        cls.clsparseInvalidValue


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_sum_ax_is_the_spmv_checksum(ctl, dtype):
    """clsparseAmd[S|D]csrSumAx = 1^T A x = sum of the SpMV's y, against the
    oracle's FP64 SpMV summed in long double; a matrix without entries sums to
    0; a short x is refused."""
    import ctypes
    lib = cls.load()
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    fn = lib.clsparseAmdScsrSumAx if dtype == np.float32 else lib.clsparseAmdDcsrSumAx
    say(f"sum_ax {dtype.__name__}")
    for rp, ci, v, m in (cls.gen_powerlaw(50000, 16.0, 1.5, 200000, 2, 20000, seed=21),
                         cls.gen_poisson2d(40, 30)):
        vt = v.astype(dtype)
        x = cls.gen_uniform(m, -1, 1, seed=8).astype(dtype)
        A = cls.CsrMatrix(rp, ci, vt, m, dtype=dtype)
        xd = cls.DenseVector(torch.tensor(x, device="cuda"))
        out = cls.Scalar(0.0, tdt)
        assert fn(ctypes.byref(A.s), ctypes.byref(xd.s), ctypes.byref(out.s), ctl.handle) == 0
        y, _ = oracle.csrmv(rp, ci, vt.astype(np.float64), x.astype(np.float64))
        ref = float(np.sum(y.astype(np.longdouble)))
        _, scale = oracle.csrmv(rp, ci, np.abs(vt.astype(np.float64)), np.abs(x.astype(np.float64)))
        bound = 1e-13 * float(np.sum(scale)) + (2.0 ** -24 if dtype == np.float32 else 0.0) * abs(ref)
        assert abs(out.item() - ref) <= bound, (out.item(), ref)
    empty = cls.CsrMatrix(np.zeros(4, np.uint32), np.zeros(0, np.uint32), np.zeros(0, dtype), 3,
                          dtype=dtype)
    xd = cls.DenseVector(torch.ones(3, dtype=tdt, device="cuda"))
    out = cls.Scalar(5.0, tdt)
    assert fn(ctypes.byref(empty.s), ctypes.byref(xd.s), ctypes.byref(out.s), ctl.handle) == 0
    assert out.item() == 0.0
    short = cls.DenseVector(torch.ones(2, dtype=tdt, device="cuda"))
    assert fn(ctypes.byref(empty.s), ctypes.byref(short.s), ctypes.byref(out.s),
              ctl.handle) == cls.clsparseInvalidSize


def _hash32(a):
    a = a & 0xFFFFFFFF
    a ^= a >> 16
    a = (a * 0x7FEB352D) & 0xFFFFFFFF
    a ^= a >> 15
    a = (a * 0x846CA68B) & 0xFFFFFFFF
    a ^= a >> 16
    return a


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_gather_sum(ctl, dtype):
    """clsparseAmd[S|D]gatherSum = sum of x[h(i + seed) mod window] over i <
    count, against the same positions formed with numpy's integer arithmetic;
    a window that is not a power of two or exceeds x is refused."""
    import ctypes
    lib = cls.load()
    fn = lib.clsparseAmdSgatherSum if dtype == np.float32 else lib.clsparseAmdDgatherSum
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    x = cls.gen_uniform(1 << 16, -1, 1, seed=12).astype(dtype)
    xd = cls.DenseVector(torch.tensor(x, device="cuda"))
    for window, count, seed in ((1 << 16, 1_000_003, 7), (1 << 10, 300_000, 4_000_000_000), (1, 5, 0)):
        say(f"gather sum {dtype.__name__} window {window} count {count}")
        out = cls.Scalar(0.0, tdt)
        assert fn(ctypes.byref(xd.s), window, count, seed, ctypes.byref(out.s), ctl.handle) == 0
        idx = _hash32(np.arange(count, dtype=np.uint64) + np.uint64(seed)) & np.uint64(window - 1)
        vals = x[idx.astype(np.int64)].astype(np.float64)
        ref = float(np.sum(vals.astype(np.longdouble)))
        bound = 1e-13 * float(np.sum(np.abs(vals))) + (2.0 ** -24 if dtype == np.float32 else 0.0) * abs(ref)
        assert abs(out.item() - ref) <= bound, (out.item(), ref)
    out = cls.Scalar(0.0, tdt)
    assert fn(ctypes.byref(xd.s), 3, 10, 0, ctypes.byref(out.s), ctl.handle) == cls.clsparseInvalidSize
    assert fn(ctypes.byref(xd.s), 1 << 17, 10, 0, ctypes.byref(out.s), ctl.handle) == cls.clsparseInvalidSize
