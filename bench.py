"""Benchmark: FP64 CSR-Adaptive SpMV on the power-law matrix (one MI355X).

Workload: n = 2^21 rows/cols, row lengths clip(round(LogNormal(ln 16 - 1.125, 1.5)), 1,
200000) plus 8 rows of 100,000 nnz at seeded positions, unique sorted uniform columns,
values and x uniform[-1, 1] (FP64). One step = one clsparseDcsrmv (y = A x) on
device-resident inputs. GB/s uses the reference's byte count
(src/benchmarks/clsparse-bench/functions/clfunc_xSpMdV.hpp L91):
    bytes = 4 (nnz + n) + sizeof(T) (nnz + ncols + n)
Prints one JSON line. The CPU baseline is the oracle's OpenMP FP64 CSR SpMV on
the host cores, timed on a bounded number of calls of the same matrix. The
line's "extra" block carries the solver workload: 3-D 7-point variable-coefficient
diffusion on 128^3 (kappa ~ LogNormal(0, 1), harmonic-mean faces), b ~ U[-1, 1],
solved to an FP64 true relative residual of 1e-10 by mixed-precision refinement
and by FP64 Jacobi-PCG.

Other modes (diagnostics, not the bench line):
  --mode copy      plain streaming kernel of known size, to calibrate rocprof counters
  --mode gather    independent gathers at hashed positions in windows of x (32 KiB .. 512 MiB),
                   the gather rate of each cache level (clsparseAmd[D|S]gatherSum)
  --mode locality  the same power-law rows with columns folded into smaller windows
                   (x-gather locality sweep), and stencil SpMVs: s64 / s32 (128^3),
                   s32:160 (160^3, z faces x100), t32:160 (its lower triangle)
  --mode cg        only the solver workload
  --mode fsai      only the learned-FSAI workload: held-out 3-D anisotropic diffusion
                   on 160^3 (z faces x100, kappa ~ LogNormal(0, 1)) solved to an FP64 true
                   relative residual of 1e-8 by learned-FSAI mixed PCG, mixed Jacobi-PCG and
                   FP64 Jacobi-PCG, with the committed trained network
  --mode fsai-train  trains the network on the seeded 2-D anisotropic family (512^2,
                   2000 matrices, the last 200 held out) and writes the weights
  --mode next      the rows after the SpMV/CG path: extended-precision SpMV, BiCGStab,
                   csrmm and the Matrix Market reader, each with its bytes and a CPU figure
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpMV GB/s & % HBM peak; CG s-to-1e-8 resid, learned FSAI vs Jacobi"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n, ncols, nnz, value_bytes=8):
    return 4 * (nnz + n) + value_bytes * (nnz + ncols + n)


def cpu_baseline(rp, ci, v, x, budget_s=3.0, max_calls=30, fp32=False):
    """Median time of the oracle's OpenMP CSR SpMV on all host cores (FP64, or
    FP32 values and sums with fp32)."""
    import numpy as np

    import oracle
    threads = oracle.max_threads()
    if fp32:
        v, x = np.ascontiguousarray(v, np.float32), np.ascontiguousarray(x, np.float32)
        y = np.empty(rp.size - 1, np.float32)
        fn = oracle.spmv_omp_s
    else:
        y = np.empty(rp.size - 1)
        fn = oracle.spmv_omp
    fn(rp, ci, v, x, y, threads)  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < max_calls and (len(times) < 3 or time.perf_counter() < t_end):
        t0 = time.perf_counter()
        fn(rp, ci, v, x, y, threads)
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2], threads, len(times)


def reference_methodology(call, torch, samples=20, calls=20):
    """The reference's SpMV timing (docs/performance/clSPARSE-0.8.0.0/W9100/README.md
    L15-20; clsparseTimer-device.cpp L377-431): `samples` host-timed batches of
    `calls` back-to-back calls (synchronised after each batch), samples more
    than one standard deviation from the mean dropped, mean of the rest per call."""
    import numpy as np
    for _ in range(3):
        call()
    torch.cuda.synchronize()
    per = []
    for _ in range(samples):
        t0 = time.perf_counter()
        for _ in range(calls):
            call()
        torch.cuda.synchronize()
        per.append((time.perf_counter() - t0) / calls)
    a = np.array(per)
    keep = a[np.abs(a - a.mean()) <= a.std()]
    return float(keep.mean()), int(keep.size)


def load_traffic():
    path = os.path.join(ROOT, "profiles", "spmv_powerlaw_fp64_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


class SpmvRun:
    """Device-resident CSR-Adaptive SpMV y = A x of one matrix."""

    def __init__(self, cls, torch, rp, ci, v, m, x, dtype, ctl):
        import numpy as np
        tdt = torch.float32 if dtype == np.float32 else torch.float64
        self.cls, self.torch, self.ctl = cls, torch, ctl
        self.A = cls.CsrMatrix(rp, ci, v, m, dtype=dtype)
        self.A.meta_create(ctl)
        self.x = cls.DenseVector(torch.tensor(np.asarray(x, dtype=dtype), device="cuda"))
        self.y = cls.DenseVector(torch.zeros(rp.size - 1, dtype=tdt, device="cuda"))
        self.alpha, self.beta = cls.Scalar(1.0, tdt), cls.Scalar(0.0, tdt)
        self.nbytes = spmv_bytes(rp.size - 1, m, int(rp[-1]), np.dtype(dtype).itemsize)

    def step(self):
        st = self.A.spmv(self.alpha, self.x, self.beta, self.y, self.ctl)
        if st != 0:
            raise self.cls.ClsparseError("csrmv", st)

    def time(self, steps, warmup, stream):
        torch = self.torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3 / steps
        per = []
        for _ in range(min(50, steps)):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            self.step()
            b.record(stream)
            per.append((a, b))
        torch.cuda.synchronize()
        launch_ms = sorted(a.elapsed_time(b) for a, b in per)[len(per) // 2]
        return wall_ms, launch_ms

    def sum_ax(self, calls, stream):
        """Median time (ms) of clsparseAmd[S|D]csrSumAx on the same A and x (one
        pass over values, columns and x gathers, no rows, no y) and its value."""
        torch = self.torch
        lib = self.cls.load()
        tdt = self.alpha.t.dtype if hasattr(self.alpha, "t") else torch.float64
        out = self.cls.Scalar(0.0, tdt)
        fn = lib.clsparseAmdDcsrSumAx if tdt == torch.float64 else lib.clsparseAmdScsrSumAx
        import ctypes

        def call():
            st = fn(ctypes.byref(self.A.s), ctypes.byref(self.x.s), ctypes.byref(out.s),
                    self.ctl.handle)
            if st != 0:
                raise self.cls.ClsparseError("csrSumAx", st)
        for _ in range(3):
            call()
        per = []
        for _ in range(calls):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            call()
            b.record(stream)
            per.append((a, b))
        torch.cuda.synchronize()
        ms = sorted(a.elapsed_time(b) for a, b in per)[len(per) // 2]
        return ms, float(out.t.double().item())

    def close(self):
        self.A.meta_delete()


def median_call_ms(torch, stream, call, calls=20, warmup=3):
    """Median of HIP-event times of single calls on `stream`."""
    for _ in range(warmup):
        call()
    per = []
    for _ in range(calls):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        call()
        b.record(stream)
        per.append((a, b))
    torch.cuda.synchronize()
    return sorted(a.elapsed_time(b) for a, b in per)[len(per) // 2]


def gather_floors(cls, torch, ctl, stream, count, calls=15):
    """clsparseAmd[D|S]gatherSum: `count` independent gathers at hashed
    positions inside windows of x of growing size (FP64 32 KiB .. 16 MiB,
    512 MiB; FP32 8 MiB): the gather rate each level of the memory hierarchy
    gives, the floor under any CSR SpMV whose columns are scattered over that
    much of x."""
    lib = cls.load()
    out = []
    for dt, bits in ((torch.float64, 12), (torch.float64, 15), (torch.float64, 18), (torch.float64, 19),
                     (torch.float64, 20), (torch.float64, 21), (torch.float64, 26), (torch.float32, 21)):
        n = 1 << bits
        x = torch.rand(n, dtype=dt, device="cuda")
        xd, res = cls.DenseVector(x), cls.Scalar(0.0, dt)
        fn = lib.clsparseAmdDgatherSum if dt == torch.float64 else lib.clsparseAmdSgatherSum

        def call(fn=fn, xd=xd, res=res, n=n):
            cls.check("gatherSum", fn(ctypes.byref(xd.s), n, count, 7, ctypes.byref(res.s), ctl.handle))
        ms = median_call_ms(torch, stream, call, calls=calls)
        item = 8 if dt == torch.float64 else 4
        out.append({"window_bytes": n * item, "dtype": "f64" if item == 8 else "f32", "gathers": count,
                    "us": round(ms * 1e3, 1), "Ggathers_per_s": round(count / (ms * 1e-3) / 1e9, 1)})
        log(f"gather floor: {out[-1]}")
        del x
    return out


def copy_peak(cls, torch, ctl, stream, n=1 << 26):
    """cldenseDscale on 512 MiB (read 8 B + write 8 B per element): the
    streaming rate one library kernel reaches, beside the 8 TB/s figure."""
    lib = cls.load()
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    r, y = cls.DenseVector(dst), cls.DenseVector(src)
    one = cls.Scalar(1.0, torch.float64)

    def call():
        cls.check("cldenseDscale", lib.cldenseDscale(ctypes.byref(r.s), ctypes.byref(one.s),
                                                     ctypes.byref(y.s), ctl.handle))
    ms = median_call_ms(torch, stream, call)
    return {"kernel": "cldenseDscale (y = 1 * x), FP64", "bytes": 16 * n, "us": round(ms * 1e3, 2),
            "GBps": round(16 * n / (ms * 1e-3) / 1e9, 1)}


def next_rows(cls, torch, ctl, stream, cpu=True):
    """Extended-precision SpMV, BiCGStab, csrmm and the Matrix Market reader."""
    import numpy as np
    import scipy.sparse as sp

    import oracle
    lib = cls.load()
    out = {}

    # Extended precision: the power-law FP64 SpMV with compensated sums.
    rp, ci, v, m = cls.gen_powerlaw(1 << 21, 16.0, 1.5, 200000, 8, 100000, seed=1)
    x = cls.gen_uniform(m, -1.0, 1.0, seed=2)
    res = {}
    for name, ext in (("plain", False), ("extended", True)):
        c = cls.Control(stream=stream, async_mode=True, extended_precision=ext)
        run = SpmvRun(cls, torch, rp, ci, v, m, x, np.float64, c)
        _, launch = run.time(50, 5, stream)
        res[name] = {"launch_us": round(launch * 1e3, 2),
                     "GBps": round(run.nbytes / (launch * 1e-3) / 1e9, 1)}
        run.close()
        c.release()
    res["bytes_per_call"] = spmv_bytes(rp.size - 1, m, int(rp[-1]))
    out["extended_precision_spmv"] = res
    log(f"extended precision: {res}")
    del rp, ci, v, x

    # BiCGStab: 3-D convection-diffusion (upwind in x), nonsymmetric, Jacobi, 1e-8.
    n1 = 96
    rp, ci, v, m = cls.gen_diffusion3d(n1, n1, n1, sigma=1.0, ax=1.0, ay=1.0, az=1.0, seed=7)
    A0 = sp.csr_matrix((v, ci, rp), shape=(m, m))
    conv = 0.5
    rows = np.arange(m)
    left = rows[rows % n1 > 0]
    U = sp.csr_matrix((np.concatenate([np.full(m, conv), np.full(left.size, -conv)]),
                       (np.concatenate([rows, left]), np.concatenate([rows, left - 1]))), shape=(m, m))
    Ab = (A0 + U).tocsr()
    Ab.sort_indices()
    rpb, cib, vb = Ab.indptr.astype(np.uint32), Ab.indices.astype(np.uint32), Ab.data
    b = cls.gen_uniform(m, -1.0, 1.0, seed=8)
    A = cls.CsrMatrix(rpb, cib, vb, m)
    A.meta_create(ctl)
    bt = torch.tensor(b, device="cuda")
    tol = 1e-8
    best = None
    for _ in range(2):
        xt = torch.zeros(m, dtype=torch.float64, device="cuda")
        sc = cls.SolverControl(cls.DIAGONAL, 20000, tol, 0.0)
        xd, bd = cls.DenseVector(xt), cls.DenseVector(bt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = lib.clsparseDcsrbicgStab(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s),
                                      sc.handle, ctl.handle)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stats = sc.stats()
        sc.release()
        cls.check("clsparseDcsrbicgStab", st)
        best = (dt, stats.n_iters + 1, stats.current_residual, stats.a_diagonals)
    nnzb = int(rpb[-1])
    # Bytes of one iteration as executed: two SpMVs (diagonal storage: K values
    # + x + y per row; CSR: the reference formula), the s, x/r and p updates
    # (40 + 64 + 48 B per row) and the dots <r*, A Mp>, <A Ms, s>, <A Ms, A Ms>
    # (fused into the diagonal SpMVs: r* and s read, 16 B; else two dot
    # kernels, 16 + 16 B).
    K = best[3]
    spmv = (K * 8 + 16) * m if K else spmv_bytes(m, m, nnzb)
    it_bytes = 2 * spmv + 152 * m + (16 if K else 32) * m
    r = b - oracle.csrmv(rpb, cib, vb, xt.cpu().numpy())[0]
    res = {"workload": f"convection_diffusion3d_{n1}^3_upwind{conv}", "n": m, "nnz": nnzb,
           "tolerance_l1_relative": tol, "seconds": round(best[0], 5), "iterations": best[1],
           "l1_residual": best[2], "true_rel_residual_l2": float(np.linalg.norm(r) / np.linalg.norm(b)),
           "a_diagonals": K, "bytes_per_iteration": it_bytes,
           "GBps": round(it_bytes * best[1] / best[0] / 1e9, 1)}
    if cpu:
        k = 4
        t0 = time.perf_counter()
        oracle.bicgstab(rpb, cib, vb, b, np.zeros(m), 1, k, 1e-30, 0.0)
        per = (time.perf_counter() - t0) / k
        res["cpu_baseline"] = {"kind": "port", "cores": 1, "seconds_per_iteration": round(per, 5),
                               "projected_seconds": round(per * best[1], 3),
                               "sample": f"{k} iterations of the oracle's FP64 BiCGStab on the same "
                                         "matrix, single thread, projected"}
    # clsparseScsrbicgStab on the same system in FP32, to 1e-6.
    A32 = cls.CsrMatrix(rpb, cib, vb.astype(np.float32), m, dtype=np.float32)
    A32.meta_create(ctl)
    bt32 = bt.float()
    best32 = None
    for _ in range(2):
        xt32 = torch.zeros(m, dtype=torch.float32, device="cuda")
        sc = cls.SolverControl(cls.DIAGONAL, 20000, 1e-6, 0.0)
        xd, bd = cls.DenseVector(xt32), cls.DenseVector(bt32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = lib.clsparseScsrbicgStab(ctypes.byref(xd.s), ctypes.byref(A32.s), ctypes.byref(bd.s),
                                      sc.handle, ctl.handle)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stats = sc.stats()
        sc.release()
        cls.check("clsparseScsrbicgStab", st)
        best32 = {"seconds": round(dt, 5), "iterations": stats.n_iters + 1,
                  "l1_residual": stats.current_residual, "tolerance_l1_relative": 1e-6,
                  "a_diagonals": stats.a_diagonals}
    res["fp32"] = best32
    A32.meta_delete()
    out["bicgstab"] = res
    log(f"bicgstab: {res}")
    A.meta_delete()

    # csrmm: 2-D Poisson 1024^2 (FP64) times a dense 32-column B.
    rp, ci, v, m = cls.gen_poisson2d(1024, 1024)
    k = 32
    nnz = int(rp[-1])
    A = cls.CsrMatrix(rp, ci, v, m)
    B = torch.tensor(cls.gen_uniform(m * k, -1.0, 1.0, seed=9), device="cuda")
    C = torch.zeros(m * k, dtype=torch.float64, device="cuda")

    def dense(t, rows, cols):
        d = cls.cldenseMatrix()
        lib.cldenseInitMatrix(ctypes.byref(d))
        d.num_rows, d.num_cols, d.lead_dim, d.major = rows, cols, cols, 1
        d.values, d.off_values = t.data_ptr(), 0
        return d
    Bd, Cd = dense(B, m, k), dense(C, m, k)
    one, zero = cls.Scalar(1.0, torch.float64), cls.Scalar(0.0, torch.float64)

    def call():
        cls.check("clsparseDcsrmm", lib.clsparseDcsrmm(ctypes.byref(one.s), ctypes.byref(A.s),
                                                       ctypes.byref(Bd), ctypes.byref(zero.s),
                                                       ctypes.byref(Cd), ctl.handle))
    ms = median_call_ms(torch, stream, call)
    nbytes = 12 * nnz + 4 * (m + 1) + 2 * 8 * m * k
    res = {"workload": f"poisson2d_1024^2_fp64_times_{k}_columns", "n": m, "nnz": nnz, "ncols": k,
           "us": round(ms * 1e3, 2), "bytes": nbytes, "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
           "GFLOPs": round(2 * nnz * k / (ms * 1e-3) / 1e9, 1),
           "bytes_rule": "A once (values, columns, row pointer) + B once + C written once"}
    rows = np.arange(0, m, 997)
    Bh = B.cpu().numpy().reshape(m, k)
    sub_rp = np.concatenate([[0], np.cumsum(np.diff(rp.astype(np.int64))[rows])]).astype(np.uint32)
    sub_ci = np.concatenate([ci[rp[r]:rp[r + 1]] for r in rows])
    sub_v = np.concatenate([v[rp[r]:rp[r + 1]] for r in rows])
    ref, scale = oracle.csrmm(sub_rp, sub_ci, sub_v, Bh, np.zeros((rows.size, k)))
    got = C.cpu().numpy().reshape(m, k)[rows]
    res["max_rel_err_sample"] = float(np.max(np.abs(got - ref) / np.maximum(scale, 1e-300)))
    if cpu:
        t0 = time.perf_counter()
        oracle.csrmm(rp, ci, v, Bh, np.zeros((m, k)))
        res["cpu_baseline"] = {"kind": "port", "cores": 1, "us": round((time.perf_counter() - t0) * 1e6, 1),
                               "sample": "one full call of the oracle's csrmm (long-double sums), "
                                         "single thread"}
    # The same product in FP32 (4 columns per lane in 16-byte groups).
    A32 = cls.CsrMatrix(rp, ci, v.astype(np.float32), m, dtype=np.float32)
    B32, C32 = B.float(), torch.zeros(m * k, dtype=torch.float32, device="cuda")
    Bd32, Cd32 = dense(B32, m, k), dense(C32, m, k)
    one32, zero32 = cls.Scalar(1.0, torch.float32), cls.Scalar(0.0, torch.float32)

    def call32():
        cls.check("clsparseScsrmm", lib.clsparseScsrmm(ctypes.byref(one32.s), ctypes.byref(A32.s),
                                                       ctypes.byref(Bd32), ctypes.byref(zero32.s),
                                                       ctypes.byref(Cd32), ctl.handle))
    ms32 = median_call_ms(torch, stream, call32)
    nb32 = 8 * nnz + 4 * (m + 1) + 2 * 4 * m * k
    res["fp32"] = {"us": round(ms32 * 1e3, 2), "bytes": nb32, "GBps": round(nb32 / (ms32 * 1e-3) / 1e9, 1),
                   "GFLOPs": round(2 * nnz * k / (ms32 * 1e-3) / 1e9, 1)}
    out["csrmm"] = res
    log(f"csrmm: {res}")
    del B, C, B32, C32

    # Matrix Market: power-law 2^16 x 2^16 written with 17 significant digits.
    import tempfile
    rp, ci, v, m = cls.gen_powerlaw(1 << 16, 16.0, 1.5, 20000, 2, 5000, seed=3)
    nnz = int(rp[-1])
    rows = np.repeat(np.arange(1, m + 1), np.diff(rp.astype(np.int64)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "powerlaw.mtx")
        with open(path, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real general\n")
            f.write(f"{m} {m} {nnz}\n")
            np.savetxt(f, np.column_stack([rows, ci.astype(np.int64) + 1, v]), fmt="%d %d %.17g")
        size = os.path.getsize(path)
        nnz_h, nr, nc = cls.mtx_header(path)
        vals = torch.zeros(nnz_h, dtype=torch.float64, device="cuda")
        colt = torch.zeros(nnz_h, dtype=torch.int32, device="cuda")
        rpt = torch.zeros(nr + 1, dtype=torch.int32, device="cuda")
        Am = cls.clsparseCsrMatrix()
        lib.clsparseInitCsrMatrix(ctypes.byref(Am))
        Am.values, Am.col_indices, Am.row_pointer = vals.data_ptr(), colt.data_ptr(), rpt.data_ptr()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            cls.check("clsparseDCsrMatrixfromFile",
                      lib.clsparseDCsrMatrixfromFile(ctypes.byref(Am), path.encode(), ctl.handle, 0))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        t = sorted(times)[1]
        ok = bool(np.array_equal(rpt.cpu().numpy().view(np.uint32), rp) and
                  np.array_equal(colt.cpu().numpy().view(np.uint32), ci) and
                  np.array_equal(vals.cpu().numpy(), v))
        res = {"file_bytes": size, "nnz": nnz, "seconds": round(t, 4),
               "MBps": round(size / t / 1e6, 1), "Mnnz_per_s": round(nnz / t / 1e6, 2),
               "round_trip_exact": ok,
               "what": "clsparseDCsrMatrixfromFile: parse, sort, CSR build and upload"}
        if cpu:
            import scipy.io
            t0 = time.perf_counter()
            scipy.io.mmread(path)
            res["cpu_comparison"] = {"kind": "scipy.io.mmread (coordinate parse only)",
                                     "seconds": round(time.perf_counter() - t0, 4)}
    out["matrix_market"] = res
    log(f"matrix market: {res}")
    return out


def poisson_workload(cls, torch, ctl, stream, cpu=True):
    """FP32 SpMV on the 2-D 5-point Poisson matrix, 256 x 256 (the reference's
    CPU-runnable case): microseconds per call beside the CPU oracle."""
    import numpy as np
    rp, ci, v, m = cls.gen_poisson2d(256, 256)
    x = cls.gen_uniform(m, -1.0, 1.0, seed=2)
    run = SpmvRun(cls, torch, rp, ci, v, m, x, np.float32, ctl)
    wall, launch = run.time(200, 20, stream)
    out = {"workload": "poisson2d_256x256_fp32_spmv", "n": m, "nnz": int(rp[-1]),
           "launch_us": round(launch * 1e3, 2), "wall_us_per_call": round(wall * 1e3, 2),
           "GBps": round(run.nbytes / (launch * 1e-3) / 1e9, 1),
           "note": "3.4 MB: L2/Infinity-Cache resident, so not an HBM figure"}
    run.close()
    if cpu:
        med, threads, calls = cpu_baseline(rp, ci, v, x, budget_s=1.0, max_calls=50, fp32=True)
        out["cpu_baseline"] = {"us_per_call": round(med * 1e6, 1), "cores": threads, "kind": "port",
                               "sample": f"median of {calls} calls, oracle OpenMP CSR SpMV with FP32 "
                                         "values and sums"}
        med1, _, calls1 = cpu_baseline_1t(rp, ci, v, x)
        out["cpu_baseline"]["single_thread_us_per_call"] = round(med1 * 1e6, 1)
    return out


def cpu_baseline_1t(rp, ci, v, x, calls=20):
    """The FP32 oracle SpMV on one thread (median of `calls`)."""
    import numpy as np

    import oracle
    v, x = np.ascontiguousarray(v, np.float32), np.ascontiguousarray(x, np.float32)
    y = np.empty(rp.size - 1, np.float32)
    ts = []
    for _ in range(calls + 1):
        t0 = time.perf_counter()
        oracle.spmv_omp_s(rp, ci, v, x, y, 1)
        ts.append(time.perf_counter() - t0)
    ts = sorted(ts[1:])
    return ts[len(ts) // 2], 1, calls


def cpu_cg_baseline(rp, ci, v, b, tol, reps=3, single_thread=False):
    """The oracle's FP64 Jacobi-PCG (the reference's recurrence,
    conjugate-gradients.hpp L57-227) run to tolerance from x = 0 on all host
    cores (OpenMP), stopping on ||r||_2 / ||b||_2 <= tol like the FP64 GPU
    solver it sits beside: median of `reps` solves; optionally one
    single-threaded solve as well."""
    import oracle
    threads = oracle.max_threads()
    times, its = [], 0
    for _ in range(reps):
        t0 = time.perf_counter()
        _, its, res = oracle.cg_omp(rp, ci, v, b, 100000, tol, crit=1, threads=threads)
        times.append(time.perf_counter() - t0)
    out = {"kind": "port", "cores": threads, "seconds": round(float(sorted(times)[len(times) // 2]), 3),
           "seconds_all": [round(t, 3) for t in times], "iterations": its, "final_rel_residual": res,
           "sample": f"median of {reps} full solves to ||r||_2/||b||_2 <= {tol:g}: oracle FP64 "
                     f"Jacobi-PCG (orc_cg_omp_d), {threads} OpenMP threads"}
    if single_thread:
        t0 = time.perf_counter()
        _, its1, _ = oracle.cg_omp(rp, ci, v, b, 100000, tol, crit=1, threads=1)
        out["single_thread"] = {"cores": 1, "seconds": round(time.perf_counter() - t0, 3),
                                "iterations": its1, "sample": "one full solve, one thread"}
    return out


class env:
    """Sets library switches (environment variables) inside the block."""

    def __init__(self, **kv):
        self.kv = kv

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kv}
        os.environ.update(self.kv)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def csr_only():
    """Within the block the solvers keep A in CSR (CLSPARSE_AMD_DIA=0): the
    solvers on the reference's CSR-Adaptive SpMV."""
    return env(CLSPARSE_AMD_DIA="0")


def timed_solves(cls, torch, ctl, fn, A, bt, m, tol, reps=5, fsai=None, max_iters=50000):
    """reps + 1 solves from x = 0; the first is reported on its own (it
    captures the solver's graphs and, for a new learned factor, runs the
    loss estimate or pilot race), the median and spread of the other reps are the figure."""
    rows = []
    for _ in range(reps + 1):
        x = torch.zeros(m, dtype=torch.float64, device="cuda")
        sc = cls.SolverControl(cls.DIAGONAL, max_iters, tol, 0.0)
        xd, bd = cls.DenseVector(x), cls.DenseVector(bt)
        args = (ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s)) + \
            ((fsai.handle,) if fsai is not None else ()) + (sc.handle, ctl.handle)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = fn(*args)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s = sc.stats()
        sc.release()
        cls.check("solve", st)
        rows.append((dt, s))
    first, rest = rows[0], rows[1:]
    ts = sorted(t for t, _ in rest)
    s = rest[-1][1]
    out = {"seconds": round(ts[len(ts) // 2], 5), "seconds_min_max": [round(ts[0], 5), round(ts[-1], 5)],
           "solves": len(rest), "iterations": s.iterations, "replacements": s.outer_iterations,
           "true_rel_residual": s.true_residual, "converged": s.status_detail == 0,
           "preconditioner": s.preconditioner, "a_diagonals": s.a_diagonals, "g_diagonals": s.g_diagonals,
           "first_solve": {"seconds": round(first[0], 5), "iterations": first[1].iterations}}
    if fsai is not None and first[1].pilot_rate_fsai > 0:
        # the first solve's pilot race: -ln(residual ratio) per second
        out["first_solve"]["pilot_rates"] = {"fsai": round(first[1].pilot_rate_fsai, 1),
                                             "jacobi": round(first[1].pilot_rate_jacobi, 1)}
    return out


def solver_workload(cls, torch, ctl, n1d=128, tol=1e-10, cpu=True):
    """Mixed-precision refinement vs FP64 Jacobi-PCG to an FP64 true relative
    residual, with A's SpMV from diagonal storage and from CSR-Adaptive, and
    with the solution kept as an FP32 increment or updated in FP64."""
    rp, ci, v, m = cls.gen_diffusion3d(n1d, n1d, n1d, sigma=1.0, seed=3)
    b = cls.gen_uniform(m, -1.0, 1.0, seed=4)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    bt = torch.tensor(b, device="cuda")
    lib = cls.load()
    out = {"workload": f"diffusion3d_{n1d}^3_lognormal1_jacobi", "n": m, "nnz": int(rp[-1]),
           "tolerance": tol, "timing": "median of 5 solves after a first one (reported apart)"}
    runs = [("mixed_fp32_inner", lib.clsparseAmdDcsrcgMixed, {}),
            ("mixed_fp32_inner_csr", lib.clsparseAmdDcsrcgMixed, {"CLSPARSE_AMD_DIA": "0"}),
            ("mixed_fp64_x", lib.clsparseAmdDcsrcgMixed, {"CLSPARSE_AMD_MIXED_X64": "1"}),
            ("mixed_fp64_x_csr", lib.clsparseAmdDcsrcgMixed,
             {"CLSPARSE_AMD_MIXED_X64": "1", "CLSPARSE_AMD_DIA": "0"}),
            ("fp64_jacobi_pcg", lib.clsparseAmdDcsrcgRefined, {}),
            ("fp64_jacobi_pcg_csr", lib.clsparseAmdDcsrcgRefined, {"CLSPARSE_AMD_DIA": "0"})]
    for name, fn, kv in runs:
        with env(**kv):
            out[name] = timed_solves(cls, torch, ctl, fn, A, bt, m, tol)
        log(f"{name}: {out[name]}")
    sec = {k: out[k]["seconds"] for k, _, _ in runs}
    out["speedup_mixed_vs_fp64"] = round(sec["fp64_jacobi_pcg"] / sec["mixed_fp32_inner"], 3)
    out["speedup_mixed_vs_fp64_csr_adaptive_both"] = round(sec["fp64_jacobi_pcg_csr"] /
                                                           sec["mixed_fp32_inner_csr"], 3)
    out["speedup_mixed_fp64_x_vs_fp64"] = round(sec["fp64_jacobi_pcg"] / sec["mixed_fp64_x"], 3)
    out["metadata_create_s"] = meta_create_seconds(cls, torch, ctl, rp, ci, v, m)
    if cpu:
        out["cpu_baseline"] = cpu_cg_baseline(rp, ci, v, b, tol, reps=3, single_thread=True)
        log(f"CPU baseline: {out['cpu_baseline']}")
    out["reference_api_cg"] = reference_api_cg(cls, torch, ctl, rp, ci, v, b, m)
    out["blas1_fp64"] = blas1_workload(cls, torch, ctl, ctl.stream)
    A.meta_delete()
    return out


def meta_create_seconds(cls, torch, ctl, rp, ci, v, m, reps=3):
    """Median wall time of clsparseCsrMetaCreate (row pointer read back, binning
    of csr-meta.hpp on the host, descriptors and long-row tables uploaded)."""
    A = cls.CsrMatrix(rp, ci, v, m)
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A.meta_create(ctl)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        A.meta_delete()
    return round(sorted(ts)[len(ts) // 2], 5)


def blas1_workload(cls, torch, ctl, stream, n=1 << 26):
    """The reference's dense BLAS-1 entry points on FP64 vectors of 2^26
    values (512 MiB each): median call time and the rate of the bytes each
    one must move (reads of x / y, the write of r)."""
    lib = cls.load()
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.rand(n, dtype=torch.float64, device="cuda")
    r = torch.empty_like(x)
    xv, yv, rv = cls.DenseVector(x), cls.DenseVector(y), cls.DenseVector(r)
    a, b, out = cls.Scalar(0.5), cls.Scalar(-2.0), cls.Scalar(0.0)
    B = ctypes.byref
    ops = {
        "dot": (lambda: lib.cldenseDdot(B(out.s), B(xv.s), B(yv.s), ctl.handle), 16),
        "nrm1": (lambda: lib.cldenseDnrm1(B(out.s), B(xv.s), ctl.handle), 8),
        "nrm2": (lambda: lib.cldenseDnrm2(B(out.s), B(xv.s), ctl.handle), 8),
        "axpy": (lambda: lib.cldenseDaxpy(B(rv.s), B(a.s), B(xv.s), B(yv.s), ctl.handle), 24),
        "axpby": (lambda: lib.cldenseDaxpby(B(rv.s), B(a.s), B(xv.s), B(b.s), B(yv.s), ctl.handle), 24),
        "scale": (lambda: lib.cldenseDscale(B(rv.s), B(a.s), B(xv.s), ctl.handle), 16),
    }
    res = {"n": n}
    for name, (fn, per) in ops.items():
        def call(fn=fn, name=name):
            cls.check(f"cldenseD{name}", fn())
        ms = median_call_ms(torch, stream, call, calls=10)
        res[name] = {"us": round(ms * 1e3, 1), "GBps": round(per * n / (ms * 1e-3) / 1e9, 1)}
    log(f"BLAS-1: {res}")
    return res


def reference_api_cg(cls, torch, ctl, rp, ci, v, b, m):
    """The drop-in clsparseDcsrcg / clsparseScsrcg (the reference's CG API:
    Jacobi, L1 relative residual, x0 = 0) on the same matrix, with A in
    diagonal storage (default) and in CSR (CLSPARSE_AMD_DIA=0). Bytes of one
    iteration as executed from diagonal storage: SpMV 4 lower diagonals + p +
    y, update 8 vectors, p update 3 vectors (S = sizeof(T) per value)."""
    import numpy as np
    lib = cls.load()
    res = {}
    for name, dtype, fn, tol in (("fp64", np.float64, lib.clsparseDcsrcg, 1e-8),
                                 ("fp32", np.float32, lib.clsparseScsrcg, 1e-6)):
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        A = cls.CsrMatrix(rp, ci, v.astype(dtype), m, dtype=dtype)
        A.meta_create(ctl)
        bt = torch.tensor(b.astype(dtype), device="cuda")

        def run():
            best = None
            for _ in range(2):  # the second solve is reported
                x = torch.zeros(m, dtype=tdt, device="cuda")
                sc = cls.SolverControl(cls.DIAGONAL, 20000, tol, 0.0)
                xd, bd = cls.DenseVector(x), cls.DenseVector(bt)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle, ctl.handle)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                s = sc.stats()
                sc.release()
                cls.check(name, st)
                it = s.iterations
                best = {"seconds": round(dt, 5), "iterations": it, "l1_residual": s.l1_residual,
                        "us_per_iteration": round(dt / it * 1e6, 2), "a_diagonals": s.a_diagonals}
            return best
        S = np.dtype(dtype).itemsize
        r = run()
        r["bytes_per_iteration"] = (4 * S + 2 * S + 8 * S + 3 * S) * m
        r["GBps"] = round(r["bytes_per_iteration"] * r["iterations"] / r["seconds"] / 1e9, 1)
        with csr_only():
            r["csr"] = run()
        r["tolerance_l1_relative"] = tol
        res[name] = r
        log(f"reference-API CG {name}: {r}")
        A.meta_delete()
    return res


def warm_solvers(cls, torch, ctl, net, fs, n1d=20):
    """One small solve of every kind (factor build, network, pilot race,
    Jacobi mixed, FP64), so that the timed first solves measure the factor's
    own costs and not the process's first use of the library's kernels."""
    rp, ci, v, m = fs.gen_target3d(n1d)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    b = torch.tensor(cls.gen_uniform(m, -1.0, 1.0, seed=5), device="cuda")
    f = cls.Fsai(A, ctl)
    f.apply_net(net, ctl)
    lib = cls.load()
    for fn, with_f in ((lib.clsparseAmdDcsrcgFsai, True), (lib.clsparseAmdDcsrcgMixed, False),
                       (lib.clsparseAmdDcsrcgRefined, False)):
        x = torch.zeros(m, dtype=torch.float64, device="cuda")
        sc = cls.SolverControl(cls.DIAGONAL, 5000, 1e-8, 0.0)
        xd, bd = cls.DenseVector(x), cls.DenseVector(b)
        args = (ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s)) + \
            ((f.handle,) if with_f else ()) + (sc.handle, ctl.handle)
        cls.check("warm-up solve", fn(*args))
        sc.release()
    torch.cuda.synchronize()
    f.release()
    A.meta_delete()


def fsai_workload(cls, torch, ctl, n1d=160, tol=1e-8, weights=None, cpu=True, sweep=True):
    """Learned-FSAI mixed PCG vs Jacobi (mixed, FP64) on the held-out 3-D
    problem: from diagonal storage (A and the factor) and with A, G and G^T on
    CSR-Adaptive; G's values in BF16 and in FP32; the solution as an FP32
    increment and in FP64. Median of 5 repeat solves each, the first solve
    and the factor's set-up reported beside them."""
    from clsparse_amd import fsai as fs
    path = weights or fs.DEFAULT_WEIGHTS
    if not os.path.exists(path):
        return {"error": f"no trained weights at {os.path.relpath(path, ROOT)}"}
    params, rec = fs.load_weights(path)
    t0 = time.time()
    rp, ci, v, m = fs.gen_target3d(n1d)
    b = cls.gen_uniform(m, -1.0, 1.0, seed=41)
    log(f"fsai target generated n={m} nnz={int(rp[-1])} in {time.time() - t0:.1f}s")
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    bt = torch.tensor(b, device="cuda")
    lib = cls.load()
    net = cls.Mlp(seed=1)
    net.set_params(params)
    warm_solvers(cls, torch, ctl, net, fs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f = cls.Fsai(A, ctl)
    t1 = time.perf_counter()
    f.apply_net(net, ctl)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    setup = t2 - t0
    out = {"workload": f"diffusion3d_{n1d}^3_az100_lognormal1", "n": m, "nnz": int(rp[-1]),
           "nnz_G": f.nnz(), "tolerance": tol, "weights_sha256": rec["sha256_float32_le"][:16],
           "fsai_pattern_setup_s": round(t1 - t0, 4), "fsai_network_apply_s": round(t2 - t1, 5),
           "timing": "median of 5 solves after a first one (reported apart)",
           "metadata_create_s": meta_create_seconds(cls, torch, ctl, rp, ci, v, m)}
    F = lib.clsparseAmdDcsrcgFsai
    csr = {"CLSPARSE_AMD_DIA": "0", "CLSPARSE_AMD_FSAI_DIA": "0"}
    runs = [("learned_fsai_mixed", F, True, {}),
            ("learned_fsai_mixed_csr", F, True, csr),
            ("learned_fsai_mixed_fp32_g_csr", F, True, dict(csr, CLSPARSE_AMD_FSAI_BF16="0")),
            ("learned_fsai_mixed_fp64_x", F, True, {"CLSPARSE_AMD_MIXED_X64": "1"}),
            ("learned_fsai_mixed_fp32_g_fp64_x_csr", F, True,
             dict(csr, CLSPARSE_AMD_FSAI_BF16="0", CLSPARSE_AMD_MIXED_X64="1")),
            ("jacobi_mixed", lib.clsparseAmdDcsrcgMixed, False, {}),
            ("fp64_jacobi_pcg", lib.clsparseAmdDcsrcgRefined, False, {}),
            ("fp64_jacobi_pcg_csr", lib.clsparseAmdDcsrcgRefined, False, {"CLSPARSE_AMD_DIA": "0"})]
    for name, fn, with_f, kv in runs:
        with env(**kv):
            if with_f:
                f.apply_net(net, ctl)  # a fresh factor: the first solve judges it (estimate or race)
            out[name] = timed_solves(cls, torch, ctl, fn, A, bt, m, tol, fsai=f if with_f else None)
        log(f"{name}: {out[name]}")
    f.apply_net(None, ctl)  # zero network: the same solver with G = D^-1/2
    out["untrained_fsai_mixed"] = timed_solves(cls, torch, ctl, F, A, bt, m, tol, fsai=f)
    log(f"untrained FSAI mixed: {out['untrained_fsai_mixed']}")
    sec = {k: out[k]["seconds"] for k, _, _, _ in runs}
    base, base_csr = sec["fp64_jacobi_pcg"], sec["fp64_jacobi_pcg_csr"]
    first = out["learned_fsai_mixed"]["first_solve"]["seconds"]
    first_csr = out["learned_fsai_mixed_csr"]["first_solve"]["seconds"]
    out["speedup"] = {
        "learned_vs_fp64_jacobi_diagonal_storage": round(base / sec["learned_fsai_mixed"], 3),
        "learned_vs_fp64_jacobi_csr_adaptive_both": round(base_csr / sec["learned_fsai_mixed_csr"], 3),
        "learned_diagonal_storage_vs_fp64_jacobi_csr_adaptive": round(base_csr / sec["learned_fsai_mixed"], 3),
        "learned_fp32_g_vs_fp64_jacobi_csr_adaptive_both": round(base_csr / sec["learned_fsai_mixed_fp32_g_csr"], 3),
        "learned_fp64_x_vs_fp64_jacobi_diagonal_storage": round(base / sec["learned_fsai_mixed_fp64_x"], 3),
        "learned_fp32_g_fp64_x_vs_fp64_jacobi_csr_adaptive_both":
            round(base_csr / sec["learned_fsai_mixed_fp32_g_fp64_x_csr"], 3),
        "first_solve_with_setup_vs_fp64_jacobi_diagonal_storage": round(base / (first + setup), 3),
        "first_solve_with_setup_vs_fp64_jacobi_csr_adaptive_both": round(base_csr / (first_csr + setup), 3),
        "jacobi_mixed_vs_fp64_jacobi": round(base / sec["jacobi_mixed"], 3),
        "setup_s": round(setup, 5),
        "what": "FP64 Jacobi-PCG median solve time over learned-FSAI mixed PCG median solve time; "
                "first_solve_with_setup adds the factor's pattern build and network evaluation to "
                "the learned solver's first solve (its loss estimate or pilot race included)"}
    out["iteration_ratio_learned_vs_untrained"] = round(
        out["learned_fsai_mixed"]["iterations"] / out["untrained_fsai_mixed"]["iterations"], 3)
    if cpu:
        out["cpu_baseline"] = cpu_cg_baseline(rp, ci, v, b, tol, reps=3)
        log(f"CPU baseline: {out['cpu_baseline']}")
    f.release()
    A.meta_delete()
    if sweep:
        out["held_out_sweep"] = fsai_sweep(cls, torch, ctl, net, tol)
    net.release()
    return out


def fsai_sweep(cls, torch, ctl, net, tol):
    """The same comparison on further held-out 3-D problems (other anisotropy,
    coefficient spread and seeds; 128^3): learned-FSAI mixed PCG (first solve
    with the factor's set-up, and the median of 3 repeat solves) against FP64
    Jacobi-PCG, both from diagonal storage."""
    lib = cls.load()
    res = []
    for az, sigma, seed in ((100.0, 1.0, 61), (10.0, 1.5, 62), (1000.0, 0.5, 63), (1.0, 1.0, 64)):
        rp, ci, v, m = cls.gen_diffusion3d(128, 128, 128, sigma=sigma, ax=1.0, ay=1.0, az=az, seed=seed)
        b = torch.tensor(cls.gen_uniform(m, -1.0, 1.0, seed=seed + 100), device="cuda")
        A = cls.CsrMatrix(rp, ci, v, m)
        A.meta_create(ctl)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = cls.Fsai(A, ctl)
        f.apply_net(net, ctl)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        lf = timed_solves(cls, torch, ctl, lib.clsparseAmdDcsrcgFsai, A, b, m, tol, reps=3, fsai=f)
        jb = timed_solves(cls, torch, ctl, lib.clsparseAmdDcsrcgRefined, A, b, m, tol, reps=3)
        res.append({"az": az, "sigma": sigma, "seed": seed, "n": m,
                    "learned_s": lf["seconds"], "learned_first_s": lf["first_solve"]["seconds"],
                    "setup_s": round(setup, 5), "learned_iterations": lf["iterations"],
                    "preconditioner": lf["preconditioner"],
                    "fp64_jacobi_s": jb["seconds"], "fp64_jacobi_iterations": jb["iterations"],
                    "converged": lf["converged"] and jb["converged"],
                    "speedup": round(jb["seconds"] / lf["seconds"], 3),
                    "speedup_first_solve_with_setup": round(jb["seconds"] / (lf["first_solve"]["seconds"] + setup), 3),
                    "pilot_rates": lf["first_solve"].get("pilot_rates")})
        log(f"held-out sweep: {res[-1]}")
        f.release()
        A.meta_delete()
    return res


def train_slice(cls, torch, ctl, stream, cpu=True, matrices=24, steps_per_matrix=4, held=4):
    """A bounded slice of the learned-FSAI training on the 512^2 family: a new
    network trained on `matrices` family members, the held-out loss before
    (zero network = Jacobi) and after, steps per second, the network's forward
    MFMA rate, and one training step checked against the CPU oracle."""
    import numpy as np
    import scipy.sparse as sp

    from clsparse_amd import fsai as fs
    held_ids = list(range(2000 - held, 2000))
    net = cls.Mlp(seed=5)
    loss0 = fs.evaluate(None, ctl, held_ids)
    rec = fs.train(net, ctl, list(range(matrices)), steps_per_matrix=steps_per_matrix, lr=2e-3)
    loss1 = fs.evaluate(net, ctl, held_ids)
    out = {"family": "aniso2d 512^2 (n = 262,144), corr 16, sigma 1, log10 ratio U[0,3], theta U[0,pi)",
           "train_matrices": matrices, "steps": rec["steps"],
           "steps_per_s": round(rec["steps"] / max(rec["train_step_seconds"], 1e-9), 1),
           "held_out_loss_jacobi_mean": round(float(np.mean(loss0)), 5),
           "held_out_loss_after_mean": round(float(np.mean(loss1)), 5), "held_out_ids": held_ids}
    # Network forward (BF16 MFMA) with G's assembly on one family matrix.
    params, _ = fs.load_weights()
    netw = cls.Mlp(seed=1)
    netw.set_params(params)
    rp, ci, v, m = fs.gen_family(7)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    f = cls.Fsai(A, ctl)
    ms = median_call_ms(torch, stream, lambda: f.apply_net(netw, ctl), calls=10)
    flops = 10240 * f.nnz()
    out["forward_with_assembly_us"] = round(ms * 1e3, 1)
    out["forward_TFLOPs"] = round(flops / (ms * 1e-3) / 1e12, 1)
    out["forward_flops_rule"] = "2 (64*16 + 64*64) per nonzero of G; the time includes G's assembly"
    f.release()
    A.meta_delete()
    # One step against the oracle on a 128^2 member-sized matrix of the family's generator.
    if cpu:
        from oracle import fsai as of
        p = fs.aniso2d_params(11)
        rp, ci, v, m = cls.gen_aniso2d(128, 128, corr_len=p["corr_len"], sigma=p["sigma"],
                                       ratio=p["ratio"], theta=p["theta"], seed=p["seed"])
        A = cls.CsrMatrix(rp, ci, v, m)
        A.meta_create(ctl)
        f = cls.Fsai(A, ctl)
        loss, gG = f.train_step(netw, A, ctl, probes=8, seed=99, lr=0.0, want_grad=True)
        o = f.output().astype(np.float64)
        X = of.bf16_bits_to_f32(f.features()).reshape(-1, 16)
        g_dev = netw.grad().astype(np.float64)
        f.release()
        A.meta_delete()
        G = of.assemble_g(rp, ci, v, o)
        Z = np.stack([of.rademacher(m, of.probe_seed(99, k)) for k in range(8)], axis=1)
        loss_ref, grad_ref = of.hutchinson(sp.csr_matrix((v, ci, rp), shape=(m, m)), G, Z)
        gi, gj, _ = of.tril_pattern(rp, ci)
        gref = np.asarray(grad_ref[gi, gj]).ravel()
        pg = of.mlp_backward(params, X, (gG * of.dhat(rp, ci, v)[gi]).astype(np.float32))
        out["gradient_check_128"] = {
            "loss_rel_err": abs(loss - loss_ref) / abs(loss_ref),
            "dL_dG_max_rel_err": float(np.max(np.abs(gG - gref)) / np.max(np.abs(gref))),
            "param_grad_max_err_over_max": float(np.max(np.abs(g_dev - pg)) / np.max(np.abs(pg))),
            "what": "one training step on a 128^2 matrix of the family's generator vs oracle/fsai.py"}
    netw.release()
    net.release()
    log(f"training slice: {out}")
    return out


def fsai_train(cls, torch, ctl, args):
    from clsparse_amd import fsai as fs
    n_total, held = 2000, 200
    train_ids = list(range(min(args.train_matrices, n_total - held)))
    held_ids = list(range(n_total - held, n_total))[: args.eval_matrices]
    net = cls.Mlp(seed=args.seed)
    log(f"held-out Jacobi loss on {len(held_ids)} matrices")
    loss_jac = fs.evaluate(None, ctl, held_ids)
    rec = fs.train(net, ctl, train_ids, steps_per_matrix=args.steps_per_matrix, lr=args.lr,
                   log=log, time_budget_s=args.budget)
    loss_net = fs.evaluate(net, ctl, held_ids)
    params = net.params()
    record = {"family": "aniso2d 512^2, corr 16, sigma 1, log10 ratio U[0,3], theta U[0,pi)",
              "train": rec, "held_out_ids": [held_ids[0], held_ids[-1]] if held_ids else [],
              "held_out_loss_jacobi_mean": float(sum(loss_jac) / max(len(loss_jac), 1)),
              "held_out_loss_learned_mean": float(sum(loss_net) / max(len(loss_net), 1)),
              "seed": args.seed}
    digest = fs.save_weights(args.weights or fs.DEFAULT_WEIGHTS, params, record)
    record["sha256_float32_le"] = digest
    net.release()
    return record


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-solver", action="store_true", help="skip the solver workload")
    ap.add_argument("--mode", default="spmv",
                    choices=["spmv", "copy", "gather", "locality", "cg", "fsai", "fsai-train", "next"])
    ap.add_argument("--no-fsai", action="store_true", help="skip the learned-FSAI workload")
    ap.add_argument("--no-sweep", action="store_true", help="--mode fsai: skip the held-out 128^3 sweep")
    ap.add_argument("--weights", default=None, help="learned-FSAI weights (.npy)")
    ap.add_argument("--train-matrices", type=int, default=1800)
    ap.add_argument("--eval-matrices", type=int, default=200)
    ap.add_argument("--steps-per-matrix", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--budget", type=float, default=None, help="fsai-train time budget (s)")
    ap.add_argument("--set", default="21,18,15,12,s64,s32",
                    help="locality mode: power-law column windows (bits) and stencils (s64, s32)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(one process drives one GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import clsparse_amd as cls

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()

    if args.mode == "gather":
        with cls.Control(stream=stream) as c:
            print(json.dumps({"mode": "gather", "floors": gather_floors(cls, torch, c, stream, 34467246),
                              "copy_peak": copy_peak(cls, torch, c, stream)}))
        return
    if args.mode == "copy":
        n = 64 << 20
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        with cls.Control() as c:
            r, y, one = cls.DenseVector(dst), cls.DenseVector(src), cls.Scalar(1.0)
            lib = cls.load()
            for _ in range(args.warmup + args.steps):
                cls.check("cldenseDscale", lib.cldenseDscale(ctypes.byref(r.s), ctypes.byref(one.s),
                                                            ctypes.byref(y.s), c.handle))
        torch.cuda.synchronize()
        assert torch.equal(dst, src)
        print(json.dumps({"mode": "copy", "bytes_read_per_launch": n * 8,
                          "bytes_written_per_launch": n * 8, "launches": args.warmup + args.steps}))
        return

    ctl = cls.Control(stream=stream, async_mode=True)

    if args.mode == "next":
        print(json.dumps({"mode": "next", "copy_peak": copy_peak(cls, torch, ctl, stream),
                          "next": next_rows(cls, torch, ctl, stream, cpu=not args.no_cpu)}))
        ctl.release()
        return
    if args.mode == "cg":
        print(json.dumps({"mode": "cg", "solver": solver_workload(cls, torch, ctl)}))
        ctl.release()
        return
    if args.mode == "fsai":
        print(json.dumps({"mode": "fsai", "fsai": fsai_workload(cls, torch, ctl, weights=args.weights,
                                                                 cpu=not args.no_cpu,
                                                                 sweep=not args.no_sweep)}))
        ctl.release()
        return
    if args.mode == "fsai-train":
        print(json.dumps({"mode": "fsai-train", "record": fsai_train(cls, torch, ctl, args)}))
        ctl.release()
        return

    t0 = time.time()
    rp, ci, v, m = cls.gen_powerlaw(1 << 21, 16.0, 1.5, 200000, 8, 100000, seed=1)
    x = cls.gen_uniform(m, -1.0, 1.0, seed=2)
    n, nnz = rp.size - 1, int(rp[-1])
    log(f"generated n={n} nnz={nnz} in {time.time() - t0:.1f}s")

    if args.mode == "locality":
        res = []
        items = args.set.split(",")
        for it in items:
            if it[0] in "st":
                continue
            bits = int(it)
            cif = (ci & np.uint32((1 << bits) - 1)) if bits < 21 else ci
            run = SpmvRun(cls, torch, rp, cif, v, m, x, np.float64, ctl)
            wall, launch = run.time(args.steps, args.warmup, stream)
            floor_ms, _ = run.sum_ax(args.steps, stream)
            res.append({"matrix": f"powerlaw cols folded to 2^{bits}",
                        "x_window_bytes": 8 << bits, "launch_us": round(launch * 1e3, 2),
                        "GBps": round(run.nbytes / (launch * 1e-3) / 1e9, 1),
                        "sum_ax_floor_us": round(floor_ms * 1e3, 2)})
            run.close()
            log(str(res[-1]))
        # Stencils: "s64" / "s32" (FP64 / FP32, 128^3), "s32:160" (size), and
        # "t32:160" for the lower triangle (the pattern of the FSAI factor).
        import scipy.sparse as sp
        stencils = [it for it in items if it[0] in "st"]
        cache = {}
        for it in stencils:
            kind, _, size = it.partition(":")
            n1d = int(size) if size else 128
            dt = np.float64 if kind[1:] == "64" else np.float32
            if n1d not in cache:
                rp3, ci3, v3, m3 = cls.gen_diffusion3d(n1d, n1d, n1d, sigma=1.0, ax=1.0, ay=1.0,
                                                        az=100.0 if size else 1.0, seed=3)
                cache[n1d] = (rp3, ci3, v3, m3, cls.gen_uniform(m3, -1.0, 1.0, seed=5))
            rp3, ci3, v3, m3, x3 = cache[n1d]
            if kind[0] == "t":
                lo = sp.tril(sp.csr_matrix((v3, ci3, rp3), shape=(m3, m3))).tocsr()
                lo.sort_indices()
                rp3, ci3, v3 = lo.indptr.astype(np.uint32), lo.indices.astype(np.uint32), lo.data
            run = SpmvRun(cls, torch, rp3, ci3, v3, m3, x3, dt, ctl)
            wall, launch = run.time(args.steps, args.warmup, stream)
            what = "tril of " if kind[0] == "t" else ""
            res.append({"matrix": f"{what}diffusion3d {n1d}^3 {np.dtype(dt).name}",
                        "nnz": int(rp3[-1]), "launch_us": round(launch * 1e3, 2),
                        "GBps": round(run.nbytes / (launch * 1e-3) / 1e9, 1)})
            run.close()
            log(str(res[-1]))
        print(json.dumps({"mode": "locality", "results": res}))
        ctl.release()
        return

    run = SpmvRun(cls, torch, rp, ci, v, m, x, np.float64, ctl)
    ms_per_step, launch_ms = run.time(args.steps, args.warmup, stream)
    nbytes = run.nbytes
    value = nbytes / (ms_per_step * 1e-3) / 1e9
    achieved = nbytes / (launch_ms * 1e-3) / 1e9

    # Parity spot check of the last result against the oracle on a row sample.
    import oracle
    yg = run.y.t.cpu().numpy()
    rows = np.sort(np.random.default_rng(0).choice(n, 4096, replace=False))
    sub_rp = np.concatenate([[0], np.cumsum(np.diff(rp.astype(np.int64))[rows])]).astype(np.uint32)
    sub_ci = np.concatenate([ci[rp[r]:rp[r + 1]] for r in rows])
    sub_v = np.concatenate([v[rp[r]:rp[r + 1]] for r in rows])
    yref, scale = oracle.csrmv(sub_rp, sub_ci, sub_v, x)
    max_err = float(np.max(np.abs(yg[rows] - yref) / np.maximum(scale, 1e-300)))
    # Floor of any CSR SpMV that gathers x directly: the same value, column and
    # x loads in storage order, with no rows and no y (1^T A x, which also
    # checks the sum of y).
    ref_s, ref_kept = reference_methodology(run.step, torch)
    reference_method = {"us_per_call": round(ref_s * 1e6, 2), "GBps": round(nbytes / ref_s / 1e9, 1),
                        "samples_kept": ref_kept,
                        "what": "20 host-timed samples of 20 back-to-back clsparseDcsrmv calls, "
                                "samples beyond 1 sigma dropped (the reference's benchmark method)"}
    meta_s = meta_create_seconds(cls, torch, ctl, rp, ci, v, m)
    floor_ms, sum_ax = run.sum_ax(20, stream)
    ysum = float(np.sum(yg))
    gather_floor = {"sum_ax_us": round(floor_ms * 1e3, 2), "spmv_launch_us": round(launch_ms * 1e3, 2),
                    "spmv_over_floor": round(launch_ms / floor_ms, 3),
                    "checksum_rel_diff": abs(sum_ax - ysum) / max(float(np.sum(np.abs(yg))), 1e-300),
                    "what": "clsparseAmdDcsrSumAx on the same matrix and x: values, columns and "
                            "x gathers in storage order, no row structure, no y"}
    run.close()

    cpu = None
    if not args.no_cpu:
        med, threads, calls = cpu_baseline(rp, ci, v, x)
        cpu = {"value": round(nbytes / med / 1e9, 3), "unit": "GB/s", "cores": threads,
               "kind": "port",
               "sample": f"{calls} calls of the same FP64 power-law SpMV (full matrix), median; "
                         f"oracle OpenMP CSR SpMV, {threads} threads"}

    extra = {"reference_methodology": reference_method,
             "metadata_create": {"seconds": meta_s, "spmv_calls_equivalent": round(meta_s / (launch_ms * 1e-3), 1),
                                 "what": "clsparseCsrMetaCreate on the power-law matrix, median of 3"},
             "gather_floor": gather_floor, "copy_peak": copy_peak(cls, torch, ctl, stream),
             "poisson_fp32": poisson_workload(cls, torch, ctl, stream, cpu=not args.no_cpu)}
    if not args.no_solver:
        extra["solver"] = solver_workload(cls, torch, ctl, cpu=not args.no_cpu)
    if not args.no_fsai:
        extra["fsai"] = fsai_workload(cls, torch, ctl, weights=args.weights, cpu=not args.no_cpu)
        extra["fsai_train"] = train_slice(cls, torch, ctl, stream, cpu=not args.no_cpu)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "powerlaw_fp64_csr_adaptive_spmv", "n": n, "ncols": m,
                   "nnz": nnz, "row_lengths": "lognormal(mean 16, sigma 1.5) clip [1,200000] "
                   "+ 8 rows x 100000", "seed": 1, "bytes_per_step": nbytes,
                   "hbm_fraction_of_value": round(value / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic()},
        "cpu_baseline": cpu,
        "parity": {"max_row_rel_err_sample": max_err, "bar": 1e-13},
        "timing": {"launch_ms_median": round(launch_ms, 5)},
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    ctl.release()


if __name__ == "__main__":
    main()
