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


def cpu_baseline(rp, ci, v, x, budget_s=3.0, max_calls=30):
    import numpy as np

    import oracle
    threads = oracle.max_threads()
    y = np.empty(rp.size - 1)
    oracle.spmv_omp(rp, ci, v, x, y, threads)  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < max_calls and (len(times) < 3 or time.perf_counter() < t_end):
        t0 = time.perf_counter()
        oracle.spmv_omp(rp, ci, v, x, y, threads)
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2], threads, len(times)


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
        med, threads, calls = cpu_baseline(rp, ci, v, x, budget_s=1.0, max_calls=50)
        out["cpu_baseline"] = {"us_per_call": round(med * 1e6, 1), "cores": threads, "kind": "port",
                               "sample": f"{calls} calls, oracle OpenMP CSR SpMV (FP64 values)"}
    return out


def cpu_cg_baseline(rp, ci, v, b, iters_needed, sample_iters=24):
    """The oracle's FP64 Jacobi-CG (single thread, the reference's algorithm) on
    `sample_iters` iterations of the same system; seconds per iteration and the
    projection to the GPU solve's iteration count."""
    import numpy as np

    import oracle
    t0 = time.perf_counter()
    _, r = oracle.cg(rp, ci, v, b, np.zeros(rp.size - 1), 1, sample_iters, 1e-300, 0.0)
    dt = time.perf_counter() - t0
    per = dt / max(r.executed, 1)
    return {"kind": "port", "cores": 1, "seconds_per_iteration": round(per, 5),
            "projected_seconds": round(per * iters_needed, 3),
            "sample": f"{r.executed} iterations of the oracle's FP64 Jacobi-CG on the same matrix, "
                      "single thread, projected to the FP64 GPU solve's iteration count"}


def solver_workload(cls, torch, ctl, n1d=128, tol=1e-10, cpu=True):
    """Mixed-precision refinement vs FP64 Jacobi-PCG to an FP64 true relative residual."""
    import numpy as np
    rp, ci, v, m = cls.gen_diffusion3d(n1d, n1d, n1d, sigma=1.0, seed=3)
    b = cls.gen_uniform(m, -1.0, 1.0, seed=4)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    bt = torch.tensor(b, device="cuda")
    lib = cls.load()
    out = {"workload": f"diffusion3d_{n1d}^3_lognormal1_jacobi", "n": m, "nnz": int(rp[-1]),
           "tolerance": tol}
    def run(name, fn):
        best = None
        for _ in range(2):  # second solve is the timed one (first builds graphs, warms caches)
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            sc = cls.SolverControl(cls.DIAGONAL, 20000, tol, 0.0)
            xd, bd = cls.DenseVector(x), cls.DenseVector(bt)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle,
                    ctl.handle)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            s = sc.stats()
            sc.release()
            cls.check(name, st)
            best = {"seconds": round(dt, 5), "iterations": s.iterations,
                    "refinements": s.outer_iterations, "true_rel_residual": s.true_residual,
                    "a_diagonals": s.a_diagonals}
        out[name] = best

    run("mixed_fp32_inner", lib.clsparseAmdDcsrcgMixed)
    run("fp64_jacobi_pcg", lib.clsparseAmdDcsrcgRefined)
    with csr_only():
        run("fp64_jacobi_pcg_csr", lib.clsparseAmdDcsrcgRefined)
    out["speedup_mixed_vs_fp64"] = round(out["fp64_jacobi_pcg"]["seconds"] /
                                         out["mixed_fp32_inner"]["seconds"], 3)
    out["speedup_mixed_vs_fp64_csr"] = round(out["fp64_jacobi_pcg_csr"]["seconds"] /
                                             out["mixed_fp32_inner"]["seconds"], 3)
    if cpu:
        out["cpu_baseline"] = cpu_cg_baseline(rp, ci, v, b, out["fp64_jacobi_pcg"]["iterations"])
    out["reference_api_cg"] = reference_api_cg(cls, torch, ctl, rp, ci, v, b, m)
    out["blas1_fp64"] = blas1_workload(cls, torch, ctl, ctl.stream)
    A.meta_delete()
    return out


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


class csr_only:
    """Within the block the solvers keep A in CSR (CLSPARSE_AMD_DIA=0): the
    FP64 baseline with the reference's SpMV."""

    def __enter__(self):
        self.old = os.environ.get("CLSPARSE_AMD_DIA")
        os.environ["CLSPARSE_AMD_DIA"] = "0"

    def __exit__(self, *exc):
        if self.old is None:
            del os.environ["CLSPARSE_AMD_DIA"]
        else:
            os.environ["CLSPARSE_AMD_DIA"] = self.old


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


def fsai_workload(cls, torch, ctl, n1d=160, tol=1e-8, weights=None, cpu=True):
    """Learned-FSAI mixed PCG vs Jacobi (mixed, FP64) on the held-out 3-D problem."""
    import numpy as np

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
    out = {"workload": f"diffusion3d_{n1d}^3_az100_lognormal1", "n": m, "nnz": int(rp[-1]),
           "nnz_G": f.nnz(), "tolerance": tol, "weights_sha256": rec["sha256_float32_le"][:16],
           "fsai_pattern_setup_s": round(t1 - t0, 4), "fsai_network_apply_s": round(t2 - t1, 5)}

    def solve(fn, with_fsai):
        best = None
        first = None
        for _ in range(2):  # the second solve is the reported one (the first captures graphs
            # and, for the learned factor, runs the pilot race whose outcome the factor keeps)
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            sc = cls.SolverControl(cls.DIAGONAL, 50000, tol, 0.0)
            xd, bd = cls.DenseVector(x), cls.DenseVector(bt)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            if with_fsai:
                st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), f.handle,
                        sc.handle, ctl.handle)
            else:
                st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle,
                        ctl.handle)
            torch.cuda.synchronize()
            dt = time.perf_counter() - ts
            s = sc.stats()
            sc.release()
            cls.check("solve", st)
            best = {"seconds": round(dt, 5), "iterations": s.iterations,
                    "replacements": s.outer_iterations, "true_rel_residual": s.true_residual,
                    "converged": s.status_detail == 0, "preconditioner": s.preconditioner,
                    "a_diagonals": s.a_diagonals, "g_diagonals": s.g_diagonals}
            if first is None:
                first = {"seconds": best["seconds"], "iterations": best["iterations"]}
                if with_fsai and s.pilot_rate_fsai > 0:
                    # the first solve's pilot race: -ln(residual ratio) per second
                    first["pilot_rates"] = {"fsai": round(s.pilot_rate_fsai, 1),
                                            "jacobi": round(s.pilot_rate_jacobi, 1)}
        best["first_solve"] = first
        return best

    out["learned_fsai_mixed"] = solve(lib.clsparseAmdDcsrcgFsai, True)
    log(f"learned FSAI mixed: {out['learned_fsai_mixed']}")
    out["jacobi_mixed"] = solve(lib.clsparseAmdDcsrcgMixed, False)
    log(f"Jacobi mixed: {out['jacobi_mixed']}")
    out["fp64_jacobi_pcg"] = solve(lib.clsparseAmdDcsrcgRefined, False)
    log(f"FP64 Jacobi: {out['fp64_jacobi_pcg']}")
    with csr_only():
        out["fp64_jacobi_pcg_csr"] = solve(lib.clsparseAmdDcsrcgRefined, False)
    log(f"FP64 Jacobi, CSR-Adaptive SpMV: {out['fp64_jacobi_pcg_csr']}")
    f.apply_net(None, ctl)  # zero network: the same solver with G = D^-1/2
    out["untrained_fsai_mixed"] = solve(lib.clsparseAmdDcsrcgFsai, True)
    log(f"untrained FSAI mixed: {out['untrained_fsai_mixed']}")
    base = out["fp64_jacobi_pcg"]["seconds"]
    out["speedup_learned_fsai_vs_fp64_jacobi"] = round(base / out["learned_fsai_mixed"]["seconds"], 3)
    out["speedup_learned_fsai_vs_fp64_jacobi_csr"] = round(
        out["fp64_jacobi_pcg_csr"]["seconds"] / out["learned_fsai_mixed"]["seconds"], 3)
    out["speedup_learned_fsai_first_solve_vs_fp64_jacobi"] = round(
        base / out["learned_fsai_mixed"]["first_solve"]["seconds"], 3)
    out["speedup_jacobi_mixed_vs_fp64_jacobi"] = round(base / out["jacobi_mixed"]["seconds"], 3)
    out["iteration_ratio_learned_vs_untrained"] = round(
        out["learned_fsai_mixed"]["iterations"] / out["untrained_fsai_mixed"]["iterations"], 3)
    if cpu:
        out["cpu_baseline"] = cpu_cg_baseline(rp, ci, v, b, out["fp64_jacobi_pcg"]["iterations"],
                                              sample_iters=12)
    f.release()
    A.meta_delete()
    out["held_out_sweep"] = fsai_sweep(cls, torch, ctl, net, tol)
    net.release()
    return out


def fsai_sweep(cls, torch, ctl, net, tol):
    """The same comparison on further held-out 3-D problems (other anisotropy,
    coefficient spread and seeds; 128^3): learned-FSAI mixed PCG (first and
    repeat solve) against FP64 Jacobi-PCG from diagonal storage."""
    lib = cls.load()
    res = []
    for az, sigma, seed in ((100.0, 1.0, 61), (10.0, 1.5, 62), (1000.0, 0.5, 63), (1.0, 1.0, 64)):
        rp, ci, v, m = cls.gen_diffusion3d(128, 128, 128, sigma=sigma, ax=1.0, ay=1.0, az=az, seed=seed)
        b = torch.tensor(cls.gen_uniform(m, -1.0, 1.0, seed=seed + 100), device="cuda")
        A = cls.CsrMatrix(rp, ci, v, m)
        A.meta_create(ctl)
        f = cls.Fsai(A, ctl)
        f.apply_net(net, ctl)

        def solve(fn, with_fsai):
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            sc = cls.SolverControl(cls.DIAGONAL, 50000, tol, 0.0)
            xd, bd = cls.DenseVector(x), cls.DenseVector(b)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if with_fsai:
                st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), f.handle, sc.handle,
                        ctl.handle)
            else:
                st = fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle, ctl.handle)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            stt = sc.stats()
            sc.release()
            cls.check("solve", st)
            return (dt, stt.iterations, stt.preconditioner, stt.status_detail == 0,
                    (round(stt.pilot_rate_fsai, 1), round(stt.pilot_rate_jacobi, 1)))

        first = solve(lib.clsparseAmdDcsrcgFsai, True)
        rep_ = solve(lib.clsparseAmdDcsrcgFsai, True)
        solve(lib.clsparseAmdDcsrcgRefined, False)  # warm
        base = solve(lib.clsparseAmdDcsrcgRefined, False)
        res.append({"az": az, "sigma": sigma, "seed": seed, "n": m,
                    "learned_s": round(rep_[0], 5), "learned_first_s": round(first[0], 5),
                    "learned_iterations": rep_[1], "preconditioner": rep_[2],
                    "fp64_jacobi_s": round(base[0], 5), "fp64_jacobi_iterations": base[1],
                    "converged": rep_[3] and base[3],
                    "speedup": round(base[0] / rep_[0], 3),
                    "speedup_first_solve": round(base[0] / first[0], 3),
                    "pilot_rates_fsai_jacobi": first[4]})
        log(f"held-out sweep: {res[-1]}")
        f.release()
        A.meta_delete()
    return res


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
                    choices=["spmv", "copy", "locality", "cg", "fsai", "fsai-train", "next"])
    ap.add_argument("--no-fsai", action="store_true", help="skip the learned-FSAI workload")
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
        print(json.dumps({"mode": "fsai", "fsai": fsai_workload(cls, torch, ctl,
                                                                 weights=args.weights)}))
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
            res.append({"matrix": f"powerlaw cols folded to 2^{bits}",
                        "x_window_bytes": 8 << bits, "launch_us": round(launch * 1e3, 2),
                        "GBps": round(run.nbytes / (launch * 1e-3) / 1e9, 1)})
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

    extra = {"gather_floor": gather_floor, "copy_peak": copy_peak(cls, torch, ctl, stream),
             "poisson_fp32": poisson_workload(cls, torch, ctl, stream, cpu=not args.no_cpu)}
    if not args.no_solver:
        extra["solver"] = solver_workload(cls, torch, ctl, cpu=not args.no_cpu)
    if not args.no_fsai:
        extra["fsai"] = fsai_workload(cls, torch, ctl, weights=args.weights, cpu=not args.no_cpu)

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
