"""ctypes wrapper of the CPU oracle (oracle.c). TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker; the product path never touches it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int, c_size_t, c_uint32, c_uint64

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class CgResult(Structure):
    _fields_ = [("status", c_int), ("n_iters", c_int), ("executed", c_int),
                ("early_exit", c_int), ("initial_residual", c_double),
                ("current_residual", c_double), ("rel_tol_out", c_double),
                ("abs_tol_out", c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built; run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        u32p, u64p, f64p = POINTER(c_uint32), POINTER(c_uint64), POINTER(c_double)
        f32p = POINTER(ctypes.c_float)
        lib.orc_row_blocks.restype = c_uint64
        lib.orc_row_blocks.argtypes = [u32p, c_uint32, u64p]
        lib.orc_csrmv_d.restype = None
        lib.orc_csrmv_d.argtypes = [c_uint32, u32p, u32p, f64p, f64p, c_double, c_double, f64p,
                                    f64p, f64p]
        lib.orc_csrmv_s.restype = None
        lib.orc_csrmv_s.argtypes = [c_uint32, u32p, u32p, f32p, f32p, ctypes.c_float,
                                    ctypes.c_float, f32p, f64p, f64p]
        lib.orc_spmv_omp_d.restype = None
        lib.orc_spmv_omp_d.argtypes = [c_uint32, u32p, u32p, f64p, f64p, f64p, c_int]
        lib.orc_spmv_omp_s.restype = None
        lib.orc_spmv_omp_s.argtypes = [c_uint32, u32p, u32p, f32p, f32p, f32p, c_int]
        lib.orc_cg_omp_d.restype = c_int
        lib.orc_cg_omp_d.argtypes = [c_uint32, u32p, u32p, f64p, f64p, f64p, c_int, c_double, c_int,
                                     c_int, f64p]
        lib.orc_max_threads.restype = c_int
        for nm in ("orc_nrm1_d", "orc_nrm2_d"):
            getattr(lib, nm).restype = c_double
            getattr(lib, nm).argtypes = [c_size_t, f64p]
        lib.orc_dot_d.restype = c_double
        lib.orc_dot_d.argtypes = [c_size_t, f64p, f64p]
        lib.orc_cg_d.restype = CgResult
        lib.orc_cg_d.argtypes = [c_uint32, u32p, u32p, f64p, f64p, f64p, c_int, c_int, c_double,
                                 c_double, c_int]
        lib.orc_cg_s.restype = CgResult
        lib.orc_cg_s.argtypes = [c_uint32, u32p, u32p, f32p, f32p, f32p, c_int, c_int, c_double,
                                 c_double, c_int]
        lib.orc_csrmm_d.restype = None
        lib.orc_csrmm_d.argtypes = [c_uint32, u32p, u32p, f64p, f64p, c_uint32, f64p, c_uint32,
                                    c_uint32, c_double, c_double, f64p]
        lib.orc_bicgstab_d.restype = CgResult
        lib.orc_bicgstab_d.argtypes = lib.orc_cg_d.argtypes
        lib.orc_bicgstab_s.restype = CgResult
        lib.orc_bicgstab_s.argtypes = lib.orc_cg_s.argtypes
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def row_blocks(row_pointer, count_only=False):
    """(entries u64, rowBlockSize) of the reference binning; or the size-only count."""
    rp = np.ascontiguousarray(row_pointer, dtype=np.uint32)
    n = rp.size - 1
    lib = load()
    size = lib.orc_row_blocks(_p(rp, c_uint32), n, None)
    if count_only:
        return int(size)
    out = np.zeros(size // 2 + 2, dtype=np.uint64)
    created = lib.orc_row_blocks(_p(rp, c_uint32), n, _p(out, c_uint64))
    return out[: created // 2].copy(), int(created)


def csrmv(rp, ci, v, x, alpha=1.0, beta=0.0, y=None):
    """FP64 oracle of y = alpha A x + beta y -> (y, per-row magnitude scale)."""
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    n = rp.size - 1
    out = np.empty(n, dtype=np.float64)
    scale = np.empty(n, dtype=np.float64)
    lib = load()
    if np.asarray(v).dtype == np.float32:
        v = np.ascontiguousarray(v, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        yin = np.ascontiguousarray(y if y is not None else np.zeros(n), dtype=np.float32)
        lib.orc_csrmv_s(n, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, ctypes.c_float),
                        _p(x, ctypes.c_float), alpha, beta, _p(yin, ctypes.c_float),
                        _p(out, c_double), _p(scale, c_double))
    else:
        v = np.ascontiguousarray(v, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        yin = np.ascontiguousarray(y if y is not None else np.zeros(n), dtype=np.float64)
        lib.orc_csrmv_d(n, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, c_double), _p(x, c_double),
                        alpha, beta, _p(yin, c_double), _p(out, c_double), _p(scale, c_double))
    return out, scale


def spmv_omp(rp, ci, v, x, y, threads):
    load().orc_spmv_omp_d(rp.size - 1, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, c_double),
                          _p(x, c_double), _p(y, c_double), threads)


def spmv_omp_s(rp, ci, v, x, y, threads):
    load().orc_spmv_omp_s(rp.size - 1, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, ctypes.c_float),
                          _p(x, ctypes.c_float), _p(y, ctypes.c_float), threads)


def cg_omp(rp, ci, v, b, max_iters, rel_tol, crit=1, threads=1):
    """FP64 Jacobi-PCG from x = 0 on `threads` host threads until the residual
    (crit 1: L2, 0: the reference's L1) relative to b is <= rel_tol
    -> (x, iterations, final residual). A timing baseline."""
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(rp.size - 1)
    res = c_double(0.0)
    it = load().orc_cg_omp_d(rp.size - 1, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, c_double),
                             _p(b, c_double), _p(x, c_double), max_iters, rel_tol, crit, threads,
                             ctypes.byref(res))
    return x, int(it), res.value


def max_threads() -> int:
    return int(load().orc_max_threads())


def dot(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return load().orc_dot_d(x.size, _p(x, c_double), _p(y, c_double))


def nrm1(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return load().orc_nrm1_d(x.size, _p(x, c_double))


def nrm2(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return load().orc_nrm2_d(x.size, _p(x, c_double))


def csrmm(rp, ci, v, B, C, alpha=1.0, beta=0.0):
    """C = alpha A B + beta C for row-major dense B (k x m) and C (n x m)
    -> (C_out, scale) in FP64 (long-double sums)."""
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    out = np.array(C, dtype=np.float64, copy=True, order="C")
    n, m = out.shape
    scale = np.zeros((n, m))
    load().orc_csrmm_d(n, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, c_double), _p(B, c_double),
                       B.shape[1], _p(out, c_double), m, m, alpha, beta, _p(scale, c_double))
    return out, scale


def bicgstab(rp, ci, v, b, x0, precond=1, max_iters=100, rel_tol=1e-2, abs_tol=1e-5,
             n_iters_prior=0):
    """CPU restatement of clsparse[S/D]csrbicgStab -> (x, CgResult)."""
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    n = rp.size - 1
    lib = load()
    if np.asarray(v).dtype == np.float32:
        t, fn, dt = ctypes.c_float, lib.orc_bicgstab_s, np.float32
    else:
        t, fn, dt = c_double, lib.orc_bicgstab_d, np.float64
    v = np.ascontiguousarray(v, dtype=dt)
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.array(x0, dtype=dt, copy=True)
    r = fn(n, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, t), _p(b, t), _p(x, t), precond,
           max_iters, rel_tol, abs_tol, n_iters_prior)
    return x, r


def cg(rp, ci, v, b, x0, precond=1, max_iters=100, rel_tol=1e-2, abs_tol=1e-5, n_iters_prior=0):
    """CPU restatement of clsparse[S/D]csrcg -> (x, CgResult)."""
    rp = np.ascontiguousarray(rp, dtype=np.uint32)
    ci = np.ascontiguousarray(ci, dtype=np.uint32)
    n = rp.size - 1
    lib = load()
    if np.asarray(v).dtype == np.float32:
        t, fn = ctypes.c_float, lib.orc_cg_s
        dt = np.float32
    else:
        t, fn = c_double, lib.orc_cg_d
        dt = np.float64
    v = np.ascontiguousarray(v, dtype=dt)
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.array(x0, dtype=dt, copy=True)
    r = fn(n, _p(rp, c_uint32), _p(ci, c_uint32), _p(v, t), _p(b, t), _p(x, t), precond,
           max_iters, rel_tol, abs_tol, n_iters_prior)
    return x, r
