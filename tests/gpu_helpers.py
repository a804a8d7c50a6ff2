"""Shared helpers of the GPU tests: device objects, oracle comparison."""
import ctypes
import sys

import numpy as np
import torch

import clsparse_amd as cls
import oracle

U32 = 2.0 ** -24
U64 = 2.0 ** -53


def say(msg: str) -> None:
    """Names the case before a GPU step (visible even if the step dies)."""
    print(msg, flush=True)
    sys.stdout.flush()


def rp_from_lengths(lengths):
    return np.concatenate([[0], np.cumsum(np.asarray(lengths, dtype=np.int64))]).astype(np.uint32)


def random_csr(rng, lengths, ncols):
    rp = rp_from_lengths(lengths)
    parts = []
    for k in lengths:
        k = int(k)
        if k == 0:
            continue
        if k * 4 < ncols:
            c = np.unique(rng.integers(0, ncols, int(k * 1.2) + 8))
            while c.size < k:
                c = np.unique(np.concatenate([c, rng.integers(0, ncols, k)]))
            c = np.sort(rng.choice(c, k, replace=False))
        else:
            c = np.sort(rng.choice(ncols, k, replace=False))
        parts.append(c)
    ci = np.concatenate(parts).astype(np.uint32) if parts else np.zeros(0, np.uint32)
    v = rng.uniform(-1, 1, int(rp[-1]))
    return rp, ci, v


def per_row_error(y_gpu, y_ref, scale):
    """|y - y_ref| / (|alpha| sum|a_ij x_j| + |beta y_i|); rows with zero scale must be exact."""
    y_gpu = np.asarray(y_gpu, dtype=np.float64)
    err = np.abs(y_gpu - y_ref)
    zero = scale == 0
    assert np.all(err[zero] == 0), "rows with no terms must be exactly zero"
    out = np.zeros_like(err)
    out[~zero] = err[~zero] / scale[~zero]
    return out


def spmv_check(rp, ci, v, ncols, dtype, *, alpha=1.0, beta=0.0, seed=1, adaptive=True,
               control=None, tol=None, y0=None, x=None, ext=False):
    """Runs clsparse[S/D]csrmv on the GPU and returns (max per-row error, y_gpu)."""
    n = rp.size - 1
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    if x is None:
        x = cls.gen_uniform(ncols, -1.0, 1.0, seed=seed)
    if y0 is None:
        y0 = cls.gen_uniform(n, -1.0, 1.0, seed=seed + 1) if beta != 0 else np.full(n, np.nan)
    own = control is None
    ctl = control or cls.Control(extended_precision=ext)
    try:
        A = cls.CsrMatrix(rp, ci, v, ncols, dtype=dtype)
        if adaptive:
            A.meta_create(ctl)
        xd = cls.DenseVector(torch.tensor(np.asarray(x, dtype=dtype), device="cuda"))
        yd = cls.DenseVector(torch.tensor(np.asarray(y0, dtype=dtype), device="cuda"))
        cls.check("csrmv", A.spmv(cls.Scalar(alpha, tdt), xd, cls.Scalar(beta, tdt), yd, ctl))
        y = yd.t.cpu().numpy()
    finally:
        if own:
            ctl.release()
    yref, scale = oracle.csrmv(rp, ci, np.asarray(v, dtype=dtype), np.asarray(x, dtype=dtype),
                               alpha, beta, np.asarray(y0, dtype=dtype))
    err = per_row_error(y, yref, scale)
    if tol is not None:
        tolv = np.broadcast_to(np.asarray(tol, dtype=np.float64), err.shape)
        over = err - tolv
        bad = int(np.argmax(over)) if err.size else 0
        assert np.all(err <= tolv), (float(err[bad]), bad, float(tolv[bad]))
    return err, y


def row_tolerance(rp, dtype):
    """Per-row parity bar: 1e-6 (FP32) and 1e-13 (FP64) relative to
    |alpha| sum|a_ij x_j| + |beta y_i|, for every row length (the FP32 kernels
    sum each row in FP64)."""
    lens = np.diff(rp.astype(np.int64))
    return np.full(lens.shape, 1e-6 if dtype == np.float32 else 1e-13)


def cg_call(A, x_t, b_t, sc, ctl):
    lib = cls.load()
    fn = lib.clsparseScsrcg if x_t.dtype == torch.float32 else lib.clsparseDcsrcg
    xd, bd = cls.DenseVector(x_t), cls.DenseVector(b_t)
    return fn(ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle, ctl.handle)
