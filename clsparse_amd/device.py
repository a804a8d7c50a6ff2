"""Device-side convenience objects over the C ABI (torch tensors hold memory).

Each wrapper owns the torch tensors whose device pointers its ctypes struct
carries, so the memory lives as long as the wrapper.
"""
from __future__ import annotations

import ctypes
from ctypes import byref

import numpy as np
import torch

from . import (DIAGONAL, NOPRECOND, QUIET, ClsparseError, check, cldenseVector,
               clsparseAmdSolverStats, clsparseCsrMatrix, clsparseScalar, load)

_TORCH = {np.float32: torch.float32, np.float64: torch.float64}


def _dev(t: torch.Tensor) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    return t.contiguous()


class Control:
    """clsparseControl bound to a torch stream (default: the current stream)."""

    def __init__(self, stream: torch.cuda.Stream | None = None, async_mode: bool = False,
                 extended_precision: bool = False):
        lib = load()
        check("clsparseSetup", lib.clsparseSetup())
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        r = lib.clsparseCreateControl(ctypes.c_void_p(self.stream.cuda_stream))
        check("clsparseCreateControl", r.status)
        self.handle = ctypes.c_void_p(r.control)
        if async_mode:
            check("clsparseEnableAsync", lib.clsparseEnableAsync(self.handle, 1))
        if extended_precision:
            check("clsparseEnableExtendedPrecision",
                  lib.clsparseEnableExtendedPrecision(self.handle, 1))

    def release(self) -> None:
        if self.handle is not None and self.handle.value:
            load().clsparseReleaseControl(self.handle)
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()


class Scalar:
    def __init__(self, value, dtype=torch.float64, device="cuda"):
        if isinstance(value, torch.Tensor):
            self.t = _dev(value.reshape(-1))
        else:
            self.t = torch.tensor([value], dtype=dtype, device=device)
        self.s = clsparseScalar(self.t.data_ptr(), 0)

    def item(self) -> float:
        return self.t[0].item()


class DenseVector:
    def __init__(self, t: torch.Tensor, offset: int = 0, num_values: int | None = None):
        self.t = _dev(t)
        n = self.t.numel() - offset if num_values is None else num_values
        self.s = cldenseVector(n, self.t.data_ptr(), offset)


class CsrMatrix:
    """CSR matrix in device memory built from host arrays."""

    def __init__(self, row_pointer, col_indices, values, num_cols: int, dtype=np.float64,
                 device="cuda"):
        rp = np.ascontiguousarray(row_pointer, dtype=np.uint32)
        ci = np.ascontiguousarray(col_indices, dtype=np.uint32)
        va = np.ascontiguousarray(values, dtype=dtype)
        self.dtype = np.dtype(dtype).type
        self.rp = torch.from_numpy(rp.view(np.int32)).to(device)
        self.ci = torch.from_numpy(ci.view(np.int32) if ci.size else np.zeros(1, np.int32)).to(device)
        self.va = torch.from_numpy(va if va.size else np.zeros(1, dtype)).to(device)
        self.s = clsparseCsrMatrix()
        load().clsparseInitCsrMatrix(byref(self.s))
        self.s.num_rows = rp.size - 1
        self.s.num_cols = num_cols
        self.s.num_nonzeros = ci.size
        self.s.values = self.va.data_ptr()
        self.s.col_indices = self.ci.data_ptr()
        self.s.row_pointer = self.rp.data_ptr()

    @property
    def num_rows(self) -> int:
        return self.s.num_rows

    def meta_create(self, control: Control) -> None:
        check("clsparseCsrMetaCreate", load().clsparseCsrMetaCreate(byref(self.s), control.handle))

    def meta_delete(self) -> None:
        load().clsparseCsrMetaDelete(byref(self.s))

    def meta_size(self, control: Control) -> int:
        r = load().clsparseCsrMetaSize(byref(self.s), control.handle)
        check("clsparseCsrMetaSize", r.status)
        return int(r.metaSize)

    def row_blocks(self) -> np.ndarray:
        lib = load()
        n = ctypes.c_uint32(0)
        check("clsparseAmdCsrMetaGetRowBlocks",
              lib.clsparseAmdCsrMetaGetRowBlocks(byref(self.s), None, 0, byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        check("clsparseAmdCsrMetaGetRowBlocks",
              lib.clsparseAmdCsrMetaGetRowBlocks(
                  byref(self.s), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n.value,
                  byref(n)))
        return out[: n.value]

    def spmv(self, alpha: Scalar, x: DenseVector, beta: Scalar, y: DenseVector,
             control: Control) -> int:
        fn = load().clsparseScsrmv if self.dtype is np.float32 else load().clsparseDcsrmv
        return fn(byref(alpha.s), byref(self.s), byref(x.s), byref(beta.s), byref(y.s),
                  control.handle)

    def __del__(self):
        try:
            if self.s.meta:
                self.meta_delete()
        except Exception:
            pass


class SolverControl:
    def __init__(self, precond=DIAGONAL, max_iters=1000, rel_tol=1e-2, abs_tol=1e-5,
                 print_mode=QUIET):
        lib = load()
        r = lib.clsparseCreateSolverControl(precond, max_iters, rel_tol, abs_tol)
        check("clsparseCreateSolverControl", r.status)
        self.handle = ctypes.c_void_p(r.control)
        if print_mode != QUIET:
            lib.clsparseSolverPrintMode(self.handle, print_mode)

    def stats(self) -> clsparseAmdSolverStats:
        st = clsparseAmdSolverStats()
        check("clsparseAmdGetSolverStats", load().clsparseAmdGetSolverStats(self.handle, byref(st)))
        return st

    def set_params(self, precond, max_iters, rel_tol, abs_tol) -> None:
        check("clsparseSetSolverParams",
              load().clsparseSetSolverParams(self.handle, precond, max_iters, rel_tol, abs_tol))

    def release(self) -> None:
        if self.handle is not None and self.handle.value:
            load().clsparseReleaseSolverControl(self.handle)
        self.handle = None


__all__ = ["Control", "Scalar", "DenseVector", "CsrMatrix", "SolverControl", "ClsparseError",
           "NOPRECOND", "DIAGONAL"]
