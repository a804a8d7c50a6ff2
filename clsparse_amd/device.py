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


class Mlp:
    """The per-nonzero network of the learned FSAI (clsparseAmdMlp)."""

    def __init__(self, seed: int = 1):
        self.handle = ctypes.c_void_p()
        check("clsparseAmdMlpCreate", load().clsparseAmdMlpCreate(byref(self.handle), seed))

    @staticmethod
    def num_params() -> int:
        return int(load().clsparseAmdMlpNumParams())

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        check("clsparseAmdMlpGetParams", load().clsparseAmdMlpGetParams(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError("wrong parameter count")
        check("clsparseAmdMlpSetParams", load().clsparseAmdMlpSetParams(
            self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def grad(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        check("clsparseAmdMlpGetGrad", load().clsparseAmdMlpGetGrad(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def release(self) -> None:
        if self.handle is not None and self.handle.value:
            load().clsparseAmdMlpDestroy(self.handle)
        self.handle = None


class Fsai:
    """Learned FSAI factor G of an FP64 matrix (clsparseAmdFsai)."""

    def __init__(self, A: CsrMatrix, control: Control):
        self.handle = ctypes.c_void_p()
        check("clsparseAmdFsaiCreate",
              load().clsparseAmdFsaiCreate(byref(self.handle), byref(A.s), control.handle))
        self.n = A.num_rows

    def update_values(self, A: CsrMatrix, control: Control) -> None:
        check("clsparseAmdFsaiUpdateValues",
              load().clsparseAmdFsaiUpdateValues(self.handle, byref(A.s), control.handle))

    def apply_net(self, net: Mlp | None, control: Control) -> None:
        h = net.handle if net is not None else None
        check("clsparseAmdFsaiApplyNet", load().clsparseAmdFsaiApplyNet(self.handle, h,
                                                                         control.handle))

    def nnz(self) -> int:
        k = ctypes.c_uint32(0)
        check("clsparseAmdFsaiGetValues",
              load().clsparseAmdFsaiGetValues(self.handle, None, None, None, byref(k)))
        return int(k.value)

    def values(self):
        """(row_pointer, col_indices, values) of G on the host."""
        k = self.nnz()
        rp = np.zeros(self.n + 1, dtype=np.uint32)
        ci = np.zeros(max(k, 1), dtype=np.uint32)
        va = np.zeros(max(k, 1), dtype=np.float64)
        P = ctypes.POINTER
        check("clsparseAmdFsaiGetValues", load().clsparseAmdFsaiGetValues(
            self.handle, rp.ctypes.data_as(P(ctypes.c_uint32)),
            ci.ctypes.data_as(P(ctypes.c_uint32)), va.ctypes.data_as(P(ctypes.c_double)), None))
        return rp, ci[:k], va[:k]

    def features(self) -> np.ndarray:
        """BF16 bit patterns [nnz][16]."""
        out = np.zeros((max(self.nnz(), 1), 16), dtype=np.uint16)
        check("clsparseAmdFsaiGetFeatures", load().clsparseAmdFsaiGetFeatures(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
        return out[: self.nnz()]

    def output(self) -> np.ndarray:
        out = np.zeros(max(self.nnz(), 1), dtype=np.float32)
        check("clsparseAmdFsaiGetOutput", load().clsparseAmdFsaiGetOutput(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out[: self.nnz()]

    def diagonals(self) -> list:
        """Column - row offsets of the factor's diagonal storage ([] = CSR)."""
        nd = ctypes.c_int32(0)
        offs = (ctypes.c_int32 * 8)()
        check("clsparseAmdFsaiGetDiagonals", load().clsparseAmdFsaiGetDiagonals(
            self.handle, ctypes.byref(nd), offs))
        return [int(offs[d]) for d in range(nd.value)]

    def train_step(self, net: Mlp, A: CsrMatrix, control: Control, probes: int = 8,
                   seed: int = 1, lr: float = 0.0, want_grad: bool = False):
        """Returns (loss, dL/dG per nonzero of G or None)."""
        loss = ctypes.c_double(0.0)
        g = np.zeros(max(self.nnz(), 1), dtype=np.float64) if want_grad else None
        gp = g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if want_grad else None
        check("clsparseAmdFsaiTrainStep", load().clsparseAmdFsaiTrainStep(
            self.handle, net.handle, byref(A.s), probes, seed, lr, control.handle, byref(loss),
            gp))
        return loss.value, (g[: self.nnz()] if want_grad else None)

    def release(self) -> None:
        if self.handle is not None and self.handle.value:
            load().clsparseAmdFsaiDestroy(self.handle)
        self.handle = None


__all__ = ["Control", "Scalar", "DenseVector", "CsrMatrix", "SolverControl", "ClsparseError",
           "NOPRECOND", "DIAGONAL", "Mlp", "Fsai"]
