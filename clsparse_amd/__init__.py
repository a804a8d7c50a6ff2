"""Python host mirror of the clSPARSE C API for the MI355X build.

Loads ``clsparse_amd/lib/libclSPARSE.so`` (built for gfx950 by
``__graft_entry__.build()``) and exposes the reference's entry points under
their own names (``clsparseScsrmv``, ``clsparseCsrMetaCreate``,
``clsparseDcsrcg`` ...), with the reference's struct layouts
(src/include/clSPARSE-1x.h L31-143) as ctypes structures whose ``cl_mem``
fields hold HIP device pointers. Device memory comes from torch tensors.

There is no CPU fallback: if the shared library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_double, c_float, c_int32, c_size_t, c_uint32, c_uint64,
                    c_void_p)

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libclSPARSE.so")

# ---- status codes (src/include/clSPARSE.h L45-88) ---------------------------
clsparseSuccess = 0
clsparseInvalidValue = -30
clsparseInvalidDevice = -33
clsparseInvalidMemObject = -38
clsparseOutOfResources = -5
clsparseNotInitialized = -1023
clsparseStructInvalid = -1022
clsparseInvalidSize = -1021
clsparseInvalidControlObject = -1018
clsparseInvalidKernelExecution = -1015
clsparseInvalidMemObj = -1020
clsparseInvalidFile = -1017
clsparseInvalidFileFormat = -1016
clsparseInvalidSolverControlObject = -2048
clsparseInvalidSystemSize = -2047

QUIET, NORMAL, VERBOSE = 0, 1, 2
NOPRECOND, DIAGONAL = 0, 1


class clsparseScalar(Structure):
    _fields_ = [("value", c_void_p), ("off_value", c_uint32)]


class cldenseVector(Structure):
    _fields_ = [("num_values", c_uint32), ("values", c_void_p), ("off_values", c_uint32)]


class clsparseCsrMatrix(Structure):
    _fields_ = [("num_rows", c_uint32), ("num_cols", c_uint32), ("num_nonzeros", c_uint32),
                ("values", c_void_p), ("col_indices", c_void_p), ("row_pointer", c_void_p),
                ("off_values", c_uint32), ("off_col_indices", c_uint32),
                ("off_row_pointer", c_uint32), ("meta", c_void_p)]


class clsparseCooMatrix(Structure):
    _fields_ = [("num_rows", c_uint32), ("num_cols", c_uint32), ("num_nonzeros", c_uint32),
                ("values", c_void_p), ("col_indices", c_void_p), ("row_indices", c_void_p),
                ("off_values", c_uint32), ("off_col_indices", c_uint32),
                ("off_row_indices", c_uint32)]


class cldenseMatrix(Structure):
    _fields_ = [("num_rows", c_uint32), ("num_cols", c_uint32), ("lead_dim", c_uint32),
                ("major", c_int32), ("values", c_void_p), ("off_values", c_uint32)]


class clsparseCreateResult(Structure):
    _fields_ = [("status", c_int32), ("control", c_void_p)]


class clsparseEventResult(Structure):
    _fields_ = [("status", c_int32), ("event", c_void_p)]


class clsparseCreateSolverResult(Structure):
    _fields_ = [("status", c_int32), ("control", c_void_p)]


class clsparseMetaSizeResult(Structure):
    _fields_ = [("status", c_int32), ("metaSize", c_uint32)]


class clsparseAmdSolverStats(Structure):
    _fields_ = [("iterations", c_int32), ("outer_iterations", c_int32),
                ("l1_residual", c_double), ("true_residual", c_double),
                ("seconds", c_double), ("status_detail", c_int32), ("preconditioner", c_int32),
                ("n_iters", c_int32), ("max_iters", c_int32),
                ("relative_tolerance", c_double), ("absolute_tolerance", c_double),
                ("initial_residual", c_double), ("current_residual", c_double),
                ("pilot_rate_fsai", c_double), ("pilot_rate_jacobi", c_double),
                ("a_diagonals", c_int32), ("g_diagonals", c_int32),
                ("loss_fsai", c_double), ("loss_jacobi", c_double)]


class clsparseAmdHostCsr(Structure):
    _fields_ = [("num_rows", c_uint32), ("num_cols", c_uint32), ("num_nonzeros", c_uint32),
                ("row_pointer", POINTER(c_uint32)), ("col_indices", POINTER(c_uint32)),
                ("values", POINTER(c_double))]


# Every exported entry point of include/clSPARSE.h and include/clSPARSE-amd-ext.h.
_V = c_void_p
_SIGS = {
    "clsparseSetup": (c_int32, []),
    "clsparseTeardown": (c_int32, []),
    "clsparseGetVersion": (c_int32, [POINTER(c_uint32)] * 4),
    "clsparseInitScalar": (c_int32, [POINTER(clsparseScalar)]),
    "clsparseInitVector": (c_int32, [POINTER(cldenseVector)]),
    "clsparseInitCooMatrix": (c_int32, [POINTER(clsparseCooMatrix)]),
    "clsparseInitCsrMatrix": (c_int32, [POINTER(clsparseCsrMatrix)]),
    "cldenseInitMatrix": (c_int32, [POINTER(cldenseMatrix)]),
    "clsparseCreateControl": (clsparseCreateResult, [_V]),
    "clsparseEnableAsync": (c_int32, [_V, c_uint32]),
    "clsparseEnableExtendedPrecision": (c_int32, [_V, c_uint32]),
    "clsparseGetEvent": (clsparseEventResult, [_V]),
    "clsparseReleaseControl": (c_int32, [_V]),
    "clsparseCreateSolverControl": (clsparseCreateSolverResult, [c_int32, c_int32, c_double, c_double]),
    "clsparseReleaseSolverControl": (c_int32, [_V]),
    "clsparseSetSolverParams": (c_int32, [_V, c_int32, c_int32, c_double, c_double]),
    "clsparseSolverPrintMode": (c_int32, [_V, c_int32]),
    "clsparseScsrcg": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                 POINTER(cldenseVector), _V, _V]),
    "clsparseDcsrcg": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                 POINTER(cldenseVector), _V, _V]),
    "clsparseScsrbicgStab": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                       POINTER(cldenseVector), _V, _V]),
    "clsparseDcsrbicgStab": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                       POINTER(cldenseVector), _V, _V]),
    "clsparseCsrMetaSize": (clsparseMetaSizeResult, [POINTER(clsparseCsrMatrix), _V]),
    "clsparseCsrMetaCreate": (c_int32, [POINTER(clsparseCsrMatrix), _V]),
    "clsparseCsrMetaDelete": (c_int32, [POINTER(clsparseCsrMatrix)]),
    "clsparseCsrMetaCompute": (c_int32, [POINTER(clsparseCsrMatrix), _V]),
    "clsparseScsrmv": (c_int32, [POINTER(clsparseScalar), POINTER(clsparseCsrMatrix),
                                 POINTER(cldenseVector), POINTER(clsparseScalar),
                                 POINTER(cldenseVector), _V]),
    "clsparseDcsrmv": (c_int32, [POINTER(clsparseScalar), POINTER(clsparseCsrMatrix),
                                 POINTER(cldenseVector), POINTER(clsparseScalar),
                                 POINTER(cldenseVector), _V]),
    "clsparseHeaderfromFile": (c_int32, [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                         ctypes.c_char_p]),
    "clsparseSCooMatrixfromFile": (c_int32, [POINTER(clsparseCooMatrix), ctypes.c_char_p, _V, c_uint32]),
    "clsparseDCooMatrixfromFile": (c_int32, [POINTER(clsparseCooMatrix), ctypes.c_char_p, _V, c_uint32]),
    "clsparseSCsrMatrixfromFile": (c_int32, [POINTER(clsparseCsrMatrix), ctypes.c_char_p, _V, c_uint32]),
    "clsparseDCsrMatrixfromFile": (c_int32, [POINTER(clsparseCsrMatrix), ctypes.c_char_p, _V, c_uint32]),
    "clsparseAmdCsrMatrixfromFileHost": (c_int32, [ctypes.c_char_p, c_uint32,
                                                   POINTER(clsparseAmdHostCsr)]),
    "clsparseScsrmm": (c_int32, [POINTER(clsparseScalar), POINTER(clsparseCsrMatrix),
                                 POINTER(cldenseMatrix), POINTER(clsparseScalar),
                                 POINTER(cldenseMatrix), _V]),
    "clsparseDcsrmm": (c_int32, [POINTER(clsparseScalar), POINTER(clsparseCsrMatrix),
                                 POINTER(cldenseMatrix), POINTER(clsparseScalar),
                                 POINTER(cldenseMatrix), _V]),
    "clsparseAmdCsrMetaGetRowBlocks": (c_int32, [POINTER(clsparseCsrMatrix), POINTER(c_uint64),
                                                 c_uint32, POINTER(c_uint32)]),
    "clsparseAmdComputeRowBlocksHost": (c_uint64, [POINTER(c_uint32), c_uint32, POINTER(c_uint64)]),
    "clsparseAmdGetSolverStats": (c_int32, [_V, POINTER(clsparseAmdSolverStats)]),
    "clsparseAmdDcsrcgMixed": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                         POINTER(cldenseVector), _V, _V]),
    "clsparseAmdDcsrcgRefined": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                           POINTER(cldenseVector), _V, _V]),
    "clsparseAmdHostCsrFree": (c_int32, [POINTER(clsparseAmdHostCsr)]),
    "clsparseAmdGenPoisson2D": (c_int32, [c_uint32, c_uint32, POINTER(clsparseAmdHostCsr)]),
    "clsparseAmdGenPowerLaw": (c_int32, [c_uint32, c_double, c_double, c_uint32, c_uint32,
                                         c_uint32, c_uint64, POINTER(clsparseAmdHostCsr)]),
    "clsparseAmdGenDiffusion3D": (c_int32, [c_uint32, c_uint32, c_uint32, c_double, c_double,
                                            c_double, c_double, c_uint64,
                                            POINTER(clsparseAmdHostCsr)]),
    "clsparseAmdGenAnisoDiffusion2D": (c_int32, [c_uint32, c_uint32, c_double, c_double, c_double,
                                                 c_double, c_uint64, POINTER(clsparseAmdHostCsr)]),
    "clsparseAmdGenUniform": (c_int32, [POINTER(c_double), c_size_t, c_double, c_double, c_uint64]),
    # learned FSAI
    "clsparseAmdMlpCreate": (c_int32, [POINTER(c_void_p), c_uint64]),
    "clsparseAmdMlpDestroy": (c_int32, [_V]),
    "clsparseAmdMlpNumParams": (c_uint32, []),
    "clsparseAmdMlpGetParams": (c_int32, [_V, POINTER(c_float)]),
    "clsparseAmdMlpSetParams": (c_int32, [_V, POINTER(c_float)]),
    "clsparseAmdMlpGetGrad": (c_int32, [_V, POINTER(c_float)]),
    "clsparseAmdFsaiCreate": (c_int32, [POINTER(c_void_p), POINTER(clsparseCsrMatrix), _V]),
    "clsparseAmdFsaiUpdateValues": (c_int32, [_V, POINTER(clsparseCsrMatrix), _V]),
    "clsparseAmdFsaiDestroy": (c_int32, [_V]),
    "clsparseAmdFsaiApplyNet": (c_int32, [_V, _V, _V]),
    "clsparseAmdFsaiGetValues": (c_int32, [_V, POINTER(c_uint32), POINTER(c_uint32),
                                           POINTER(c_double), POINTER(c_uint32)]),
    "clsparseAmdFsaiGetFeatures": (c_int32, [_V, POINTER(ctypes.c_uint16)]),
    "clsparseAmdFsaiGetOutput": (c_int32, [_V, POINTER(c_float)]),
    "clsparseAmdFsaiGetDiagonals": (c_int32, [_V, POINTER(c_int32), POINTER(c_int32)]),
    "clsparseAmdFsaiTrainStep": (c_int32, [_V, _V, POINTER(clsparseCsrMatrix), c_uint32, c_uint64,
                                           c_float, _V, POINTER(c_double), POINTER(c_double)]),
    "clsparseAmdDcsrcgFsai": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                        POINTER(cldenseVector), _V, _V, _V]),
    "clsparseAmdDcsrcgFsai64": (c_int32, [POINTER(cldenseVector), POINTER(clsparseCsrMatrix),
                                          POINTER(cldenseVector), _V, _V, _V]),
}
for _s in ("S", "D"):
    _SIGS[f"cldense{_s}scale"] = (c_int32, [POINTER(cldenseVector), POINTER(clsparseScalar),
                                            POINTER(cldenseVector), _V])
    _SIGS[f"cldense{_s}axpy"] = (c_int32, [POINTER(cldenseVector), POINTER(clsparseScalar),
                                           POINTER(cldenseVector), POINTER(cldenseVector), _V])
    _SIGS[f"cldense{_s}axpby"] = (c_int32, [POINTER(cldenseVector), POINTER(clsparseScalar),
                                            POINTER(cldenseVector), POINTER(clsparseScalar),
                                            POINTER(cldenseVector), _V])
    for _op in ("nrm1", "nrm2", "reduce"):
        _SIGS[f"cldense{_s}{_op}"] = (c_int32, [POINTER(clsparseScalar), POINTER(cldenseVector), _V])
    _SIGS[f"cldense{_s}dot"] = (c_int32, [POINTER(clsparseScalar), POINTER(cldenseVector),
                                          POINTER(cldenseVector), _V])
    for _op in ("add", "sub", "mul", "div"):
        _SIGS[f"cldense{_s}{_op}"] = (c_int32, [POINTER(cldenseVector), POINTER(cldenseVector),
                                                POINTER(cldenseVector), _V])
    _SIGS[f"clsparseAmd{_s}csrSumAx"] = (c_int32, [POINTER(clsparseCsrMatrix), POINTER(cldenseVector),
                                                   POINTER(clsparseScalar), _V])
    _SIGS[f"clsparseAmd{_s}gatherSum"] = (c_int32, [POINTER(cldenseVector), c_uint32, c_uint64, c_uint32,
                                                    POINTER(clsparseScalar), _V])
    _SIGS[f"clsparseAmd{_s}csrmvDiag"] = (c_int32, [POINTER(clsparseCsrMatrix), POINTER(cldenseVector),
                                                    POINTER(cldenseVector), c_int32, POINTER(c_int32), _V])
_SIGS["cldenseIreduce"] = (c_int32, [POINTER(clsparseScalar), POINTER(cldenseVector), _V])
_SIGS["clsparseAmdDcsrSumAxNoAlloc"] = (c_int32, [POINTER(clsparseCsrMatrix), POINTER(cldenseVector),
                                                 POINTER(clsparseScalar), _V])

EXPORTED = tuple(sorted(_SIGS))

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library (once) and declares every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class ClsparseError(RuntimeError):
    def __init__(self, name: str, status: int):
        super().__init__(f"{name} returned status {status}")
        self.status = status


def check(name: str, status: int) -> None:
    if status != clsparseSuccess:
        raise ClsparseError(name, status)


# ---- host generators ----------------------------------------------------------

def _take_host_csr(m: clsparseAmdHostCsr):
    n, nnz = m.num_rows, m.num_nonzeros
    rp = np.ctypeslib.as_array(m.row_pointer, shape=(n + 1,)).copy()
    ci = np.ctypeslib.as_array(m.col_indices, shape=(max(nnz, 1),))[:nnz].copy()
    va = np.ctypeslib.as_array(m.values, shape=(max(nnz, 1),))[:nnz].copy()
    ncols = m.num_cols
    load().clsparseAmdHostCsrFree(ctypes.byref(m))
    return rp, ci, va, ncols


def gen_poisson2d(nx: int, ny: int):
    """(row_pointer u32, col_indices u32, values f64, num_cols) of the 5-point Laplacian."""
    m = clsparseAmdHostCsr()
    check("clsparseAmdGenPoisson2D", load().clsparseAmdGenPoisson2D(nx, ny, ctypes.byref(m)))
    return _take_host_csr(m)


def gen_powerlaw(n: int, mean_len: float = 16.0, sigma: float = 1.5, max_len: int = 200000,
                 n_dense: int = 8, dense_len: int = 100000, seed: int = 1):
    m = clsparseAmdHostCsr()
    check("clsparseAmdGenPowerLaw",
          load().clsparseAmdGenPowerLaw(n, mean_len, sigma, max_len, n_dense, dense_len, seed,
                                        ctypes.byref(m)))
    return _take_host_csr(m)


def gen_diffusion3d(nx: int, ny: int, nz: int, sigma: float = 1.0, ax: float = 1.0,
                    ay: float = 1.0, az: float = 1.0, seed: int = 1):
    m = clsparseAmdHostCsr()
    check("clsparseAmdGenDiffusion3D",
          load().clsparseAmdGenDiffusion3D(nx, ny, nz, sigma, ax, ay, az, seed, ctypes.byref(m)))
    return _take_host_csr(m)


def gen_aniso2d(nx: int, ny: int, corr_len: float = 16.0, sigma: float = 1.0,
                ratio: float = 10.0, theta: float = 0.0, seed: int = 1):
    """2-D anisotropic Q1 diffusion (9-point, SPD) on nx x ny interior nodes."""
    m = clsparseAmdHostCsr()
    check("clsparseAmdGenAnisoDiffusion2D",
          load().clsparseAmdGenAnisoDiffusion2D(nx, ny, corr_len, sigma, ratio, theta, seed,
                                                ctypes.byref(m)))
    return _take_host_csr(m)


def read_mtx_host(path: str, read_explicit_zeroes: bool = False):
    """(row_pointer, col_indices, values f64, num_cols) of a Matrix Market file."""
    m = clsparseAmdHostCsr()
    check("clsparseAmdCsrMatrixfromFileHost",
          load().clsparseAmdCsrMatrixfromFileHost(path.encode(), 1 if read_explicit_zeroes else 0,
                                                  ctypes.byref(m)))
    return _take_host_csr(m)


def mtx_header(path: str):
    """(nnz upper bound, rows, cols) from the file header."""
    nnz, r, c = c_uint32(), c_uint32(), c_uint32()
    check("clsparseHeaderfromFile", load().clsparseHeaderfromFile(
        ctypes.byref(nnz), ctypes.byref(r), ctypes.byref(c), path.encode()))
    return nnz.value, r.value, c.value


def gen_uniform(n: int, lo: float = -1.0, hi: float = 1.0, seed: int = 1) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    check("clsparseAmdGenUniform",
          load().clsparseAmdGenUniform(out.ctypes.data_as(POINTER(c_double)), n, lo, hi, seed))
    return out


def row_blocks_host(row_pointer: np.ndarray, count_only: bool = False):
    """Row-block array of the library's binning for a host row pointer.

    Returns (entries u64 array, rowBlockSize) -- or, with count_only, the
    size-only pass result (the value clsparseCsrMetaSize reports before
    clsparseCsrMetaCreate)."""
    rp = np.ascontiguousarray(row_pointer, dtype=np.uint32)
    n = rp.size - 1
    lib = load()
    size = lib.clsparseAmdComputeRowBlocksHost(rp.ctypes.data_as(POINTER(c_uint32)), n, None)
    if count_only:
        return int(size)
    out = np.zeros(size // 2 + 2, dtype=np.uint64)
    created = lib.clsparseAmdComputeRowBlocksHost(rp.ctypes.data_as(POINTER(c_uint32)), n,
                                                  out.ctypes.data_as(POINTER(c_uint64)))
    return out[: created // 2].copy(), int(created)


from .device import Control, CsrMatrix, DenseVector, Fsai, Mlp, Scalar, SolverControl  # noqa: E402,F401
