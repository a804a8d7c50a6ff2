// Library setup, struct initialisers, control and solver-control objects.
//
// Reference: src/library/clsparse-init.cpp L25-107,
// src/library/internal/clsparse-control.cpp L98-235,
// src/library/solvers/solver-control.cpp L20-102.
#include <cstring>
#include <iostream>
#include <new>

#include "common.hpp"

int clsparseInitialized = 0;

// Version of the reference this library stands in for (CMakeLists.txt L40).
extern "C" clsparseStatus clsparseGetVersion(cl_uint* major, cl_uint* minor, cl_uint* patch,
                                             cl_uint* tweak)
{
    if (major) *major = 0;
    if (minor) *minor = 10;
    if (patch) *patch = 2;
    if (tweak) *tweak = 0;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseSetup(void)
{
    clsparseInitialized = 1;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseTeardown(void)
{
    clsparseInitialized = 0;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseInitScalar(clsparseScalar* scalar)
{
    if (scalar) std::memset(scalar, 0, sizeof(*scalar));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseInitVector(cldenseVector* vec)
{
    if (vec) std::memset(vec, 0, sizeof(*vec));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseInitCooMatrix(clsparseCooMatrix* cooMatx)
{
    if (cooMatx) std::memset(cooMatx, 0, sizeof(*cooMatx));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseInitCsrMatrix(clsparseCsrMatrix* csrMatx)
{
    if (csrMatx) std::memset(csrMatx, 0, sizeof(*csrMatx));
    return clsparseSuccess;
}

extern "C" clsparseStatus cldenseInitMatrix(cldenseMatrix* denseMatx)
{
    if (denseMatx) {
        std::memset(denseMatx, 0, sizeof(*denseMatx));
        denseMatx->major = rowMajor;
    }
    return clsparseSuccess;
}

// ---- control ---------------------------------------------------------------

extern "C" clsparseCreateResult clsparseCreateControl(hipStream_t stream)
{
    clsparseCreateResult r{clsparseSuccess, nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        r.status = clsparseInvalidDevice;
        return r;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        r.status = clsparseInvalidDevice;
        return r;
    }
    // Kernels are compiled for gfx950 only (wave64, CDNA4).
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0 || prop.warpSize != 64) {
        std::fprintf(stderr, "clSPARSE: device %d is %s; the library is compiled for gfx950 only\n", dev,
                     prop.gcnArchName);
        r.status = clsparseInvalidDevice;
        return r;
    }
    _clsparseControl* c = new (std::nothrow) _clsparseControl;
    if (c == nullptr) {
        r.status = clsparseOutOfHostMemory;
        return r;
    }
    c->stream = stream;
    c->device = dev;
    c->num_cus = prop.multiProcessorCount;
    if (hipEventCreateWithFlags(&c->event, hipEventDisableTiming) != hipSuccess) {
        delete c;
        r.status = clsparseOutOfResources;
        return r;
    }
    if (hipMalloc(&c->red_partials, kReducePartialsMax * sizeof(double)) != hipSuccess) {
        (void)hipEventDestroy(c->event);
        delete c;
        r.status = clsparseOutOfResources;
        return r;
    }
    // Default priority (the same as the caller's stream).
    if (hipStreamCreateWithFlags(&c->solver_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(c->red_partials);
        (void)hipEventDestroy(c->event);
        delete c;
        r.status = clsparseOutOfResources;
        return r;
    }
    if (const clsparseStatus st = sum_ax_scratch_setup(dev); st != clsparseSuccess) {
        (void)hipStreamDestroy(c->solver_stream);
        (void)hipFree(c->red_partials);
        (void)hipEventDestroy(c->event);
        delete c;
        r.status = st;
        return r;
    }
    r.control = c;
    return r;
}

extern "C" clsparseStatus clsparseEnableAsync(clsparseControl control, cl_bool async)
{
    if (control == nullptr) return clsparseInvalidControlObject;
    control->async = async != 0;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseEnableExtendedPrecision(clsparseControl control,
                                                          cl_bool extPrecision)
{
    if (control == nullptr) return clsparseInvalidControlObject;
    control->extended_precision = extPrecision != 0;
    return clsparseSuccess;
}

extern "C" clsparseEventResult clsparseGetEvent(clsparseControl control)
{
    clsparseEventResult r{clsparseSuccess, nullptr};
    if (control == nullptr) {
        r.status = clsparseInvalidControlObject;
        return r;
    }
    r.event = control->event;
    return r;
}

extern "C" clsparseStatus clsparseReleaseControl(clsparseControl control)
{
    if (control == nullptr) return clsparseInvalidControlObject;
    (void)hipStreamSynchronize(control->stream);
    if (control->work) (void)hipFree(control->work);
    if (control->red_partials) (void)hipFree(control->red_partials);
    if (control->pinned) (void)hipHostFree(control->pinned);
    if (control->event) (void)hipEventDestroy(control->event);
    if (control->solver_stream) (void)hipStreamDestroy(control->solver_stream);
    delete control;
    return clsparseSuccess;
}

clsparseStatus finish_op(clsparseControl control)
{
    CLS_HIP_TRY(hipEventRecord(control->event, control->stream));
    if (!control->async) CLS_HIP_TRY(hipEventSynchronize(control->event));
    return clsparseSuccess;
}

clsparseStatus control_workspace(clsparseControl control, size_t bytes, void** out)
{
    if (bytes > control->work_bytes) {
        if (control->work) {
            CLS_HIP_TRY(hipStreamSynchronize(control->stream));
            CLS_HIP_TRY(hipFree(control->work));
            control->work = nullptr;
            control->work_bytes = 0;
        }
        size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
        if (hipMalloc(&control->work, want) != hipSuccess) {
            control->work = nullptr;
            return clsparseOutOfResources;
        }
        control->work_bytes = want;
    }
    *out = control->work;
    return clsparseSuccess;
}

clsparseStatus control_pinned(clsparseControl control, size_t bytes, void** out)
{
    if (bytes > control->pinned_bytes) {
        if (control->pinned) {
            CLS_HIP_TRY(hipStreamSynchronize(control->stream));
            CLS_HIP_TRY(hipHostFree(control->pinned));
            control->pinned = nullptr;
            control->pinned_bytes = 0;
        }
        size_t want = bytes < 4096 ? 4096 : bytes;
        if (hipHostMalloc(&control->pinned, want, hipHostMallocDefault) != hipSuccess) {
            control->pinned = nullptr;
            return clsparseOutOfHostMemory;
        }
        control->pinned_bytes = want;
    }
    *out = control->pinned;
    return clsparseSuccess;
}

// ---- solver control --------------------------------------------------------

void _solverControl::print() const
{
    if (printMode == VERBOSE)
        std::cout << "Iteration: " << nIters << " Residuum: " << currentResidual << std::endl;
}

void _solverControl::printSummary(clsparseStatus status) const
{
    if (printMode == VERBOSE || printMode == NORMAL) {
        std::cout << "Solver constraints: \n"
                  << "\trelative tolerance = " << relativeTolerance
                  << "\n\tabsolute tolerance = " << absoluteTolerance
                  << "\n\tmax iterations = " << maxIters << "\n\tPreconditioner: "
                  << (preconditioner == DIAGONAL ? "Diagonal" : "No preconditioner")
                  << std::endl;
        std::cout << "Solver finished calculations with status " << status << std::endl;
        std::cout << "\tfinal residual = " << currentResidual << "\titerations = " << nIters
                  << std::endl;
    }
}

extern "C" clsparseCreateSolverResult clsparseCreateSolverControl(PRECONDITIONER precond,
                                                                  cl_int maxIters,
                                                                  cl_double relTol,
                                                                  cl_double absTol)
{
    clsparseCreateSolverResult r{clsparseSuccess, nullptr};
    _solverControl* s = new (std::nothrow) _solverControl;
    if (s == nullptr) {
        r.status = clsparseOutOfHostMemory;
        return r;
    }
    s->absoluteTolerance = absTol;
    s->relativeTolerance = relTol;
    s->nIters = 0;
    s->maxIters = maxIters;
    s->initialResidual = 0;
    s->preconditioner = precond;
    r.control = s;
    return r;
}

extern "C" clsparseStatus clsparseReleaseSolverControl(clSParseSolverControl solverControl)
{
    if (solverControl == nullptr) return clsparseInvalidSolverControlObject;
    delete solverControl;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseSetSolverParams(clSParseSolverControl solverControl,
                                                  PRECONDITIONER precond, cl_int maxIters,
                                                  cl_double relTol, cl_double absTol)
{
    if (solverControl == nullptr) return clsparseInvalidSolverControlObject;
    solverControl->absoluteTolerance = absTol;
    solverControl->relativeTolerance = relTol;
    solverControl->nIters = 0;
    solverControl->maxIters = maxIters;
    solverControl->initialResidual = 0;
    solverControl->preconditioner = precond;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseSolverPrintMode(clSParseSolverControl solverControl,
                                                  PRINT_MODE mode)
{
    if (solverControl == nullptr) return clsparseInvalidSolverControlObject;
    solverControl->printMode = mode;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdGetSolverStats(clSParseSolverControl solverControl,
                                                    clsparseAmdSolverStats* out)
{
    if (solverControl == nullptr) return clsparseInvalidSolverControlObject;
    if (out) {
        *out = solverControl->stats;
        out->n_iters = solverControl->nIters;
        out->max_iters = solverControl->maxIters;
        out->relative_tolerance = solverControl->relativeTolerance;
        out->absolute_tolerance = solverControl->absoluteTolerance;
        out->initial_residual = solverControl->initialResidual;
        out->current_residual = solverControl->currentResidual;
    }
    return clsparseSuccess;
}
