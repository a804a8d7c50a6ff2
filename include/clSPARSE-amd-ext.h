/*
 * clSPARSE-amd-ext.h -- extension entry points of the MI355X library.
 *
 * None of these has a counterpart in the reference's public header; they are
 * additive and never change the behaviour of a reference entry point.
 *   - metadata read-back (lets callers check the row-block array bit for bit),
 *   - solver statistics (executed iterations, FP64 true residual, timings),
 *   - the mixed-precision (FP32 inner / FP64 outer) Jacobi-PCG solver,
 *   - seeded synthetic matrix and vector generators used by tests and benches.
 */
#ifndef CLSPARSE_AMD_EXT_H_
#define CLSPARSE_AMD_EXT_H_

#include "clSPARSE.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copies the row-block entries written by clsparseCsrMetaCreate (before any
 * SpMV; this build never modifies them) into host memory. *entries receives
 * the number of 64-bit entries (= rowBlockSize / 2); at most `capacity` are
 * copied. clsparseStructInvalid if the matrix has no metadata. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdCsrMetaGetRowBlocks(const clsparseCsrMatrix* csrMatx, cl_ulong* host_out,
                                   clsparseIdx_t capacity, clsparseIdx_t* entries);

/* The same binning on a host row pointer, without a device: writes the
 * entries to out (NULL = count only) and returns the reference's rowBlockSize
 * (2 * entries written, or 2 * entries counted by the size-only pass). */
CLSPARSE_EXPORT cl_ulong
    clsparseAmdComputeRowBlocksHost(const cl_uint* row_pointer, cl_uint num_rows, cl_ulong* out);

/* Per-solve statistics kept in the solver control object. */
typedef struct clsparseAmdSolverStats_
{
    cl_int iterations;          /* CG iterations executed (inner, all refinements) */
    cl_int outer_iterations;    /* refinement steps (mixed solver), 0 otherwise */
    cl_double l1_residual;      /* final ||r||_1 / ||b||_1 of the recurrence */
    cl_double true_residual;    /* final FP64 ||b - A x||_2 / ||b||_2 (mixed solver) */
    cl_double seconds;          /* wall time of the solve, host clock */
    cl_int status_detail;       /* 0 converged, 1 maxIters reached, 2 stagnated */
    cl_int reserved;
    /* Snapshot of the solver control's reference fields
     * (src/library/solvers/solver-control.hpp L108-127). */
    cl_int n_iters;
    cl_int max_iters;
    cl_double relative_tolerance;
    cl_double absolute_tolerance;
    cl_double initial_residual;
    cl_double current_residual;
} clsparseAmdSolverStats;

CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGetSolverStats(clSParseSolverControl solverControl, clsparseAmdSolverStats* out);

/* Mixed-precision iterative refinement with Jacobi-PCG (no reference
 * counterpart). A, x and b are FP64. The inner solves run in FP32 values with
 * FP64 dot/norm reductions; the outer loop forms r = b - A x with the FP64
 * CSR-Adaptive SpMV and updates x in FP64. Stops when the FP64 true relative
 * residual ||b - A x||_2 / ||b||_2 <= relTol of the solver control, or when
 * maxIters inner iterations in total have run. The preconditioner field of the
 * solver control selects DIAGONAL or NOPRECOND for the inner solves. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdDcsrcgMixed(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                           clSParseSolverControl solverControl, clsparseControl control);

/* The same refinement loop with FP64 inner solves: FP64 Jacobi-PCG judged by
 * the FP64 true relative residual (the baseline the mixed solver is set
 * beside). */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdDcsrcgRefined(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                             clSParseSolverControl solverControl, clsparseControl control);

/* ---- synthetic generators (host memory, seeded, deterministic) ----------- */
typedef struct clsparseAmdHostCsr_
{
    clsparseIdx_t num_rows;
    clsparseIdx_t num_cols;
    clsparseIdx_t num_nonzeros;
    cl_uint* row_pointer;      /* num_rows + 1 */
    cl_uint* col_indices;      /* num_nonzeros, sorted and unique within each row */
    cl_double* values;         /* num_nonzeros */
} clsparseAmdHostCsr;

CLSPARSE_EXPORT clsparseStatus clsparseAmdHostCsrFree(clsparseAmdHostCsr* m);

/* 2-D 5-point Poisson Laplacian on an nx x ny grid: diagonal 4, neighbours -1. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenPoisson2D(cl_uint nx, cl_uint ny, clsparseAmdHostCsr* out);

/* Power-law rows: length = clip(round(LogNormal(ln(mean_len) - sigma^2/2, sigma)),
 * 1, max_len); then n_dense rows at seeded positions get dense_len entries.
 * Columns uniform without replacement over [0, n), sorted; values U[-1, 1]. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenPowerLaw(cl_uint n, cl_double mean_len, cl_double sigma, cl_uint max_len,
                           cl_uint n_dense, cl_uint dense_len, cl_ulong seed,
                           clsparseAmdHostCsr* out);

/* 3-D 7-point variable-coefficient diffusion on nx*ny*nz cells with Dirichlet
 * boundaries: cell coefficient kappa ~ LogNormal(0, sigma), face coefficient
 * the harmonic mean of the two cells (the boundary face uses the cell's own
 * kappa), scaled per axis by (ax, ay, az). SPD, diagonally dominant. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenDiffusion3D(cl_uint nx, cl_uint ny, cl_uint nz, cl_double sigma,
                              cl_double ax, cl_double ay, cl_double az, cl_ulong seed,
                              clsparseAmdHostCsr* out);

/* Fills out[0..n) with U[lo, hi) from a SplitMix64 stream of `seed`. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenUniform(cl_double* out, size_t n, cl_double lo, cl_double hi, cl_ulong seed);

#ifdef __cplusplus
}
#endif

#endif /* CLSPARSE_AMD_EXT_H_ */
