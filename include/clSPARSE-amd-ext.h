/*
 * clSPARSE-amd-ext.h -- extension entry points of the MI355X library.
 *
 * None of these has a counterpart in the reference's public header; they are
 * additive and never change the behaviour of a reference entry point.
 *   - metadata read-back (lets callers check the row-block array bit for bit),
 *   - solver statistics (executed iterations, FP64 true residual, timings),
 *   - the mixed-precision (FP32 inner / FP64 outer) Jacobi-PCG solver,
 *   - the learned factorized sparse approximate inverse preconditioner (FSAI),
 *   - seeded synthetic matrix and vector generators used by tests and benches.
 */
#ifndef CLSPARSE_AMD_EXT_H_
#define CLSPARSE_AMD_EXT_H_

#include "clSPARSE.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Copies the row-block entries written by clsparseCsrMetaCreate (before any
 * SpMV; this library never modifies them) into host memory. *entries receives
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
    cl_int preconditioner;      /* refinement solvers: 0 none, 1 Jacobi, 2 learned FSAI,
                                   3 learned FSAI that fell back to Jacobi */
    /* Snapshot of the solver control's reference fields
     * (src/library/solvers/solver-control.hpp L108-127). */
    cl_int n_iters;
    cl_int max_iters;
    cl_double relative_tolerance;
    cl_double absolute_tolerance;
    cl_double initial_residual;
    cl_double current_residual;
    /* Learned-FSAI solver: residual reduction rate -ln(r_end / r_start) per
     * second of the pilot race, for the factor and for Jacobi (0 if not run). */
    cl_double pilot_rate_fsai;
    cl_double pilot_rate_jacobi;
    /* CG solvers (clsparse[S/D]csrcg and the refinement solvers): diagonals
     * of the diagonal storage A's values were moved to for the iterations of
     * the solve (0: CSR-Adaptive SpMV); learned-FSAI solver: diagonals of the
     * factor's storage (0: CSR SpMVs or no factor). */
    cl_int a_diagonals;
    cl_int g_diagonals;
    /* Learned-FSAI mixed solver, factor in diagonal storage: Hutchinson
     * estimates of ||F^T A F - I||_F^2 / n for the factor and for Jacobi
     * (F = D^-1/2) taken before the first solve with a new factor (0 if not
     * taken); the pilot race is skipped when the factor's is at most half of
     * Jacobi's. */
    cl_double loss_fsai;
    cl_double loss_jacobi;
} clsparseAmdSolverStats;

CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGetSolverStats(clSParseSolverControl solverControl, clsparseAmdSolverStats* out);

/* Mixed-precision Jacobi-PCG with FP64 residual replacement (no reference
 * counterpart). A, x and b are FP64. One CG run keeps A's values and the
 * vectors r, z, p, A p in FP32 and accumulates every dot product and norm in
 * FP64; the solution x is updated in FP64. Whenever the recurrence residual
 * has fallen tenfold since the last replacement, or claims convergence, the
 * residual is replaced by the FP64 true residual b - A x (FP64 CSR-Adaptive
 * SpMV) and the search direction is re-based on it, without a restart
 * (reliable updates; CLSPARSE_AMD_RELIABLE_DELTA sets the factor). Stops when
 * the FP64 true relative residual ||b - A x||_2 / ||b||_2 <= relTol of the
 * solver control, or when maxIters iterations have run. The preconditioner
 * field of the solver control selects DIAGONAL or NOPRECOND. Statistics:
 * iterations, residual replacements (outer_iterations), true residual. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdDcsrcgMixed(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                           clSParseSolverControl solverControl, clsparseControl control);

/* The same solver with FP64 values and vectors: FP64 PCG judged by the FP64
 * true relative residual (the baseline the mixed solver is set beside). */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdDcsrcgRefined(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                             clSParseSolverControl solverControl, clsparseControl control);

/* ---- learned factorized sparse approximate inverse ----------------------
 * G is lower triangular on the pattern of tril(A), G_ij = (delta_ij + o_ij) /
 * sqrt(a_ii), where o_ij is the output of a per-nonzero MLP (16 features ->
 * 64 -> 64 -> 1, ReLU; BF16 weights and activations on MFMA, FP32
 * accumulation). The preconditioner is M = G G^T. A network with zero output
 * gives exactly the Jacobi preconditioner; so does a NULL network. */
typedef struct _clsparseAmdMlp* clsparseAmdMlp;
typedef struct _clsparseAmdFsai* clsparseAmdFsai;

/* Network on the current device: He-normal hidden layers from `seed`, output
 * layer zero (so a new network reproduces Jacobi). */
CLSPARSE_EXPORT clsparseStatus clsparseAmdMlpCreate(clsparseAmdMlp* net, cl_ulong seed);
CLSPARSE_EXPORT clsparseStatus clsparseAmdMlpDestroy(clsparseAmdMlp net);
/* Number of FP32 parameters: W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3. */
CLSPARSE_EXPORT cl_uint clsparseAmdMlpNumParams(void);
CLSPARSE_EXPORT clsparseStatus clsparseAmdMlpGetParams(clsparseAmdMlp net, cl_float* host);
CLSPARSE_EXPORT clsparseStatus clsparseAmdMlpSetParams(clsparseAmdMlp net, const cl_float* host);
/* Parameter gradient of the last clsparseAmdFsaiTrainStep. */
CLSPARSE_EXPORT clsparseStatus clsparseAmdMlpGetGrad(clsparseAmdMlp net, cl_float* host);

/* Builds G's pattern (and that of G^T) from a square FP64 matrix A with
 * metadata-free CSR input, the diagonal scaling and the per-nonzero features
 * (BF16, 16 per nonzero of G), all on the device. G starts as Jacobi.
 * clsparseInvalidValue for a row pointer that decreases or passes nnz, or a
 * column index >= n; clsparseInvalidSize for n or nnz >= 2^31. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdFsaiCreate(clsparseAmdFsai* fsai, const clsparseCsrMatrix* A, clsparseControl control);
/* Recomputes scaling and features for new values of A on the same pattern. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdFsaiUpdateValues(clsparseAmdFsai fsai, const clsparseCsrMatrix* A, clsparseControl control);
CLSPARSE_EXPORT clsparseStatus clsparseAmdFsaiDestroy(clsparseAmdFsai fsai);
/* Evaluates the network (NULL = Jacobi) and assembles G and G^T in FP64 and FP32. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdFsaiApplyNet(clsparseAmdFsai fsai, clsparseAmdMlp net, clsparseControl control);
/* Host copies of G (CSR: num_rows + 1 row pointers, *nnz columns and FP64
 * values); any output may be NULL. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdFsaiGetValues(clsparseAmdFsai fsai, cl_uint* row_pointer, cl_uint* col_indices,
                             cl_double* values, cl_uint* nnz);
CLSPARSE_EXPORT clsparseStatus clsparseAmdFsaiGetFeatures(clsparseAmdFsai fsai, uint16_t* bf16);
/* Network output o per nonzero of G from the last evaluation. */
CLSPARSE_EXPORT clsparseStatus clsparseAmdFsaiGetOutput(clsparseAmdFsai fsai, cl_float* out);
/* Storage of the BF16 factor used by the FP32 solver: *ndiag = number of
 * diagonals when G's pattern lies on at most 8 of them (structured grids) and
 * the factor is applied from diagonal storage, else 0 (CSR SpMVs). offsets
 * (8 entries, may be NULL) receives the diagonals' column - row offsets. */
CLSPARSE_EXPORT clsparseStatus clsparseAmdFsaiGetDiagonals(clsparseAmdFsai fsai, cl_int* ndiag,
                                                           cl_int* offsets);

/* One training step on A: assembles G from the network, estimates
 * L = ||(G^T A G - I) Z||_F^2 / (probes * n) with `probes` Rademacher columns
 * Z drawn from `seed`, back-propagates dL/dG through the network, and, when
 * lr > 0, applies one Adam step (beta 0.9 / 0.999). loss and grad_g (dL/dG
 * per nonzero of G, FP64) may be NULL. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdFsaiTrainStep(clsparseAmdFsai fsai, clsparseAmdMlp net, const clsparseCsrMatrix* A,
                             cl_uint probes, cl_ulong seed, cl_float lr, clsparseControl control,
                             cl_double* loss, cl_double* grad_g);

/* The mixed-precision solver of clsparseAmdDcsrcgMixed with M = G G^T in
 * place of Jacobi: two CSR-Adaptive SpMVs per iteration, with G's values
 * stored in BF16 (M stays symmetric positive definite; CLSPARSE_AMD_FSAI_BF16=0
 * keeps FP32). Jacobi takes over for the rest of the solve when G holds a
 * non-finite value or lacks a diagonal entry; when the recurrence residual
 * turns non-finite (x goes back to the last replacement point); when the
 * residual fails to halve within 16 chunks of 32 iterations; or when a pilot
 * race -- 2 chunks with G, then 2 with Jacobi from the same start, each
 * timed -- finds Jacobi reducing the residual more than 1.15 times faster per
 * second (CLSPARSE_AMD_RACE_MARGIN sets the factor: Jacobi's first
 * iterations run ahead of its whole-solve rate; CLSPARSE_AMD_FSAI_POLICY=always
 * skips the race). The race's outcome is kept with the factor and reused by
 * later solves until the factor changes (clsparseAmdFsaiApplyNet,
 * clsparseAmdFsaiUpdateValues). The FP64 true residual decides convergence;
 * stats.preconditioner reports 2 (G to the end) or 3 (Jacobi took over) and
 * stats.pilot_rate_* the two measured rates. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdDcsrcgFsai(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                          clsparseAmdFsai fsai, clSParseSolverControl solverControl,
                          clsparseControl control);
/* The same with FP64 inner solves. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdDcsrcgFsai64(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                            clsparseAmdFsai fsai, clSParseSolverControl solverControl,
                            clsparseControl control);

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

/* y = A x for a square stencil matrix through the diagonal storage the CG
 * solvers (clsparse[S/D]csrcg, the refinement, learned-FSAI and BiCGStab
 * solvers) move its values to: the same detection, fill and SpMV kernels, so
 * that the solvers' SpMV can be checked against clsparse[S/D]csrmv. fused != 0
 * runs the FP32 solvers' kernel that forms its operand on the fly (first
 * iteration: the operand is x itself). *ndiag receives the number of
 * diagonals; 0 when A is not a stencil of at most 9 diagonals, and then y is
 * not written. No reference counterpart. */
CLSPARSE_EXPORT clsparseStatus clsparseAmdScsrmvDiag(const clsparseCsrMatrix* A, const cldenseVector* x,
                                                     cldenseVector* y, cl_int fused, cl_int* ndiag,
                                                     const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus clsparseAmdDcsrmvDiag(const clsparseCsrMatrix* A, const cldenseVector* x,
                                                     cldenseVector* y, cl_int fused, cl_int* ndiag,
                                                     const clsparseControl control);

/* sum = sum over the stored entries of values[j] * x[col_indices[j]] =
 * 1^T (A x), accumulated in FP64 in a fixed order and written to the device
 * scalar in the matrix's precision. One pass in storage order that ignores the
 * rows: the checksum of an SpMV, and the time of the loads an SpMV of A cannot
 * avoid (values, columns and x gathers, no row structure, no y). */
CLSPARSE_EXPORT clsparseStatus clsparseAmdScsrSumAx(const clsparseCsrMatrix* A, const cldenseVector* x,
                                                    clsparseScalar* sum, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus clsparseAmdDcsrSumAx(const clsparseCsrMatrix* A, const cldenseVector* x,
                                                    clsparseScalar* sum, const clsparseControl control);

/* sum = sum over i < count of x[h(i + seed) mod window]: independent gathers
 * of x at hashed positions inside its first `window` values (a power of two,
 * at most x's length; count < 2^32), accumulated in FP64. h is the 32-bit
 * xor-shift-multiply hash (MurmurHash3 finaliser, constants 0x7feb352d and
 * 0x846ca68b). The rate at which the device gathers into a window of a given
 * size: the floor that scattered columns set for any CSR SpMV. No reference
 * counterpart. */
CLSPARSE_EXPORT clsparseStatus clsparseAmdSgatherSum(const cldenseVector* x, cl_uint window, cl_ulong count,
                                                     cl_uint seed, clsparseScalar* sum,
                                                     const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus clsparseAmdDgatherSum(const cldenseVector* x, cl_uint window, cl_ulong count,
                                                     cl_uint seed, clsparseScalar* sum,
                                                     const clsparseControl control);

/* clsparseAmdDcsrSumAx without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Its kernels keep their
 * per-workgroup partials in a library-owned device array of 4096 doubles per
 * GPU (created with the first control object on that GPU and held until the
 * process ends); the grid never needs more, so no size is refused for it.
 * Every call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams, or from different threads) race and must be ordered by
 * the caller. Same arguments, checks and result as clsparseAmdDcsrSumAx. */
CLSPARSE_EXPORT clsparseStatus clsparseAmdDcsrSumAxNoAlloc(const clsparseCsrMatrix* A, const cldenseVector* x,
                                                           clsparseScalar* sum, const clsparseControl control);

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
 * kappa), scaled per axis by (ax, ay, az). SPD, diagonally dominant, and
 * exactly symmetric in floating point. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenDiffusion3D(cl_uint nx, cl_uint ny, cl_uint nz, cl_double sigma,
                              cl_double ax, cl_double ay, cl_double az, cl_ulong seed,
                              clsparseAmdHostCsr* out);

/* 2-D anisotropic diffusion -div(kappa R diag(1, 1/ratio) R^T grad u) on an
 * nx x ny grid of interior nodes, bilinear (Q1) finite elements, Dirichlet
 * boundary: a symmetric positive definite 9-point stencil. log(kappa) is a
 * Gaussian random field of standard deviation sigma and correlation length
 * corr_len (cells), sampled with 128 seeded random Fourier modes; R rotates
 * by theta. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenAnisoDiffusion2D(cl_uint nx, cl_uint ny, cl_double corr_len, cl_double sigma,
                                   cl_double ratio, cl_double theta, cl_ulong seed,
                                   clsparseAmdHostCsr* out);

/* The FP64 CSR Matrix Market reader into host memory (no device): the same
 * parsing, symmetric expansion, zero filtering and ordering as
 * clsparseDCsrMatrixfromFile. Free with clsparseAmdHostCsrFree. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdCsrMatrixfromFileHost(const char* filePath, cl_bool read_explicit_zeroes,
                                     clsparseAmdHostCsr* out);

/* Fills out[0..n) with U[lo, hi) from a SplitMix64 stream of `seed`. */
CLSPARSE_EXPORT clsparseStatus
    clsparseAmdGenUniform(cl_double* out, size_t n, cl_double lo, cl_double hi, cl_ulong seed);

#ifdef __cplusplus
}
#endif

#endif /* CLSPARSE_AMD_EXT_H_ */
