/*
 * clSPARSE.h -- C ABI of the MI355X (gfx950) CSR-Adaptive SpMV / CG library.
 *
 * Drop-in for the clSPARSE hot path: the same entry-point names, argument
 * order, struct field order and numeric status codes as the reference's
 * public header (src/include/clSPARSE.h, src/include/clSPARSE-1x.h), with
 * OpenCL handles replaced by HIP objects:
 *
 *   cl_mem            -> void*        (a HIP device pointer, same size on LP64)
 *   cl_command_queue  -> hipStream_t  (0 = the legacy default stream)
 *   cl_event          -> hipEvent_t
 *   cl_int/uint/bool/float/double -> fixed-width C types
 *
 * Every scalar argument (alpha, beta, results) lives in device memory, as in
 * the reference ("pointer mode device"). Element offsets (off_*) are honoured
 * everywhere, including the adaptive SpMV and CG paths where the reference
 * silently ignores them (src/library/blas2/csrmv-adaptive.hpp L93-97,
 * src/library/solvers/conjugate-gradients.hpp L65-66).
 *
 * The library is built for gfx950 only and links against the HIP runtime.
 * It has no CPU fallback: every compute entry point launches HIP kernels.
 */
#ifndef CLSPARSE_AMD_CLSPARSE_H_
#define CLSPARSE_AMD_CLSPARSE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define CLSPARSE_EXPORT
#else
#define CLSPARSE_EXPORT __attribute__((visibility("default")))
#endif

/* Opaque HIP handle types, spelled exactly as hip_runtime_api.h spells them so
 * that including both headers is legal. */
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;

/* OpenCL scalar typedefs used by the reference signatures. */
typedef int32_t  cl_int;
typedef uint32_t cl_uint;
typedef uint32_t cl_bool;
typedef uint64_t cl_ulong;
typedef float    cl_float;
typedef double   cl_double;

/* Index type: 32-bit, as the reference (src/include/clSPARSE-xx.h L45-50). */
typedef cl_uint clsparseIdx_t;

/* Status codes: numeric values of src/include/clSPARSE.h L45-88 (the OpenCL
 * codes come from CL/cl.h). */
typedef enum clsparseStatus_
{
    clsparseSuccess                 = 0,
    clsparseInvalidValue            = -30,
    clsparseInvalidCommandQueue     = -36,
    clsparseInvalidContext          = -34,
    clsparseInvalidMemObject        = -38,
    clsparseInvalidDevice           = -33,
    clsparseInvalidEventWaitList    = -57,
    clsparseInvalidEvent            = -58,
    clsparseOutOfResources          = -5,
    clsparseOutOfHostMemory         = -6,
    clsparseInvalidOperation        = -59,
    clsparseCompilerNotAvailable    = -3,
    clsparseBuildProgramFailure     = -11,
    clsparseInvalidKernelArgs       = -52,

    clsparseNotImplemented          = -1024,
    clsparseNotInitialized          = -1023,
    clsparseStructInvalid           = -1022,
    clsparseInvalidSize             = -1021,
    clsparseInvalidMemObj           = -1020,
    clsparseInsufficientMemory      = -1019,
    clsparseInvalidControlObject    = -1018,
    clsparseInvalidFile             = -1017,
    clsparseInvalidFileFormat       = -1016,
    clsparseInvalidKernelExecution  = -1015,
    clsparseInvalidType             = -1014,

    clsparseInvalidSolverControlObject = -2048,
    clsparseInvalidSystemSize          = -2047,
    clsparseIterationsExceeded         = -2046,
    clsparseToleranceNotReached        = -2045,
    clsparseSolverError                = -2044
} clsparseStatus;

typedef enum _cldenseMajor
{
    rowMajor = 1,
    columnMajor
} cldenseMajor;

/* Data structures: field order of src/include/clSPARSE-1x.h L31-143. */
typedef struct clsparseScalar_
{
    void* value;                 /* device pointer */
    clsparseIdx_t off_value;     /* element offset added to value */
} clsparseScalar;

typedef struct cldenseVector_
{
    clsparseIdx_t num_values;
    void* values;                /* device pointer */
    clsparseIdx_t off_values;
} cldenseVector;

typedef struct clsparseCsrMatrix_
{
    clsparseIdx_t num_rows;
    clsparseIdx_t num_cols;
    clsparseIdx_t num_nonzeros;
    void* values;                /* T[num_nonzeros] */
    void* col_indices;           /* uint32[num_nonzeros], sorted per row not required */
    void* row_pointer;           /* uint32[num_rows + 1] */
    clsparseIdx_t off_values;
    clsparseIdx_t off_col_indices;
    clsparseIdx_t off_row_pointer;
    void* meta;                  /* library-owned CSR-Adaptive metadata */
} clsparseCsrMatrix;

typedef struct clsparseCooMatrix_
{
    clsparseIdx_t num_rows;
    clsparseIdx_t num_cols;
    clsparseIdx_t num_nonzeros;
    void* values;
    void* col_indices;
    void* row_indices;
    clsparseIdx_t off_values;
    clsparseIdx_t off_col_indices;
    clsparseIdx_t off_row_indices;
} clsparseCooMatrix;

typedef struct cldenseMatrix_
{
    clsparseIdx_t num_rows;
    clsparseIdx_t num_cols;
    clsparseIdx_t lead_dim;
    cldenseMajor major;
    void* values;
    clsparseIdx_t off_values;
} cldenseMatrix;

/* ---- library setup: src/include/clSPARSE.h L108-135 ---------------------- */
CLSPARSE_EXPORT clsparseStatus clsparseSetup(void);                       /* L108-109 */
CLSPARSE_EXPORT clsparseStatus clsparseTeardown(void);                    /* L119-120 */
CLSPARSE_EXPORT clsparseStatus clsparseGetVersion(cl_uint* major, cl_uint* minor,
                                                  cl_uint* patch, cl_uint* tweak); /* L134-135 */

/* ---- struct initialisers: src/include/clSPARSE.h L155-208 ----------------- */
CLSPARSE_EXPORT clsparseStatus clsparseInitScalar(clsparseScalar* scalar);          /* L155-156 */
CLSPARSE_EXPORT clsparseStatus clsparseInitVector(cldenseVector* vec);              /* L168-169 */
CLSPARSE_EXPORT clsparseStatus clsparseInitCooMatrix(clsparseCooMatrix* cooMatx);   /* L181-182 */
CLSPARSE_EXPORT clsparseStatus clsparseInitCsrMatrix(clsparseCsrMatrix* csrMatx);   /* L194-195 */
CLSPARSE_EXPORT clsparseStatus cldenseInitMatrix(cldenseMatrix* denseMatx);         /* L207-208 */

/* ---- control object: src/include/clSPARSE.h L224-304 ---------------------- */
typedef struct _clsparseControl* clsparseControl;

typedef struct _clsparseCreateResult
{
    clsparseStatus status;
    clsparseControl control;
} clsparseCreateResult;

typedef struct _clsparseEventResult
{
    clsparseStatus status;
    hipEvent_t event;
} clsparseEventResult;

/* L243-244: the queue becomes the HIP stream every call of this control is
 * ordered on. Fails with clsparseInvalidDevice if the current device is not a
 * gfx950 part. */
CLSPARSE_EXPORT clsparseCreateResult clsparseCreateControl(hipStream_t stream);
/* L256-257: async=false (default) makes every call synchronous with the host. */
CLSPARSE_EXPORT clsparseStatus clsparseEnableAsync(clsparseControl control, cl_bool async);
/* L270-271: compensated (two-sum / two-fma) accumulation inside SpMV. */
CLSPARSE_EXPORT clsparseStatus clsparseEnableExtendedPrecision(clsparseControl control,
                                                               cl_bool extPrecision);
/* L291-292: event recorded after the last operation issued on this control. */
CLSPARSE_EXPORT clsparseEventResult clsparseGetEvent(clsparseControl control);
/* L303-304 */
CLSPARSE_EXPORT clsparseStatus clsparseReleaseControl(clsparseControl control);

/* ---- iterative solvers: src/include/clSPARSE.h L321-448 ------------------- */
typedef enum _print_mode
{
    QUIET = 0,
    NORMAL,
    VERBOSE
} PRINT_MODE;

typedef enum _precond
{
    NOPRECOND = 0,
    DIAGONAL
} PRECONDITIONER;

typedef struct _solverControl* clSParseSolverControl;

typedef struct _clsparseCreateSolverResult
{
    clsparseStatus status;
    clSParseSolverControl control;
} clsparseCreateSolverResult;

CLSPARSE_EXPORT clsparseCreateSolverResult
    clsparseCreateSolverControl(PRECONDITIONER precond, cl_int maxIters,
                                cl_double relTol, cl_double absTol);          /* L369-371 */
CLSPARSE_EXPORT clsparseStatus
    clsparseReleaseSolverControl(clSParseSolverControl solverControl);        /* L382-383 */
CLSPARSE_EXPORT clsparseStatus
    clsparseSetSolverParams(clSParseSolverControl solverControl, PRECONDITIONER precond,
                            cl_int maxIters, cl_double relTol, cl_double absTol); /* L398-401 */
CLSPARSE_EXPORT clsparseStatus
    clsparseSolverPrintMode(clSParseSolverControl solverControl, PRINT_MODE mode); /* L413-414 */

/* L429-431 / L446-448: Jacobi- or un-preconditioned CG, L1 relative residual
 * stopping rule of src/library/solvers/solver-control.hpp L47-64. */
CLSPARSE_EXPORT clsparseStatus
    clsparseScsrcg(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                   clSParseSolverControl solverControl, clsparseControl control);
CLSPARSE_EXPORT clsparseStatus
    clsparseDcsrcg(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                   clSParseSolverControl solverControl, clsparseControl control);
/* BiCGStab with right Jacobi / no preconditioning, same solver control and
 * ||r||_1 / ||b||_1 stopping rule (src/include/clSPARSE.h L450-482). */
CLSPARSE_EXPORT clsparseStatus
    clsparseScsrbicgStab(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                         clSParseSolverControl solverControl, clsparseControl control);  /* L463-465 */
CLSPARSE_EXPORT clsparseStatus
    clsparseDcsrbicgStab(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                         clSParseSolverControl solverControl, clsparseControl control);  /* L480-482 */

/* ---- Matrix Market file input: src/include/clSPARSE.h L490-611 ------------
 * Coordinate .mtx files (real, integer or pattern; general or symmetric).
 * The caller allocates the device arrays (sizes from clsparseHeaderfromFile);
 * entries are sorted by (row, column), symmetric files are expanded, and
 * explicit zeroes are dropped unless read_explicit_zeroes is set. */
CLSPARSE_EXPORT clsparseStatus
    clsparseHeaderfromFile(clsparseIdx_t* nnz, clsparseIdx_t* row, clsparseIdx_t* col,
                           const char* filePath);  /* L508-509 */
CLSPARSE_EXPORT clsparseStatus
    clsparseSCooMatrixfromFile(clsparseCooMatrix* cooMatx, const char* filePath,
                               clsparseControl control, cl_bool read_explicit_zeroes);  /* L533-534 */
CLSPARSE_EXPORT clsparseStatus
    clsparseDCooMatrixfromFile(clsparseCooMatrix* cooMatx, const char* filePath,
                               clsparseControl control, cl_bool read_explicit_zeroes);  /* L559-560 */
CLSPARSE_EXPORT clsparseStatus
    clsparseSCsrMatrixfromFile(clsparseCsrMatrix* csrMatx, const char* filePath,
                               clsparseControl control, cl_bool read_explicit_zeroes);  /* L584-585 */
CLSPARSE_EXPORT clsparseStatus
    clsparseDCsrMatrixfromFile(clsparseCsrMatrix* csrMatx, const char* filePath,
                               clsparseControl control, cl_bool read_explicit_zeroes);  /* L610-611 */

/* ---- CSR-Adaptive metadata: src/include/clSPARSE.h L612-659 --------------- */
typedef struct _clsparseMetaSizeResult
{
    clsparseStatus status;
    clsparseIdx_t metaSize;
} clsparseMetaSizeResult;

/* L633-634: size (in 64-bit words) of the row-block buffer; the stored size
 * when meta already exists (src/library/internal/data-types/csr-meta.cpp L30-35). */
CLSPARSE_EXPORT clsparseMetaSizeResult
    clsparseCsrMetaSize(clsparseCsrMatrix* csrMatx, clsparseControl control);
/* L648-649: bit-exact row-block metadata of csr-meta.hpp L93-261. */
CLSPARSE_EXPORT clsparseStatus
    clsparseCsrMetaCreate(clsparseCsrMatrix* csrMatx, clsparseControl control);
/* L658-659 */
CLSPARSE_EXPORT clsparseStatus
    clsparseCsrMetaDelete(clsparseCsrMatrix* csrMatx);
/* Alias of clsparseCsrMetaCreate under the name used by the reference's
 * samples/sample-spmv.cpp L57 and src/tests/test_readMMcoo.cpp L93. */
CLSPARSE_EXPORT clsparseStatus
    clsparseCsrMetaCompute(clsparseCsrMatrix* csrMatx, clsparseControl control);

/* ---- dense BLAS-1: src/include/clSPARSE.h L694-1028 ----------------------- */
/* r = alpha * y                                     L694-713 */
CLSPARSE_EXPORT clsparseStatus cldenseSscale(cldenseVector* r, const clsparseScalar* alpha,
                                             const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDscale(cldenseVector* r, const clsparseScalar* alpha,
                                             const cldenseVector* y, const clsparseControl control);
/* r = alpha * x + y                                 L727-749 */
CLSPARSE_EXPORT clsparseStatus cldenseSaxpy(cldenseVector* r, const clsparseScalar* alpha,
                                            const cldenseVector* x, const cldenseVector* y,
                                            const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDaxpy(cldenseVector* r, const clsparseScalar* alpha,
                                            const cldenseVector* x, const cldenseVector* y,
                                            const clsparseControl control);
/* r = alpha * x + beta * y                          L762-789 */
CLSPARSE_EXPORT clsparseStatus cldenseSaxpby(cldenseVector* r, const clsparseScalar* alpha,
                                             const cldenseVector* x, const clsparseScalar* beta,
                                             const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDaxpby(cldenseVector* r, const clsparseScalar* alpha,
                                             const cldenseVector* x, const clsparseScalar* beta,
                                             const cldenseVector* y, const clsparseControl control);
/* s = sum(x)                                        L797-836 */
CLSPARSE_EXPORT clsparseStatus cldenseIreduce(clsparseScalar* s, const cldenseVector* x,
                                              const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseSreduce(clsparseScalar* s, const cldenseVector* x,
                                              const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDreduce(clsparseScalar* s, const cldenseVector* x,
                                              const clsparseControl control);
/* s = sum |x_i|                                     L838-862 */
CLSPARSE_EXPORT clsparseStatus cldenseSnrm1(clsparseScalar* s, const cldenseVector* x,
                                            const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDnrm1(clsparseScalar* s, const cldenseVector* x,
                                            const clsparseControl control);
/* s = sqrt(sum x_i^2)                               L864-889 */
CLSPARSE_EXPORT clsparseStatus cldenseSnrm2(clsparseScalar* s, const cldenseVector* x,
                                            const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDnrm2(clsparseScalar* s, const cldenseVector* x,
                                            const clsparseControl control);
/* r = <x, y>                                        L891-910 */
CLSPARSE_EXPORT clsparseStatus cldenseSdot(clsparseScalar* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDdot(clsparseScalar* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
/* element-wise r = x op y, op in {+, -, *, /}       L923-1028 */
CLSPARSE_EXPORT clsparseStatus cldenseSadd(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDadd(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseSsub(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDsub(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseSmul(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDmul(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseSdiv(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus cldenseDdiv(cldenseVector* r, const cldenseVector* x,
                                           const cldenseVector* y, const clsparseControl control);

/* ---- sparse BLAS-2: src/include/clSPARSE.h L1062-1091 --------------------- */
/* y = alpha * A * x + beta * y. meta == NULL selects CSR-Vector
 * (src/library/blas2/clsparse-csrmv.hpp L35-38), otherwise CSR-Adaptive;
 * meta with rowBlockSize 0 gives clsparseStructInvalid (L42-46). */
CLSPARSE_EXPORT clsparseStatus
    clsparseScsrmv(const clsparseScalar* alpha, const clsparseCsrMatrix* matx,
                   const cldenseVector* x, const clsparseScalar* beta,
                   cldenseVector* y, const clsparseControl control);
CLSPARSE_EXPORT clsparseStatus
    clsparseDcsrmv(const clsparseScalar* alpha, const clsparseCsrMatrix* matx,
                   const cldenseVector* x, const clsparseScalar* beta,
                   cldenseVector* y, const clsparseControl control);

/* ---- sparse BLAS-3: src/include/clSPARSE.h L1145-1187 ----------------------
 * C = alpha * A * B + beta * C with B and C row-major (lead_dim = row stride;
 * the major field is not read, as in the reference). Size mismatches between
 * A, B and C give clsparseInvalidSize. */
CLSPARSE_EXPORT clsparseStatus
    clsparseScsrmm(const clsparseScalar* alpha, const clsparseCsrMatrix* sparseMatA,
                   const cldenseMatrix* denseMatB, const clsparseScalar* beta,
                   cldenseMatrix* denseMatC, const clsparseControl control);  /* L1159-1165 */
CLSPARSE_EXPORT clsparseStatus
    clsparseDcsrmm(const clsparseScalar* alpha, const clsparseCsrMatrix* sparseMatA,
                   const cldenseMatrix* denseMatB, const clsparseScalar* beta,
                   cldenseMatrix* denseMatC, const clsparseControl control);  /* L1181-1187 */

#ifdef __cplusplus
}
#endif

#endif /* CLSPARSE_AMD_CLSPARSE_H_ */
