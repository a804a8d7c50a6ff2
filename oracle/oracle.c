/*
 * oracle.c -- CPU restatement of the clSPARSE hot path. TEST INFRASTRUCTURE.
 *
 * Used only as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. Nothing in the product library links,
 * loads or calls this file.
 *
 * Functions and the reference code they restate:
 *   orc_row_blocks   src/library/internal/data-types/csr-meta.hpp L93-269
 *                    (ComputeRowBlocks / ComputeRowBlocksSize), with
 *                    numThreadsForReduction L52-68 (GCC branch).
 *   orc_csrmv_*      y = alpha*A*x + beta*y as csrmv_adaptive.cl / csrmv_general.cl
 *                    define it (beta == 0 never reads y, L481-483), evaluated
 *                    in long double with Neumaier compensation: an FP64 oracle.
 *   orc_cg_*         src/library/solvers/conjugate-gradients.hpp L47-229 and
 *                    solver-control.hpp L47-64 (L1 residual, nIters quirk),
 *                    Jacobi of preconditioners/diagonal.hpp L45-77; dot and
 *                    norm reductions in double.
 *   orc_spmv_omp_d   plain FP64 / FP32 row-parallel CSR SpMV (OpenMP): the CPU
 *   orc_spmv_omp_s   baselines timed beside the GPU.
 *   orc_cg_omp_d     the FP64 Jacobi-PCG recurrence above on all host cores
 *                    (OpenMP), run to a tolerance: the solvers' CPU baseline.
 *
 * Parity pins (which checks fix each function) are listed in the header of
 * tests/test_oracle.py.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- metadata ------------------------------------------------------------ */

static uint64_t orc_nt(uint64_t num_rows)
{
    /* 256 >> (8*sizeof(int) - __builtin_clz(num_rows - 1)); the argument is
     * truncated to unsigned int as the builtin's parameter type does. */
    unsigned int a = (unsigned int)(num_rows - 1);
    int sh = 32 - __builtin_clz(a);
    return (uint64_t)(256 >> (sh & 31));
}

uint64_t orc_row_blocks(const uint32_t* rowDelimiters, uint32_t nRows, uint64_t* rowBlocks)
{
    const int allocate = rowBlocks != NULL;
    const uint64_t blkSize = 1024, blkMultiplier = 3, rows_for_vector = 1;
    uint64_t* base = rowBlocks;
    uint64_t total_row_blocks = 1;
    uint64_t sum = 0, i, last_i = 0;
    uint32_t consecutive_long_rows = 0;

    if (allocate) {
        *rowBlocks = 0;
        rowBlocks++;
    }
    for (i = 1; i <= nRows; i++) {
        uint32_t row_length = rowDelimiters[i] - rowDelimiters[i - 1];
        sum += row_length;

        if (row_length > 128)
            consecutive_long_rows++;
        else if (consecutive_long_rows > 0) {
            if (row_length < 32)
                consecutive_long_rows = (uint32_t)-1;
            else
                consecutive_long_rows++;
        }

        if (consecutive_long_rows == 1) {
            if (i - last_i > 1) {
                if (allocate) {
                    *rowBlocks = (i - 1) << 32;
                    if (((i - 1) - last_i) > rows_for_vector)
                        *(rowBlocks - 1) |= orc_nt((i - 1) - last_i);
                    rowBlocks++;
                }
                total_row_blocks++;
                last_i = i - 1;
                sum = row_length;
            }
        } else if (consecutive_long_rows == (uint32_t)-1) {
            if (allocate) {
                *rowBlocks = (i - 1) << 32;
                if (((i - 1) - last_i) > rows_for_vector)
                    *(rowBlocks - 1) |= orc_nt((i - 1) - last_i);
                rowBlocks++;
            }
            total_row_blocks++;
            last_i = i - 1;
            sum = row_length;
            consecutive_long_rows = 0;
        }

        if ((i - last_i == 1) && sum > blkSize) {
            int numWGReq = (int)ceil((double)row_length / (double)(blkMultiplier * blkSize));
            numWGReq = (numWGReq < (1 << 24)) ? numWGReq : (1 << 24);
            if (allocate) {
                for (int w = 1; w < numWGReq; w++) {
                    *rowBlocks = ((i - 1) << 32) | (uint64_t)w;
                    rowBlocks++;
                }
                *rowBlocks = i << 32;
                rowBlocks++;
            }
            total_row_blocks += (uint64_t)numWGReq;
            last_i = i;
            sum = 0;
            consecutive_long_rows = 0;
        } else if ((i - last_i > 1) && sum > blkSize) {
            i--;
            if (allocate) {
                *rowBlocks = i << 32;
                if ((i - last_i) > rows_for_vector) *(rowBlocks - 1) |= orc_nt(i - last_i);
                rowBlocks++;
            }
            total_row_blocks++;
            last_i = i;
            sum = 0;
            consecutive_long_rows = 0;
        } else if (sum == blkSize) {
            if (allocate) {
                *rowBlocks = i << 32;
                if ((i - last_i) > rows_for_vector) *(rowBlocks - 1) |= orc_nt(i - last_i);
                rowBlocks++;
            }
            total_row_blocks++;
            last_i = i;
            sum = 0;
            consecutive_long_rows = 0;
        }
    }

    if (allocate && (*(rowBlocks - 1) >> 32) != (uint64_t)nRows) {
        *rowBlocks = (uint64_t)nRows << 32;
        if (((uint64_t)nRows - last_i) > rows_for_vector) *(rowBlocks - 1) |= orc_nt(i - last_i);
        rowBlocks++;
    }
    total_row_blocks++;

    if (allocate) return 2 * (uint64_t)(rowBlocks - base);
    return 2 * total_row_blocks;
}

/* ---- SpMV ------------------------------------------------------------------ */

/* Compensated (Neumaier) accumulation in long double. */
typedef struct {
    long double s, c;
} ksum;

static void ks_add(ksum* k, long double v)
{
    long double t = k->s + v;
    if (fabsl(k->s) >= fabsl(v))
        k->c += (k->s - t) + v;
    else
        k->c += (v - t) + k->s;
    k->s = t;
}

/* y_out[i] = alpha * sum_j a_ij x_j + beta * y_in[i]  (y_in unread if beta == 0).
 * scale[i] = |alpha| * sum_j |a_ij x_j| + |beta y_in[i]|: the magnitude the
 * per-row relative error is measured against. */
void orc_csrmv_d(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v,
                 const double* x, double alpha, double beta, const double* y_in, double* y_out,
                 double* scale)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024)
#endif
    for (uint32_t r = 0; r < n; ++r) {
        ksum s = {0, 0};
        long double mag = 0;
        for (uint32_t j = rp[r]; j < rp[r + 1]; ++j) {
            /* A double product needs 106 bits; long double keeps 64, so the
             * rounding error of the product (exact, via fmal) is added too. */
            long double p = (long double)v[j] * (long double)x[ci[j]];
            long double e = fmal((long double)v[j], (long double)x[ci[j]], -p);
            ks_add(&s, p);
            ks_add(&s, e);
            mag += fabsl(p);
        }
        long double out = (long double)alpha * (s.s + s.c);
        long double m = fabsl((long double)alpha) * mag;
        if (beta != 0.0) {
            out += (long double)beta * (long double)y_in[r];
            m += fabsl((long double)beta * (long double)y_in[r]);
        }
        y_out[r] = (double)out;
        if (scale) scale[r] = (double)m;
    }
}

void orc_csrmv_s(uint32_t n, const uint32_t* rp, const uint32_t* ci, const float* v,
                 const float* x, float alpha, float beta, const float* y_in, double* y_out,
                 double* scale)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024)
#endif
    for (uint32_t r = 0; r < n; ++r) {
        ksum s = {0, 0};
        long double mag = 0;
        for (uint32_t j = rp[r]; j < rp[r + 1]; ++j) {
            long double p = (long double)v[j] * (long double)x[ci[j]];
            ks_add(&s, p);
            mag += fabsl(p);
        }
        long double out = (long double)alpha * (s.s + s.c);
        long double m = fabsl((long double)alpha) * mag;
        if (beta != 0.0f) {
            out += (long double)beta * (long double)y_in[r];
            m += fabsl((long double)beta * (long double)y_in[r]);
        }
        y_out[r] = (double)out;
        if (scale) scale[r] = (double)m;
    }
}

/* Plain row-parallel FP64 CSR SpMV (CPU baseline). */
void orc_spmv_omp_d(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v,
                    const double* x, double* y, int threads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads)
#endif
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        double s = 0.0;
        for (uint32_t j = rp[r]; j < rp[r + 1]; ++j) s += v[j] * x[ci[j]];
        y[r] = s;
    }
    (void)threads;
}

/* Plain row-parallel FP32 CSR SpMV with FP32 values and sums (CPU baseline of
 * the FP32 workloads). */
void orc_spmv_omp_s(uint32_t n, const uint32_t* rp, const uint32_t* ci, const float* v,
                    const float* x, float* y, int threads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 512) num_threads(threads)
#endif
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        float s = 0.0f;
        for (uint32_t j = rp[r]; j < rp[r + 1]; ++j) s += v[j] * x[ci[j]];
        y[r] = s;
    }
    (void)threads;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- BLAS-1 ---------------------------------------------------------------- */

double orc_dot_d(size_t n, const double* x, const double* y)
{
    ksum s = {0, 0};
    for (size_t i = 0; i < n; ++i) ks_add(&s, (long double)x[i] * (long double)y[i]);
    return (double)(s.s + s.c);
}

double orc_nrm1_d(size_t n, const double* x)
{
    ksum s = {0, 0};
    for (size_t i = 0; i < n; ++i) ks_add(&s, fabsl((long double)x[i]));
    return (double)(s.s + s.c);
}

double orc_nrm2_d(size_t n, const double* x)
{
    ksum s = {0, 0};
    for (size_t i = 0; i < n; ++i) ks_add(&s, (long double)x[i] * (long double)x[i]);
    return (double)sqrtl(s.s + s.c);
}

/* ---- CG ---------------------------------------------------------------------- */

typedef struct {
    int status;          /* 0 = clsparseSuccess */
    int n_iters;         /* solverControl->nIters after the call */
    int executed;        /* loop iterations run */
    int early_exit;      /* 1: returned before the loop */
    double initial_residual;
    double current_residual;
    double rel_tol_out;  /* tolerances after the call (b == 0 zeroes them) */
    double abs_tol_out;
} orc_cg_result;

/* fused multiply-add in the vector's own precision */
#define ORC_FMA(a, b, c) _Generic((a), float: fmaf, default: fma)((a), (b), (c))

#define ORC_CG_BODY(T)                                                                      \
    orc_cg_result res;                                                                      \
    memset(&res, 0, sizeof(res));                                                           \
    res.rel_tol_out = relTol;                                                               \
    res.abs_tol_out = absTol;                                                               \
    res.n_iters = nIters_prior;                                                             \
    double norm_b = 0;                                                                      \
    for (uint32_t i = 0; i < n; ++i) norm_b += fabs((double)b[i]);                          \
    if (norm_b == 0) {                                                                      \
        res.n_iters = 0;                                                                    \
        res.rel_tol_out = 0;                                                                \
        res.abs_tol_out = 0;                                                                \
        for (uint32_t i = 0; i < n; ++i) x[i] = b[i];                                       \
        res.early_exit = 1;                                                                 \
        return res;                                                                         \
    }                                                                                       \
    T* y = (T*)calloc(n, sizeof(T));                                                        \
    T* z = (T*)calloc(n, sizeof(T));                                                        \
    T* r = (T*)calloc(n, sizeof(T));                                                        \
    T* p = (T*)calloc(n, sizeof(T));                                                        \
    T* invD = (T*)malloc(n * sizeof(T));                                                    \
    for (uint32_t i = 0; i < n; ++i) {                                                      \
        T d = (T)1;                                                                         \
        for (uint32_t j = rp[i]; j < rp[i + 1]; ++j)                                        \
            if (ci[j] == i) {                                                               \
                d = (T)1 / v[j];                                                            \
                break;                                                                      \
            }                                                                               \
        invD[i] = d;                                                                        \
    }                                                                                       \
    /* y = A x; r = b - y */                                                                \
    for (uint32_t i = 0; i < n; ++i) {                                                      \
        T s = 0;                                                                            \
        for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) s = ORC_FMA(v[j], x[ci[j]], s);         \
        y[i] = s;                                                                           \
        r[i] = b[i] - y[i];                                                                 \
    }                                                                                       \
    double norm_r = 0;                                                                      \
    for (uint32_t i = 0; i < n; ++i) norm_r += fabs((double)r[i]);                          \
    double residuum = norm_r / norm_b;                                                      \
    res.initial_residual = residuum;                                                        \
    res.current_residual = residuum;                                                        \
    if (residuum <= relTol || residuum <= absTol * residuum || nIters_prior >= maxIters) {  \
        res.n_iters = 0;                                                                    \
        res.early_exit = 1;                                                                 \
        goto done;                                                                          \
    }                                                                                       \
    for (uint32_t i = 0; i < n; ++i) {                                                      \
        z[i] = precond ? r[i] * invD[i] : r[i];                                             \
        p[i] = z[i];                                                                        \
    }                                                                                       \
    double rz = 0;                                                                          \
    for (uint32_t i = 0; i < n; ++i) rz += (double)r[i] * (double)z[i];                     \
    int iteration = 0;                                                                      \
    for (;;) {                                                                              \
        res.n_iters = iteration;                                                            \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            T s = 0;                                                                        \
            for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) s = ORC_FMA(v[j], p[ci[j]], s);     \
            y[i] = s;                                                                       \
        }                                                                                   \
        double yp = 0;                                                                      \
        for (uint32_t i = 0; i < n; ++i) yp += (double)y[i] * (double)p[i];                 \
        const T alpha = (T)(rz / yp);                                                       \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            x[i] = x[i] + alpha * p[i];                                                     \
            r[i] = r[i] - alpha * y[i];                                                     \
            z[i] = precond ? r[i] * invD[i] : r[i];                                         \
        }                                                                                   \
        const double rz_old = rz;                                                           \
        rz = 0;                                                                             \
        for (uint32_t i = 0; i < n; ++i) rz += (double)r[i] * (double)z[i];                 \
        const T beta = (T)(rz / rz_old);                                                    \
        for (uint32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];                         \
        norm_r = 0;                                                                         \
        for (uint32_t i = 0; i < n; ++i) norm_r += fabs((double)r[i]);                      \
        residuum = norm_r / norm_b;                                                         \
        iteration++;                                                                        \
        res.current_residual = residuum;                                                    \
        if (residuum <= relTol || residuum <= absTol * res.initial_residual ||              \
            res.n_iters >= maxIters)                                                        \
            break;                                                                          \
    }                                                                                       \
    res.executed = iteration;                                                               \
done:                                                                                       \
    free(y);                                                                                \
    free(z);                                                                                \
    free(r);                                                                                \
    free(p);                                                                                \
    free(invD);                                                                             \
    return res;

orc_cg_result orc_cg_d(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v,
                       const double* b, double* x, int precond, int maxIters, double relTol,
                       double absTol, int nIters_prior)
{
    ORC_CG_BODY(double)
}

orc_cg_result orc_cg_s(uint32_t n, const uint32_t* rp, const uint32_t* ci, const float* v,
                       const float* b, float* x, int precond, int maxIters, double relTol,
                       double absTol, int nIters_prior)
{
    ORC_CG_BODY(float)
}

/* ---- CG on all host cores (CPU baseline) ----------------------------------------
 * The FP64 Jacobi-PCG recurrence of conjugate-gradients.hpp L57-227 with every
 * vector loop and reduction split over `threads` OpenMP threads (loops fused
 * where the recurrence allows: SpMV with <y,p>, the x / r / z update with <r,z>
 * and the norm). crit 0: the reference's ||r||_1 / ||b||_1 <= relTol; crit 1:
 * ||r||_2 / ||b||_2 <= relTol (the stopping rule of the FP64 refinement solver
 * it is timed against). Returns the iterations run; x in/out (x0 = x). The
 * reduction order depends on the thread count: a timing baseline, not a parity
 * oracle (orc_cg_d is that). */
int orc_cg_omp_d(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v,
                 const double* b, double* x, int maxIters, double relTol, int crit, int threads,
                 double* final_residual)
{
    double* y = (double*)malloc((size_t)n * sizeof(double));
    double* z = (double*)malloc((size_t)n * sizeof(double));
    double* r = (double*)malloc((size_t)n * sizeof(double));
    double* p = (double*)malloc((size_t)n * sizeof(double));
    double* invD = (double*)malloc((size_t)n * sizeof(double));
    double norm_b = 0, norm_r = 0, rz = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : norm_b, norm_r, rz)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double d = 1.0, s = 0.0;
        for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) {
            if (ci[j] == (uint32_t)i && d == 1.0) d = 1.0 / v[j];
            s = fma(v[j], x[ci[j]], s);
        }
        invD[i] = d;
        r[i] = b[i] - s;
        z[i] = r[i] * d;
        p[i] = z[i];
        rz += r[i] * z[i];
        norm_b += crit ? b[i] * b[i] : fabs(b[i]);
        norm_r += crit ? r[i] * r[i] : fabs(r[i]);
    }
    if (crit) {
        norm_b = sqrt(norm_b);
        norm_r = sqrt(norm_r);
    }
    double res = norm_b > 0 ? norm_r / norm_b : 0.0;
    int it = 0;
    while (res > relTol && it < maxIters) {
        double yp = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : yp)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            double s = 0.0;
            for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) s = fma(v[j], p[ci[j]], s);
            y[i] = s;
            yp += s * p[i];
        }
        const double alpha = rz / yp;
        double rz_new = 0, nr = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) reduction(+ : rz_new, nr)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * y[i];
            z[i] = r[i] * invD[i];
            rz_new += r[i] * z[i];
            nr += crit ? r[i] * r[i] : fabs(r[i]);
        }
        const double beta = rz_new / rz;
        rz = rz_new;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
        for (int64_t i = 0; i < (int64_t)n; ++i) p[i] = z[i] + beta * p[i];
        res = (crit ? sqrt(nr) : nr) / norm_b;
        ++it;
    }
    if (final_residual) *final_residual = res;
    free(y);
    free(z);
    free(r);
    free(p);
    free(invD);
    (void)threads;
    return it;
}

/* ---- BiCGStab ------------------------------------------------------------------
 * Restates bicgStab<T> of src/library/solvers/biconjugate-gradients-stabilized.hpp
 * L41-268 (right preconditioning with M = Jacobi or identity, L1 residual
 * relative to ||b||_1, two finished() checks per iteration; finished uses the
 * solver control's nIters, which holds the value from before the call until
 * the end of the first iteration). Dots and norms accumulate in double. */

#define ORC_BICG_BODY(T)                                                                    \
    orc_cg_result res;                                                                      \
    memset(&res, 0, sizeof(res));                                                           \
    res.rel_tol_out = relTol;                                                               \
    res.abs_tol_out = absTol;                                                               \
    res.n_iters = nIters_prior;                                                             \
    double norm_b = 0;                                                                      \
    for (uint32_t i = 0; i < n; ++i) norm_b += fabs((double)b[i]);                          \
    if (norm_b <= (double)ORC_MIN(T)) { /* L75-83 */                                        \
        res.n_iters = 0;                                                                    \
        res.rel_tol_out = 0;                                                                \
        res.abs_tol_out = 0;                                                                \
        for (uint32_t i = 0; i < n; ++i) x[i] = b[i];                                       \
        res.early_exit = 1;                                                                 \
        return res;                                                                         \
    }                                                                                       \
    T* invD = (T*)malloc(n * sizeof(T));                                                    \
    T* r = (T*)calloc(n, sizeof(T));                                                        \
    T* rs = (T*)calloc(n, sizeof(T));                                                       \
    T* p = (T*)calloc(n, sizeof(T));                                                        \
    T* s = (T*)calloc(n, sizeof(T));                                                        \
    T* Mp = (T*)calloc(n, sizeof(T));                                                       \
    T* AMp = (T*)calloc(n, sizeof(T));                                                      \
    T* Ms = (T*)calloc(n, sizeof(T));                                                       \
    T* AMs = (T*)calloc(n, sizeof(T));                                                      \
    for (uint32_t i = 0; i < n; ++i) {                                                      \
        T d = (T)1;                                                                         \
        for (uint32_t j = rp[i]; j < rp[i + 1]; ++j)                                        \
            if (ci[j] == i) {                                                               \
                d = (T)1 / v[j];                                                            \
                break;                                                                      \
            }                                                                               \
        invD[i] = precond ? d : (T)1;                                                       \
    }                                                                                       \
    for (uint32_t i = 0; i < n; ++i) { /* L103-109: r = b - A x */                          \
        T acc = 0;                                                                          \
        for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) acc = ORC_FMA(v[j], x[ci[j]], acc);     \
        r[i] = b[i] - acc;                                                                  \
    }                                                                                       \
    double nr = 0;                                                                          \
    for (uint32_t i = 0; i < n; ++i) nr += fabs((double)r[i]);                              \
    double residuum = nr / norm_b;                                                          \
    res.initial_residual = residuum;                                                        \
    res.current_residual = residuum;                                                        \
    if (residuum <= relTol || residuum <= absTol * residuum || nIters_prior >= maxIters) {  \
        res.n_iters = 0;                                                                    \
        res.early_exit = 1;                                                                 \
        goto done;                                                                          \
    }                                                                                       \
    for (uint32_t i = 0; i < n; ++i) {                                                      \
        p[i] = r[i];                                                                        \
        rs[i] = r[i];                                                                       \
    }                                                                                       \
    double rr_old = 0;                                                                      \
    for (uint32_t i = 0; i < n; ++i) rr_old += (double)rs[i] * (double)r[i];                \
    int iteration = 0;                                                                      \
    int cur_n = nIters_prior;                                                               \
    for (;;) {                                                                              \
        for (uint32_t i = 0; i < n; ++i) Mp[i] = p[i] * invD[i];                            \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            T acc = 0;                                                                      \
            for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) acc = ORC_FMA(v[j], Mp[ci[j]], acc);\
            AMp[i] = acc;                                                                   \
        }                                                                                   \
        double ra = 0;                                                                      \
        for (uint32_t i = 0; i < n; ++i) ra += (double)rs[i] * (double)AMp[i];              \
        const T alpha = (T)(rr_old / ra);                                                   \
        double ns = 0;                                                                      \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            s[i] = r[i] - alpha * AMp[i];                                                   \
            ns += fabs((double)s[i]);                                                       \
        }                                                                                   \
        residuum = ns / norm_b;                                                             \
        res.current_residual = residuum;                                                    \
        if (residuum <= relTol || residuum <= absTol * res.initial_residual ||              \
            cur_n >= maxIters) { /* L186-195 */                                             \
            res.n_iters = iteration;                                                        \
            for (uint32_t i = 0; i < n; ++i) x[i] = x[i] + alpha * Mp[i];                   \
            res.executed = iteration + 1;                                                   \
            break;                                                                          \
        }                                                                                   \
        for (uint32_t i = 0; i < n; ++i) Ms[i] = s[i] * invD[i];                            \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            T acc = 0;                                                                      \
            for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) acc = ORC_FMA(v[j], Ms[ci[j]], acc);\
            AMs[i] = acc;                                                                   \
        }                                                                                   \
        double us = 0, ww = 0;                                                              \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            us += (double)AMs[i] * (double)s[i];                                            \
            ww += (double)AMs[i] * (double)AMs[i];                                          \
        }                                                                                   \
        const T omega = (T)(us / ww);                                                       \
        nr = 0;                                                                             \
        double rr_new = 0;                                                                  \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            x[i] = x[i] + alpha * Mp[i];                                                    \
            x[i] = x[i] + omega * Ms[i];                                                    \
            r[i] = s[i] - omega * AMs[i];                                                   \
            nr += fabs((double)r[i]);                                                       \
            rr_new += (double)rs[i] * (double)r[i];                                         \
        }                                                                                   \
        residuum = nr / norm_b;                                                             \
        res.current_residual = residuum;                                                    \
        if (residuum <= relTol || residuum <= absTol * res.initial_residual ||              \
            cur_n >= maxIters) { /* L233-238 */                                             \
            res.n_iters = iteration;                                                        \
            res.executed = iteration + 1;                                                   \
            break;                                                                          \
        }                                                                                   \
        T beta = (T)(rr_new / rr_old);                                                      \
        beta = beta * alpha;                                                                \
        beta = beta / omega;                                                                \
        rr_old = rr_new;                                                                    \
        const T bo = beta * omega;                                                          \
        for (uint32_t i = 0; i < n; ++i) {                                                  \
            p[i] = beta * p[i] + r[i];                                                      \
            p[i] = p[i] - bo * AMp[i];                                                      \
        }                                                                                   \
        iteration++;                                                                        \
        cur_n = iteration;                                                                  \
        res.n_iters = iteration;                                                            \
    }                                                                                       \
done:                                                                                       \
    free(invD);                                                                             \
    free(r);                                                                                \
    free(rs);                                                                               \
    free(p);                                                                                \
    free(s);                                                                                \
    free(Mp);                                                                               \
    free(AMp);                                                                              \
    free(Ms);                                                                               \
    free(AMs);                                                                              \
    return res;

#define ORC_MIN(T) _Generic((T)0, float: FLT_MIN, default: DBL_MIN)

orc_cg_result orc_bicgstab_d(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v,
                             const double* b, double* x, int precond, int maxIters, double relTol,
                             double absTol, int nIters_prior)
{
    ORC_BICG_BODY(double)
}

orc_cg_result orc_bicgstab_s(uint32_t n, const uint32_t* rp, const uint32_t* ci, const float* v,
                             const float* b, float* x, int precond, int maxIters, double relTol,
                             double absTol, int nIters_prior)
{
    ORC_BICG_BODY(float)
}

/* ---- csrmm -----------------------------------------------------------------------
 * C = alpha * A * B + beta * C with row-major B and C (element (k, c) at
 * k * ld + c), as csrmv_batched of src/library/kernels/csrmm_general.cl
 * L59-167 computes it column by column: alpha == 0 contributes nothing,
 * beta == 0 never reads C. Long-double sums; scale[r*ncols+c] receives
 * |alpha| sum |a b| + |beta c| for error bounds. */
void orc_csrmm_d(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* v, const double* B,
                 uint32_t ldb, double* C, uint32_t ldc, uint32_t ncols, double alpha, double beta,
                 double* scale)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (uint32_t r = 0; r < n; ++r) {
        for (uint32_t c = 0; c < ncols; ++c) {
            long double s = 0, a = 0;
            if (alpha != 0)
                for (uint32_t j = rp[r]; j < rp[r + 1]; ++j) {
                    const long double t = (long double)alpha * v[j] * B[(size_t)ci[j] * ldb + c];
                    s += t;
                    a += fabsl(t);
                }
            double* out = &C[(size_t)r * ldc + c];
            if (beta != 0) {
                s += (long double)beta * *out;
                a += fabsl((long double)beta * *out);
            }
            *out = (double)s;
            scale[(size_t)r * ncols + c] = (double)a;
        }
    }
}
