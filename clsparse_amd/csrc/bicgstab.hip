// BiCGStab on gfx950 with a device-resident iteration loop.
//
// Algorithm and stopping rule follow the reference
// (src/library/solvers/biconjugate-gradients-stabilized.hpp L41-268,
//  src/library/solvers/solver-control.hpp L47-64): right preconditioning with
// M = Jacobi or identity, residual ||.||_1 / ||b||_1, r* = r0, and two
// finished() checks per iteration:
//   Mp = M p; AMp = A Mp; alpha = <r*, r> / <r*, AMp>; s = r - alpha AMp;
//   if finished(||s||/||b||): x += alpha Mp, stop;
//   Ms = M s; AMs = A Ms; omega = <AMs, s> / <AMs, AMs>;
//   x += alpha Mp; x += omega Ms; r = s - omega AMs;
//   if finished(||r||/||b||): stop;
//   beta = (<r*, r_new> / <r*, r>) * alpha / omega;
//   p = beta p + r; p = p - (beta omega) AMp.
// finished() reads the solver control's nIters, which keeps its value from
// before the call until the first iteration ends (then = iteration index).
// One iteration is five kernels for a stencil matrix in diagonal storage (two
// SpMVs with their dot products fused, the s update with M s, the first check
// with the x / r update, the p update with M p and the second check), seven
// with the CSR SpMV (two dot kernels more); the scalars stay on the device;
// 32 iterations are captured once in a hipGraph and replayed, and kernels
// return at once after the stop. Reductions are fixed-order (FP64 partials),
// so solves are reproducible. The reference's function ends without a return
// statement; here it returns clsparseSuccess.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>

#include "common.hpp"
#include "solver_common.hpp"
#include "dia.hpp"

namespace {

constexpr int kBiChunk = 32;  // iterations per captured graph (even: both parities)
// Rows per thread per step of the grid-stride loops (their loads issued
// together; rows and partial sums still in grid-stride order).
constexpr int kBiRows = 4;

struct BiDevState
{
    double norm_b;
    double rel_tol;
    double abs_tol;
    double init_res;
    double cur_res;
    double alpha_stop;  // alpha of the iteration that stopped at the first check
    int max_iters;
    int n_prior;        // the solver control's nIters before the call
    int iter;           // iterations completed
    int done;
    int stop_kind;      // 1: at the first check (x += alpha Mp pending), 2: at the second
    int n_iters;
    int converged;
    int precond;
};

__device__ __forceinline__ bool bi_finished(const BiDevState* st, double res, int it)
{
    const int n = (it == 0) ? st->n_prior : it;
    return res <= st->rel_tol || res <= st->abs_tol * st->init_res || n >= st->max_iters;
}

template <typename T>
__device__ __forceinline__ T precond_apply(const BiDevState* st, const T* invD, T v, size_t i)
{
    return st->precond ? v * invD[i] : v;
}

// r = b - y; r* = r; p = r; Mp = M p; partials ||r||_1 and <r*, r>.
template <typename T>
__global__ __launch_bounds__(256) void bi_init(size_t n, const T* __restrict__ b, const T* __restrict__ y,
                                               T* __restrict__ r, T* __restrict__ rs, T* __restrict__ p,
                                               T* __restrict__ Mp, const T* __restrict__ invD,
                                               const BiDevState* __restrict__ st,
                                               double* __restrict__ part_n, double* __restrict__ part_rr)
{
    __shared__ double sh[4];
    double nr = 0.0, rr = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
        const T ri = b[i] - y[i];
        r[i] = ri;
        rs[i] = ri;
        p[i] = ri;
        Mp[i] = precond_apply(st, invD, ri, i);
        nr += fabs(static_cast<double>(ri));
        rr += static_cast<double>(ri) * static_cast<double>(ri);
    }
    nr = block_sum256(nr, sh);
    rr = block_sum256(rr, sh);
    if (threadIdx.x == 0) {
        part_n[blockIdx.x] = nr;
        part_rr[blockIdx.x] = rr;
    }
}

__global__ __launch_bounds__(256) void bi_init_finalize(BiDevState* __restrict__ st,
                                                        const double* __restrict__ part_n, uint32_t g)
{
    __shared__ double sh[4];
    const double s = sum_partials(part_n, g, sh);
    if (threadIdx.x == 0) {
        st->init_res = s / st->norm_b;
        st->cur_res = st->init_res;
    }
}

// partials of <a, b> (and, with c, of <c, d>).
template <typename T>
__global__ __launch_bounds__(256) void bi_dot(size_t n, const T* __restrict__ a, const T* __restrict__ b,
                                              double* __restrict__ part_ab, const T* __restrict__ c,
                                              const T* __restrict__ d, double* __restrict__ part_cd,
                                              const BiDevState* __restrict__ st)
{
    __shared__ double sh[4];
    if (st->done) return;
    double s1 = 0.0, s2 = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i0 = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i0 < n; i0 += kBiRows * stride) {
        T av[kBiRows], bv[kBiRows], cv[kBiRows], dv[kBiRows];
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i < n) {
                av[u] = a[i];
                bv[u] = b[i];
                if (c) {
                    cv[u] = c[i];
                    dv[u] = d[i];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            if (i0 + u * stride >= n) continue;
            s1 += static_cast<double>(av[u]) * static_cast<double>(bv[u]);
            if (c) s2 += static_cast<double>(cv[u]) * static_cast<double>(dv[u]);
        }
    }
    s1 = block_sum256(s1, sh);
    if (c) s2 = block_sum256(s2, sh);
    if (threadIdx.x == 0) {
        part_ab[blockIdx.x] = s1;
        if (c) part_cd[blockIdx.x] = s2;
    }
}

// alpha = <r*, r> / <r*, AMp>; s = r - alpha AMp; Ms = M s; partials ||s||_1.
template <typename T>
__global__ __launch_bounds__(256) void bi_s_update(size_t n, const T* __restrict__ r, const T* __restrict__ AMp,
                                                   T* __restrict__ s, T* __restrict__ Ms,
                                                   const T* __restrict__ invD,
                                                   const double* __restrict__ part_rr_old,
                                                   const double* __restrict__ part_ra,
                                                   double* __restrict__ part_s, uint32_t g,
                                                   const BiDevState* __restrict__ st)
{
    __shared__ double sh[4];
    if (st->done) return;
    const T alpha = static_cast<T>(sum_partials(part_rr_old, g, sh) / sum_partials(part_ra, g, sh));
    double ns = 0.0;
    const bool pre = st->precond != 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i0 = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i0 < n; i0 += kBiRows * stride) {
        T rv[kBiRows], av[kBiRows], dv[kBiRows];
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i < n) {
                rv[u] = r[i];
                av[u] = AMp[i];
                if (pre) dv[u] = invD[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i >= n) continue;
            const T si = rv[u] - alpha * av[u];
            s[i] = si;
            Ms[i] = pre ? si * dv[u] : si;
            ns += fabs(static_cast<double>(si));
        }
    }
    ns = block_sum256(ns, sh);
    if (threadIdx.x == 0) part_s[blockIdx.x] = ns;
}

// First check, then omega = <AMs, s> / <AMs, AMs>; x += alpha Mp;
// x += omega Ms; r = s - omega AMs; partials ||r||_1 and <r*, r>.
// Every workgroup evaluates the first check, finished(||s||_1 / ||b||_1),
// from the same partials and so reaches the same decision; on a stop,
// workgroup 0 records it (x += alpha Mp follows after the loop) and none
// updates x or r. The A Ms product and its dots, launched before this
// kernel, are then unused.
template <typename T>
__global__ __launch_bounds__(256) void bi_xr_update(
    size_t n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ rs, const T* __restrict__ s,
    const T* __restrict__ Mp, const T* __restrict__ Ms, const T* __restrict__ AMs,
    const double* __restrict__ part_rr_old, const double* __restrict__ part_ra,
    const double* __restrict__ part_us, const double* __restrict__ part_ww,
    const double* __restrict__ part_s, double* __restrict__ part_n, double* __restrict__ part_rr_new,
    uint32_t g, BiDevState* __restrict__ st)
{
    __shared__ double sh[4];
    if (st->done) return;
    const double rr_old = sum_partials(part_rr_old, g, sh);
    const double ra = sum_partials(part_ra, g, sh);
    const double res_s = sum_partials(part_s, g, sh) / st->norm_b;
    const int it = st->iter;
    if (bi_finished(st, res_s, it)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->cur_res = res_s;
            st->done = 1;
            st->stop_kind = 1;
            st->n_iters = it;
            st->alpha_stop = rr_old / ra;
            st->converged = (res_s <= st->rel_tol || res_s <= st->abs_tol * st->init_res) ? 1 : 0;
        }
        return;
    }
    const T alpha = static_cast<T>(rr_old / ra);
    const T omega = static_cast<T>(sum_partials(part_us, g, sh) / sum_partials(part_ww, g, sh));
    double nr = 0.0, rr = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i0 = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i0 < n; i0 += kBiRows * stride) {
        T xv[kBiRows], mp[kBiRows], ms[kBiRows], sv[kBiRows], ams[kBiRows], rsv[kBiRows];
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i < n) {
                xv[u] = x[i];
                mp[u] = Mp[i];
                ms[u] = Ms[i];
                sv[u] = s[i];
                ams[u] = AMs[i];
                rsv[u] = rs[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i >= n) continue;
            T xi = xv[u] + alpha * mp[u];
            xi = xi + omega * ms[u];
            x[i] = xi;
            const T ri = sv[u] - omega * ams[u];
            r[i] = ri;
            nr += fabs(static_cast<double>(ri));
            rr += static_cast<double>(rsv[u]) * static_cast<double>(ri);
        }
    }
    nr = block_sum256(nr, sh);
    rr = block_sum256(rr, sh);
    if (threadIdx.x == 0) {
        part_n[blockIdx.x] = nr;
        part_rr_new[blockIdx.x] = rr;
    }
}

// beta = (<r*, r_new> / <r*, r_old>) * alpha / omega; p = beta p + r;
// p = p - (beta omega) AMp; Mp = M p. Block 0 evaluates the second check.
template <typename T>
__global__ __launch_bounds__(256) void bi_p_update(
    size_t n, T* __restrict__ p, T* __restrict__ Mp, const T* __restrict__ r, const T* __restrict__ AMp,
    const T* __restrict__ invD, const double* __restrict__ part_rr_old,
    const double* __restrict__ part_rr_new, const double* __restrict__ part_ra,
    const double* __restrict__ part_us, const double* __restrict__ part_ww,
    const double* __restrict__ part_n, uint32_t g, BiDevState* __restrict__ st)
{
    __shared__ double sh[4];
    if (st->done) return;
    const double rr_old = sum_partials(part_rr_old, g, sh);
    const T alpha = static_cast<T>(rr_old / sum_partials(part_ra, g, sh));
    const T omega = static_cast<T>(sum_partials(part_us, g, sh) / sum_partials(part_ww, g, sh));
    T beta = static_cast<T>(sum_partials(part_rr_new, g, sh) / rr_old);
    beta = beta * alpha;
    beta = beta / omega;
    const T bo = beta * omega;
    const bool pre = st->precond != 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i0 = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i0 < n; i0 += kBiRows * stride) {
        T pv[kBiRows], rv[kBiRows], av[kBiRows], dv[kBiRows];
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i < n) {
                pv[u] = p[i];
                rv[u] = r[i];
                av[u] = AMp[i];
                if (pre) dv[u] = invD[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kBiRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i >= n) continue;
            T pi = beta * pv[u] + rv[u];
            pi = pi - bo * av[u];
            p[i] = pi;
            Mp[i] = pre ? pi * dv[u] : pi;
        }
    }
    if (blockIdx.x == 0) {
        const double res = sum_partials(part_n, g, sh) / st->norm_b;
        if (threadIdx.x == 0) {
            const int it = st->iter;
            st->cur_res = res;
            if (bi_finished(st, res, it)) {
                st->done = 1;
                st->stop_kind = 2;
                st->n_iters = it;
                st->converged = (res <= st->rel_tol || res <= st->abs_tol * st->init_res) ? 1 : 0;
            } else {
                st->iter = it + 1;
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void bi_xfix(size_t n, T* __restrict__ x, const T* __restrict__ Mp,
                                               const BiDevState* __restrict__ st)
{
    const T alpha = static_cast<T>(st->alpha_stop);
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
        x[i] = x[i] + alpha * Mp[i];
}

template <typename T>
struct BiBuffers
{
    T *r, *rs, *p, *Mp, *AMp, *s, *Ms, *AMs, *invD, *one, *zero;
    double *part_n, *part_ra, *part_s, *part_us, *part_ww;
    double* part_rr[2];
    BiDevState* st;

    static size_t bytes(size_t n)
    {
        return 9 * align_up(n * sizeof(T)) + 2 * align_up(sizeof(T)) +
               7 * align_up(kReducePartialsMax * sizeof(double)) + align_up(sizeof(BiDevState));
    }
    void carve(Carve& c, size_t n)
    {
        r = c.take<T>(n);
        rs = c.take<T>(n);
        p = c.take<T>(n);
        Mp = c.take<T>(n);
        AMp = c.take<T>(n);
        s = c.take<T>(n);
        Ms = c.take<T>(n);
        AMs = c.take<T>(n);
        invD = c.take<T>(n);
        one = c.take<T>(1);
        zero = c.take<T>(1);
        part_n = c.take<double>(kReducePartialsMax);
        part_ra = c.take<double>(kReducePartialsMax);
        part_s = c.take<double>(kReducePartialsMax);
        part_us = c.take<double>(kReducePartialsMax);
        part_ww = c.take<double>(kReducePartialsMax);
        part_rr[0] = c.take<double>(kReducePartialsMax);
        part_rr[1] = c.take<double>(kReducePartialsMax);
        st = c.take<BiDevState>(1);
    }
};

// adia: A in diagonal storage (stencils), else the CSR SpMV; its first SpMV
// forms the <r*, A Mp> partials itself (same grid-stride order as bi_dot).
template <typename T>
clsparseStatus enqueue_bi_iteration(const clsparseCsrMatrix* A, const DiaMatrix<T>& adia, T* x,
                                    BiBuffers<T>& B, size_t n, uint32_t g, int par, hipStream_t s, bool ext)
{
    const int* done = &B.st->done;
    double* rr_old = B.part_rr[par ^ 1];
    double* rr_new = B.part_rr[par];
    if (adia.vals != nullptr) {
        CLS_TRY(launch_dia_spmv<T>(adia, n, g, B.Mp, B.AMp, B.rs, B.part_ra, done, s));
    } else {
        CLS_TRY(launch_csrmv<T>(B.one, A, B.Mp, B.zero, B.AMp, s, ext, done));
        hipLaunchKernelGGL((bi_dot<T>), dim3(g), dim3(256), 0, s, n, B.rs, B.AMp, B.part_ra,
                           static_cast<const T*>(nullptr), static_cast<const T*>(nullptr), nullptr, B.st);
    }
    hipLaunchKernelGGL((bi_s_update<T>), dim3(g), dim3(256), 0, s, n, B.r, B.AMp, B.s, B.Ms, B.invD, rr_old,
                       B.part_ra, B.part_s, g, B.st);
    if (adia.vals != nullptr) {
        // <A Ms, s> and <A Ms, A Ms> partials fused (bi_dot's order and values).
        CLS_TRY(launch_dia_spmv<T>(adia, n, g, B.Ms, B.AMs, B.s, B.part_us, done, s, B.part_ww));
    } else {
        CLS_TRY(launch_csrmv<T>(B.one, A, B.Ms, B.zero, B.AMs, s, ext, done));
        hipLaunchKernelGGL((bi_dot<T>), dim3(g), dim3(256), 0, s, n, B.AMs, B.s, B.part_us,
                           static_cast<const T*>(B.AMs), static_cast<const T*>(B.AMs), B.part_ww, B.st);
    }
    hipLaunchKernelGGL((bi_xr_update<T>), dim3(g), dim3(256), 0, s, n, x, B.r, B.rs, B.s, B.Mp, B.Ms, B.AMs,
                       rr_old, B.part_ra, B.part_us, B.part_ww, B.part_s, B.part_n, rr_new, g, B.st);
    hipLaunchKernelGGL((bi_p_update<T>), dim3(g), dim3(256), 0, s, n, B.p, B.Mp, B.r, B.AMp, B.invD, rr_old,
                       rr_new, B.part_ra, B.part_us, B.part_ww, B.part_n, g, B.st);
    CLS_HIP_TRY(hipGetLastError());
    return clsparseSuccess;
}

template <typename T>
clsparseStatus bicg_solve(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                          clSParseSolverControl sc, clsparseControl control)
{
    // Size rule of biconjugate-gradients-stabilized.hpp L50-55; a square A as
    // in the CG solver (the temporaries have num_cols entries).
    if (A->num_cols != b->num_values || A->num_rows != x->num_values || A->num_rows != A->num_cols)
        return clsparseInvalidSystemSize;
    const size_t n = A->num_cols;
    hipStream_t s = control->solver_stream;
    T* xp = dev_ptr<T>(x->values, x->off_values);
    const T* bp = dev_ptr<T>(b->values, b->off_values);
    if (n > 0 && (xp == nullptr || bp == nullptr)) return clsparseInvalidMemObject;

    // Stencil matrices: A's values in diagonal storage for the iterations
    // (not with extended precision, which needs the compensated CSR sums).
    DiaOffsets a_off{};
    int a_ndiag = 0;
    const bool ext = control->extended_precision;
    // (after the caller's queued work: A's arrays and the partials buffer)
    CLS_TRY(enter_solver_stream(control));
    if (!ext) CLS_TRY(dia_candidates(A, reinterpret_cast<int32_t*>(control->red_partials), s, &a_off, &a_ndiag));
    void* work = nullptr;
    CLS_TRY(control_workspace(control, BiBuffers<T>::bytes(n) +
                                           (a_ndiag ? align_up(a_ndiag * n * sizeof(T)) + align_up(sizeof(int)) : 0),
                              &work));
    void* pinned = nullptr;
    CLS_TRY(control_pinned(control, sizeof(BiDevState) + 64, &pinned));
    BiDevState* hs = static_cast<BiDevState*>(pinned);
    Carve c{static_cast<char*>(work)};
    BiBuffers<T> B;
    B.carve(c, n);
    T* a_vals = a_ndiag ? c.take<T>(a_ndiag * n) : nullptr;
    int* a_bad = a_ndiag ? c.take<int>(1) : nullptr;
    const uint32_t g = grid_for(n);

    CLS_TRY(enter_solver_stream(control));
    const auto t0 = std::chrono::steady_clock::now();
    // ||b||_1 (L66-73)
    CLS_TRY(launch_reduce<T, double>(RedOp::Abs, bp, nullptr, n, &B.st->norm_b, false,
                                     control->red_partials, s));
    CLS_HIP_TRY(hipMemcpyAsync(hs, B.st, sizeof(BiDevState), hipMemcpyDeviceToHost, s));
    CLS_HIP_TRY(hipStreamSynchronize(s));
    const double norm_b = hs->norm_b;
    sc->stats = clsparseAmdSolverStats{};
    if (norm_b <= static_cast<double>(std::numeric_limits<T>::min())) {
        // x = b and zero tolerances (L75-83)
        sc->nIters = 0;
        sc->absoluteTolerance = 0.0;
        sc->relativeTolerance = 0.0;
        if (n > 0) CLS_HIP_TRY(hipMemcpyAsync(xp, bp, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        CLS_TRY(leave_solver_stream(control));
        return clsparseSuccess;
    }

    const int prec = (sc->preconditioner == DIAGONAL) ? 1 : 0;
    if (prec) {
        hipLaunchKernelGGL((extract_inv_diag<T, T>), dim3((A->num_rows + 255) / 256), dim3(256), 0, s,
                           A->num_rows, dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer),
                           dev_ptr<uint32_t>(A->col_indices, A->off_col_indices),
                           dev_ptr<T>(A->values, A->off_values), B.invD);
    }
    hipLaunchKernelGGL((fill_kernel<T>), dim3(1), dim3(256), 0, s, size_t(1), B.one, T(1));
    hipLaunchKernelGGL((fill_kernel<T>), dim3(1), dim3(256), 0, s, size_t(1), B.zero, T(0));
    BiDevState init{};
    init.norm_b = norm_b;
    init.rel_tol = sc->relativeTolerance;
    init.abs_tol = sc->absoluteTolerance;
    init.max_iters = sc->maxIters;
    init.n_prior = sc->nIters;
    init.precond = prec;
    *hs = init;
    CLS_HIP_TRY(hipMemcpyAsync(B.st, hs, sizeof(BiDevState), hipMemcpyHostToDevice, s));

    DiaMatrix<T> adia{};
    if (a_ndiag) {
        CLS_TRY((dia_fill<T, T>(A, n, &a_ndiag, a_off, a_vals, a_bad, s, &adia.sym)));
        if (a_ndiag) adia = DiaMatrix<T>{a_vals, a_ndiag, adia.sym, a_off};
    }

    // y = A x (into AMp), r = b - y, r* = r, p = r, Mp = M p (L103-134)
    CLS_TRY(launch_csrmv<T>(B.one, A, xp, B.zero, B.AMp, s, ext, nullptr));
    hipLaunchKernelGGL((bi_init<T>), dim3(g), dim3(256), 0, s, n, bp, B.AMp, B.r, B.rs, B.p, B.Mp, B.invD,
                       B.st, B.part_n, B.part_rr[1]);
    hipLaunchKernelGGL(bi_init_finalize, dim3(1), dim3(256), 0, s, B.st, B.part_n, g);
    CLS_HIP_TRY(hipGetLastError());
    CLS_HIP_TRY(hipMemcpyAsync(hs, B.st, sizeof(BiDevState), hipMemcpyDeviceToHost, s));
    CLS_HIP_TRY(hipStreamSynchronize(s));
    sc->initialResidual = hs->init_res;
    if (sc->finished(sc->initialResidual)) {  // L124-128 (uses the stored nIters)
        sc->nIters = 0;
        sc->stats.l1_residual = sc->initialResidual;
        CLS_TRY(leave_solver_stream(control));
        return clsparseSuccess;
    }

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const bool use_graph = graphs_enabled();
    clsparseStatus status = clsparseSuccess;
    if (use_graph) {
        CLS_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kBiChunk && status == clsparseSuccess; ++k)
            status = enqueue_bi_iteration<T>(A, adia, xp, B, n, g, k & 1, s, ext);
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (status == clsparseSuccess && e != hipSuccess) status = clsparseInvalidKernelExecution;
        if (status == clsparseSuccess && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            status = clsparseInvalidKernelExecution;
    }
    int printed = 0;
    while (status == clsparseSuccess) {
        if (use_graph) {
            if (hipGraphLaunch(exec, s) != hipSuccess) status = clsparseInvalidKernelExecution;
        } else {
            for (int k = 0; k < kBiChunk && status == clsparseSuccess; ++k)
                status = enqueue_bi_iteration<T>(A, adia, xp, B, n, g, k & 1, s, ext);
        }
        if (status != clsparseSuccess) break;
        CLS_HIP_TRY(hipMemcpyAsync(hs, B.st, sizeof(BiDevState), hipMemcpyDeviceToHost, s));
        CLS_HIP_TRY(hipStreamSynchronize(s));
        if (sc->printMode == VERBOSE) {
            // the reference prints after every completed iteration (L262-265)
            for (int k = printed; k < hs->iter; ++k) {
                sc->nIters = k + 1;
                sc->currentResidual = hs->cur_res;
                sc->print();
            }
            printed = hs->iter;
        }
        if (hs->done) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    CLS_TRY(status);
    if (hs->stop_kind == 1) {
        // x = x + alpha Mp at the first check (L191-193)
        hipLaunchKernelGGL((bi_xfix<T>), dim3(ew_grid(n)), dim3(256), 0, s, n, xp, B.Mp, B.st);
        CLS_HIP_TRY(hipGetLastError());
    }
    CLS_TRY(leave_solver_stream(control));
    const auto t1 = std::chrono::steady_clock::now();
    sc->nIters = hs->n_iters;
    sc->currentResidual = hs->cur_res;
    sc->stats.iterations = hs->n_iters + 1;
    sc->stats.l1_residual = hs->cur_res;
    sc->stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    sc->stats.status_detail = hs->converged ? 0 : 1;
    sc->stats.preconditioner = prec;
    sc->stats.a_diagonals = a_ndiag;
    return clsparseSuccess;
}

template <typename T>
clsparseStatus bicg_entry(cldenseVector* x, const clsparseCsrMatrix* A, const cldenseVector* b,
                          clSParseSolverControl solverControl, clsparseControl control)
{
    // Check order of biconjugate-gradients-stabilized.cpp L26-41.
    if (!clsparseInitialized) return clsparseNotInitialized;
    if (control == nullptr) return clsparseInvalidControlObject;
    if (solverControl == nullptr) return clsparseInvalidSolverControlObject;
    if (x == nullptr || A == nullptr || b == nullptr) return clsparseInvalidValue;
    const clsparseStatus status = bicg_solve<T>(x, A, b, solverControl, control);
    solverControl->printSummary(status);
    return status;
}

}  // namespace

// src/library/solvers/biconjugate-gradients-stabilized.cpp L20-67
extern "C" clsparseStatus clsparseScsrbicgStab(cldenseVector* x, const clsparseCsrMatrix* A,
                                               const cldenseVector* b, clSParseSolverControl solverControl,
                                               clsparseControl control)
{
    return bicg_entry<float>(x, A, b, solverControl, control);
}

// src/library/solvers/biconjugate-gradients-stabilized.cpp L69-116
extern "C" clsparseStatus clsparseDcsrbicgStab(cldenseVector* x, const clsparseCsrMatrix* A,
                                               const cldenseVector* b, clSParseSolverControl solverControl,
                                               clsparseControl control)
{
    return bicg_entry<double>(x, A, b, solverControl, control);
}
