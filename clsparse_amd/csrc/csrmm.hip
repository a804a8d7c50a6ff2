// Sparse times dense: C = alpha * A * B + beta * C (clsparse[S/D]csrmm).
//
// Semantics of the reference (src/library/blas3/clsparse-csrmm.cpp L25-76,
// clsparse-csrmm.hpp L84-179, kernels/csrmm_general.cl L59-167): B and C are
// row-major with lead_dim as the row stride (the major field is not read),
// element (k, c) of B at off_values + k * lead_dim + c; alpha == 0 gives
// beta * C without reading B, beta == 0 never reads C, beta == 1 adds C.
// The reference runs one CSR-Vector SpMV per column of C (A re-read for every
// column). Here W lanes (the column tile, a power of two >= the columns of C up
// to 64) form a row slot, and a wave runs 64 / W rows of C at once: each slot
// loads W nonzeros of its row with one coalesced read, hands each nonzero's
// value and column to its lanes through lane permutes, and keeps 8 B row
// segments in flight per step, every one a coalesced read of W columns. Waves
// walk consecutive rows and workgroups are remapped so each XCD covers one
// contiguous range of rows (neighbouring rows share B rows in its L2). Each
// element of C is summed over the row's nonzeros in storage order, so results
// are deterministic.
#include <algorithm>

#include "common.hpp"

namespace {

constexpr int kBInFlight = 8;

template <typename T, int W>
__global__ __launch_bounds__(256) void csrmm_kernel(uint32_t num_rows, uint32_t ncols,
                                                    const uint32_t* __restrict__ rp,
                                                    const uint32_t* __restrict__ ci,
                                                    const T* __restrict__ v, const T* __restrict__ B,
                                                    uint32_t ldb, T* __restrict__ C, uint32_t ldc,
                                                    const T* __restrict__ pAlpha,
                                                    const T* __restrict__ pBeta)
{
    constexpr uint32_t kSlots = 64 / W;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t slot = lane / W;
    const uint32_t t = lane % W;
    const int base = static_cast<int>(slot * W);  // first lane of this slot
    const T alpha = *pAlpha;
    const T beta = *pBeta;
    // Each XCD (gridDim.x is a multiple of 8; logical workgroups [x G/8,
    // (x+1) G/8) run on XCD x) owns one contiguous eighth of the rows and its
    // waves sweep it in steps of waves * kSlots rows: a B row shared by
    // nearby rows is then fetched into one L2, not one per XCD.
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t per_xcd = gridDim.x / 8;
    const uint32_t xcd = lb / per_xcd;
    const uint32_t rows_per_xcd = (num_rows + 7) / 8;
    const uint32_t first = min(num_rows, xcd * rows_per_xcd);
    const uint32_t stop = min(num_rows, first + rows_per_xcd);
    const uint32_t wave = (lb - xcd * per_xcd) * 4 + (threadIdx.x >> 6);
    const uint32_t waves = per_xcd * 4;
    // The next step's row bounds are loaded at the top of each step, so their
    // latency hides behind this step's nonzeros and B rows.
    uint32_t row0 = first + wave * kSlots;
    uint32_t nbeg = 0, nend = 0;
    if (row0 + slot < stop) {
        nbeg = rp[row0 + slot];
        nend = rp[row0 + slot + 1];
    }
    for (; row0 < stop; row0 += waves * kSlots) {
        const uint32_t row = row0 + slot;
        const bool live = row < stop;
        const uint32_t beg = nbeg, end = nend;
        const uint32_t next = row0 + waves * kSlots + slot;
        if (next < stop) {
            nbeg = rp[next];
            nend = rp[next + 1];
        }
        for (uint32_t c0 = 0; c0 < ncols; c0 += W) {
            const uint32_t c = c0 + t;
            const bool colok = c < ncols;
            T sum = T(0);
            if (alpha != T(0)) {
                for (uint32_t j0 = beg; j0 < end; j0 += W) {
                    // Lane t of the slot holds nonzero j0 + t of the row.
                    const uint32_t jj = j0 + t;
                    T av = T(0);
                    uint32_t cc = 0;
                    if (jj < end) {
                        av = alpha * v[jj];
                        cc = ci[jj];
                    }
                    const uint32_t cnt = min(static_cast<uint32_t>(W), end - j0);
                    for (uint32_t q = 0; q < cnt; q += kBInFlight) {
                        T a8[kBInFlight], b8[kBInFlight];
#pragma unroll
                        for (int u = 0; u < kBInFlight; ++u) {
                            const int src = base + static_cast<int>(min(q + u, static_cast<uint32_t>(W - 1)));
                            a8[u] = __shfl(av, src, 64);
                            const uint32_t col = __shfl(cc, src, 64);
                            b8[u] = (colok && q + u < cnt) ? B[static_cast<size_t>(col) * ldb + c] : T(0);
                        }
#pragma unroll
                        for (int u = 0; u < kBInFlight; ++u)
                            if (q + u < cnt) sum = fma(a8[u], b8[u], sum);
                    }
                }
            }
            if (live && colok) {
                T* out = C + static_cast<size_t>(row) * ldc + c;
                if (beta == T(0))
                    *out = sum;
                else if (beta == T(1))
                    *out = sum + *out;
                else
                    *out = sum + beta * *out;
            }
        }
    }
}

template <typename T>
clsparseStatus csrmm_entry(const clsparseScalar* alpha, const clsparseCsrMatrix* A,
                           const cldenseMatrix* B, const clsparseScalar* beta, cldenseMatrix* C,
                           clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (alpha == nullptr || A == nullptr || B == nullptr || beta == nullptr || C == nullptr)
        return clsparseInvalidValue;
    // Shapes the reference reads without checking: B has A's columns as rows,
    // C has A's rows, and the leading dimensions cover the used columns.
    if (B->num_rows != A->num_cols || C->num_rows != A->num_rows || C->num_cols > B->num_cols ||
        B->lead_dim < C->num_cols || C->lead_dim < C->num_cols)
        return clsparseInvalidSize;
    const uint32_t n = A->num_rows, ncols = C->num_cols;
    if (n == 0 || ncols == 0) return finish_op(control);
    const T* pa = dev_ptr<T>(alpha->value, alpha->off_value);
    const T* pb = dev_ptr<T>(beta->value, beta->off_value);
    const uint32_t* rp = dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer);
    const uint32_t* ci = dev_ptr<uint32_t>(A->col_indices, A->off_col_indices);
    const T* v = dev_ptr<T>(A->values, A->off_values);
    const T* Bp = dev_ptr<T>(B->values, B->off_values);
    T* Cp = dev_ptr<T>(C->values, C->off_values);
    if (rp == nullptr || Bp == nullptr || Cp == nullptr || pa == nullptr || pb == nullptr)
        return clsparseInvalidMemObject;
    // At most as many waves as the chip holds at once (2048 workgroups of 4
    // waves at 8 waves per SIMD): the resident waves of an XCD then work on one
    // compact window of its rows that advances together, so the B rows that
    // neighbouring rows share stay in its L2.
    const uint32_t slots = ncols > 32 ? 1u : ncols > 16 ? 2u : ncols > 8 ? 4u : ncols > 4 ? 8u
                                       : ncols > 2 ? 16u : ncols > 1 ? 32u : 64u;
    const uint32_t want = (n + 4 * slots - 1) / (4 * slots);
    const uint32_t grid = std::min<uint32_t>(2048, (want + 7) / 8 * 8);  // a multiple of 8 (one share per XCD)
    hipStream_t s = control->stream;
#define CLS_CSRMM(W)                                                                                 \
    hipLaunchKernelGGL((csrmm_kernel<T, W>), dim3(grid), dim3(256), 0, s, n, ncols, rp, ci, v, Bp,   \
                       B->lead_dim, Cp, C->lead_dim, pa, pb)
    if (ncols > 32)
        CLS_CSRMM(64);
    else if (ncols > 16)
        CLS_CSRMM(32);
    else if (ncols > 8)
        CLS_CSRMM(16);
    else if (ncols > 4)
        CLS_CSRMM(8);
    else if (ncols > 2)
        CLS_CSRMM(4);
    else if (ncols > 1)
        CLS_CSRMM(2);
    else
        CLS_CSRMM(1);
#undef CLS_CSRMM
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}

}  // namespace

// src/library/blas3/clsparse-csrmm.cpp L25-49
extern "C" clsparseStatus clsparseScsrmm(const clsparseScalar* alpha, const clsparseCsrMatrix* sparseMatA,
                                         const cldenseMatrix* denseMatB, const clsparseScalar* beta,
                                         cldenseMatrix* denseMatC, const clsparseControl control)
{
    return csrmm_entry<float>(alpha, sparseMatA, denseMatB, beta, denseMatC, control);
}

// src/library/blas3/clsparse-csrmm.cpp L51-76
extern "C" clsparseStatus clsparseDcsrmm(const clsparseScalar* alpha, const clsparseCsrMatrix* sparseMatA,
                                         const cldenseMatrix* denseMatB, const clsparseScalar* beta,
                                         cldenseMatrix* denseMatC, const clsparseControl control)
{
    return csrmm_entry<double>(alpha, sparseMatA, denseMatB, beta, denseMatC, control);
}
