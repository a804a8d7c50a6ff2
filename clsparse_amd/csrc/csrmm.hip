// Sparse times dense: C = alpha * A * B + beta * C (clsparse[S/D]csrmm).
//
// Semantics of the reference (src/library/blas3/clsparse-csrmm.cpp L25-76,
// clsparse-csrmm.hpp L84-179, kernels/csrmm_general.cl L59-167): B and C are
// row-major with lead_dim as the row stride (the major field is not read),
// element (k, c) of B at off_values + k * lead_dim + c; alpha == 0 gives
// beta * C without reading B, beta == 0 never reads C, beta == 1 adds C.
// The reference runs one CSR-Vector SpMV per column of C (A re-read for every
// column). Here a wave owns a row of C: it walks the row's nonzeros once and
// its lanes cover up to 64 columns of B at a time, so every B row segment is
// one coalesced read; with fewer columns the spare lanes take every
// (64/width)-th nonzero and the partial sums meet through lane permutes.
// Per element of C the sum is fixed-order, so results are deterministic.
#include <algorithm>

#include "common.hpp"

namespace {

// W = lanes per column group (power of two, 1..64); 64 / W nonzero slices.
template <typename T, int W>
__global__ __launch_bounds__(256) void csrmm_kernel(uint32_t num_rows, uint32_t ncols,
                                                    const uint32_t* __restrict__ rp,
                                                    const uint32_t* __restrict__ ci,
                                                    const T* __restrict__ v, const T* __restrict__ B,
                                                    uint32_t ldb, T* __restrict__ C, uint32_t ldc,
                                                    const T* __restrict__ pAlpha,
                                                    const T* __restrict__ pBeta)
{
    constexpr int kSlices = 64 / W;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t c_in = lane % W;  // column inside the tile
    const uint32_t slice = lane / W;
    const T alpha = *pAlpha;
    const T beta = *pBeta;
    const uint32_t waves = gridDim.x * 4;
    for (uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6); row < num_rows; row += waves) {
        const uint32_t beg = rp[row], end = rp[row + 1];
        for (uint32_t c0 = 0; c0 < ncols; c0 += W) {
            const uint32_t c = c0 + c_in;
            T sum = T(0);
            if (alpha != T(0) && c < ncols) {
                for (uint32_t j = beg + slice; j < end; j += kSlices)
                    sum = fma(alpha * v[j], B[static_cast<size_t>(ci[j]) * ldb + c], sum);
            }
            // lanes with the same column and different slices: lane, lane+W, ...
            if constexpr (kSlices > 1) {
                for (int off = W; off < 64; off <<= 1) sum += __shfl_xor(sum, off, 64);
            }
            if (slice == 0 && c < ncols) {
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
    const uint32_t grid = std::min<uint32_t>((n + 3) / 4, 65536);
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
