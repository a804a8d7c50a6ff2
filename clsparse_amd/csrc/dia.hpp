// Stencil matrices in diagonal storage for the iterative solvers (CG,
// BiCGStab): detection from sampled rows, a checked fill from the caller's CSR
// values, the symmetric-half check, and the SpMV with an optional fused dot.
// Included by the solver sources after common.hpp and solver_common.hpp.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "common.hpp"
#include "line_part.hpp"
#include "solver_common.hpp"

namespace {

// Rows per thread per step (loads of all of them issued together; rows and
// partial sums still in grid-stride order).
constexpr int kDiaRows = 4;

// Diagonal storage of the FP32 copy of A in the mixed solvers (stencils of
// structured grids): at most kDiaMax diagonals.
constexpr int kDiaMax = 9;
static_assert(kDiaMax >= kFsaiMaxDiags, "offset table holds the factor's diagonals");

struct DiaOffsets
{
    int32_t o[kDiaMax];
};

template <typename T>
__device__ __forceinline__ T mul_rounded_t(T a, T b)
{
#pragma clang fp contract(off)
    return a * b;
}

// y = A p from diagonal storage (vals[d n + i] = A(i, i + off[d]), off
// ascending, zero outside the pattern) with partials of <y, p> in the
// grid-stride order of cg_dot_kernel (g workgroups). SYM: A is symmetric with
// offsets -o..o, and the upper diagonals are read from the lower ones,
// A(i, i + o) = A(i + o, i) = vals[d' n + i + o] with off[d'] = -o, so only
// (K + 1) / 2 diagonals are read. Workgroups are remapped so that each XCD
// walks contiguous row ranges (the shifted reads of p -- and of the mirrored
// diagonals -- then meet lines its own L2 holds).
template <int K, bool SYM, typename T>
__global__ __launch_bounds__(256) void dia_spmv_dot_kernel(size_t n, DiaOffsets off,
                                                           const T* __restrict__ vals,
                                                           const T* __restrict__ p,
                                                           T* __restrict__ y,
                                                           const T* __restrict__ w,
                                                           double* __restrict__ part,
                                                           double* __restrict__ part2,
                                                           const int* __restrict__ done)
{
    constexpr int KL = SYM ? (K + 1) / 2 : K;
    __shared__ double sh[4];
    if (done != nullptr && *done != 0) return;
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    double a = 0.0, a2 = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i0 = static_cast<size_t>(blk) * 256 + threadIdx.x; i0 < n; i0 += kDiaRows * stride) {
        T v[kDiaRows][K], xv[kDiaRows][K], pi[kDiaRows];
#pragma unroll
        for (int u = 0; u < kDiaRows; ++u) {
            const size_t i = i0 + u * stride;
#pragma unroll
            for (int d = 0; d < K; ++d) {
                const int64_t j = static_cast<int64_t>(i) + off.o[d];
                const bool in = i < n && j >= 0 && j < static_cast<int64_t>(n);
                const size_t at = d < KL ? static_cast<size_t>(d) * n + i
                                         : static_cast<size_t>(K - 1 - d) * n + static_cast<size_t>(j);
                v[u][d] = in ? vals[at] : T(0);
                xv[u][d] = in ? p[j] : T(0);
            }
            pi[u] = (i < n && part != nullptr) ? w[i] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kDiaRows; ++u) {
            const size_t i = i0 + u * stride;
            if (i >= n) continue;
            double sum = 0.0;  // FP32 products summed in FP64, as the CSR SpMV sums them
#pragma unroll
            for (int d = 0; d < K; ++d)
                if (v[u][d] != T(0)) sum += static_cast<double>(mul_rounded_t(v[u][d], xv[u][d]));
            const T acc = static_cast<T>(sum);
            y[i] = acc;
            a += static_cast<double>(acc) * static_cast<double>(pi[u]);
            a2 += static_cast<double>(acc) * static_cast<double>(acc);
        }
    }
    if (part == nullptr) return;
    a = block_sum256(a, sh);
    if (threadIdx.x == 0) part[blk] = a;
    if (part2 == nullptr) return;
    a2 = block_sum256(a2, sh);
    if (threadIdx.x == 0) part2[blk] = a2;
}

// R rows iu[0..R) of y = A x from diagonal storage through buffer loads,
// their <y, w> and <y, y> terms added to a and a2 (rows with live[u] false are
// computed from in-range loads and dropped: a branch around a row lets the
// compiler sink its loads into it, one row at a time).
template <int K, bool SYM, typename T, int R>
__device__ __forceinline__ void dia_rows(uint32_t n, const DiaOffsets& off, __amdgpu_buffer_rsrc_t rv,
                                         __amdgpu_buffer_rsrc_t rx, __amdgpu_buffer_rsrc_t rw, T* __restrict__ y,
                                         const uint32_t (&iu)[R], const bool (&live)[R], double& a, double& a2)
{
    constexpr int KL = SYM ? (K + 1) / 2 : K;
    constexpr uint32_t S = sizeof(T);
    T v[R][K], xv[R][K], pi[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const uint32_t i = iu[u];
#pragma unroll
        for (int d = 0; d < K; ++d) {
            const uint32_t j = i + static_cast<uint32_t>(off.o[d]);  // wraps below 0: out of range
            xv[u][d] = buf_load<T>(rx, j * S, 0);
            if (d < KL) {
                // Rows past n read row n - 1: the diagonal's base is a scalar
                // offset, which the buffer's range check may not cover.
                const uint32_t il = i < n ? i : n - 1;
                v[u][d] = buf_load<T>(rv, il * S, static_cast<uint32_t>(d) * n * S);
            } else {
                // Mirrored value at row j of the lower diagonal; an offset past
                // the buffer (no branch) reads 0 where j >= n.
                const uint32_t at = j < n ? (static_cast<uint32_t>(K - 1 - d) * n + j) * S : 0xFFFFFFFFu;
                v[u][d] = buf_load<T>(rv, at, 0);
            }
        }
        pi[u] = buf_load<T>(rw, i * S, 0);
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
        double sum = 0.0;  // FP32 products summed in FP64, as the CSR SpMV sums them
#pragma unroll
        for (int d = 0; d < K; ++d)
            if (v[u][d] != T(0)) sum += static_cast<double>(mul_rounded_t(v[u][d], xv[u][d]));
        const T acc = static_cast<T>(sum);
        if (live[u]) y[iu[u]] = acc;
        a += live[u] ? static_cast<double>(acc) * static_cast<double>(pi[u]) : 0.0;
        a2 += live[u] ? static_cast<double>(acc) * static_cast<double>(acc) : 0.0;
    }
}

// Two rows i and i + line of a symmetric stencil (offsets ascending, so
// diagonal 0 is -line and K - 1 is +line) in line order, with the far
// diagonals' operands carried in registers from row to row: pm = p[i - line],
// p0 = p[i] and f0 = the lower far value at row i come in, and the same three
// for row i + 2 line go out in their place. The values and the summation
// order are those of dia_rows; only the loads differ.
template <int K, typename T>
__device__ __forceinline__ void dia_line_pair(uint32_t n, const DiaOffsets& off, __amdgpu_buffer_rsrc_t rv,
                                              __amdgpu_buffer_rsrc_t rx, __amdgpu_buffer_rsrc_t rw,
                                              T* __restrict__ y, uint32_t i, uint32_t line, bool live0, bool live1,
                                              T& pm, T& p0, T& f0, double& a, double& a2)
{
    constexpr int KL = (K + 1) / 2;
    constexpr uint32_t S = sizeof(T);
    const uint32_t i1 = i + line, i2 = i1 + line;
    // Lower far values at rows i1 and i2 (0 past n, as the mirrored load reads).
    const T p1 = buf_load<T>(rx, i1 * S, 0), p2 = buf_load<T>(rx, i2 * S, 0);
    const T f1 = buf_load<T>(rv, i1 < n ? i1 * S : 0xFFFFFFFFu, 0);
    const T f2 = buf_load<T>(rv, i2 < n ? i2 * S : 0xFFFFFFFFu, 0);
    const uint32_t iu[2] = {i, i1};
    T v[2][K], xv[2][K], pi[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
#pragma unroll
        for (int d = 1; d < K - 1; ++d) {
            if (d != KL - 1) {  // p[i] itself is carried
                const uint32_t j = iu[u] + static_cast<uint32_t>(off.o[d]);  // wraps below 0: out of range
                xv[u][d] = buf_load<T>(rx, j * S, 0);
            }
            if (d < KL) {
                const uint32_t il = iu[u] < n ? iu[u] : n - 1;
                v[u][d] = buf_load<T>(rv, il * S, static_cast<uint32_t>(d) * n * S);
            } else {
                const uint32_t j = iu[u] + static_cast<uint32_t>(off.o[d]);
                const uint32_t at = j < n ? (static_cast<uint32_t>(K - 1 - d) * n + j) * S : 0xFFFFFFFFu;
                v[u][d] = buf_load<T>(rv, at, 0);
            }
        }
        pi[u] = buf_load<T>(rw, iu[u] * S, 0);
    }
    xv[0][0] = pm;
    v[0][0] = f0;
    xv[0][KL - 1] = p0;
    xv[0][K - 1] = p1;
    v[0][K - 1] = f1;
    xv[1][0] = p0;
    v[1][0] = f1;
    xv[1][KL - 1] = p1;
    xv[1][K - 1] = p2;
    v[1][K - 1] = f2;
    const bool live[2] = {live0, live1};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        double sum = 0.0;  // FP32 products summed in FP64, as the CSR SpMV sums them
#pragma unroll
        for (int d = 0; d < K; ++d)
            if (v[u][d] != T(0)) sum += static_cast<double>(mul_rounded_t(v[u][d], xv[u][d]));
        const T acc = static_cast<T>(sum);
        if (live[u]) y[iu[u]] = acc;
        a += live[u] ? static_cast<double>(acc) * static_cast<double>(pi[u]) : 0.0;
        a2 += live[u] ? static_cast<double>(acc) * static_cast<double>(acc) : 0.0;
    }
    pm = p1;
    p0 = p2;
    f0 = f2;
}

// The same product through buffer loads: the value of diagonal d at row i is
// the buffer offset d*n*sizeof(T) (one scalar register per diagonal) plus
// i*sizeof(T) (one vector register per row), where the 64-bit flat addresses
// took two vector registers per load and left two waves per SIMD; reads of p
// outside [0, n) return 0 from the buffer's range check. Needs
// K * n * sizeof(T) < 2^32.
// The grid (at most kReducePartialsMax workgroups of 4 waves) puts 4 waves on
// a SIMD: tell the scheduler so, or it trades the loads in flight for an
// occupancy the launch never reaches.
// line > 0: each lane walks rows c, c + line, c + 2 line, ... (line = the
// largest offset: a z line of a 3-D grid), a stretch of them per lane, so the
// rows +-line away and the mirrored values of the far diagonal are the ones
// the same lane loads one step before or after; lanes of a wave take
// consecutive c. line == 0: grid-stride rows.
template <int K, bool SYM, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void dia_spmv_dot_buf_kernel(uint32_t n, DiaOffsets off,
                                                               const T* __restrict__ vals,
                                                               const T* __restrict__ p,
                                                               T* __restrict__ y,
                                                               const T* __restrict__ w,
                                                               double* __restrict__ part,
                                                               double* __restrict__ part2,
                                                               const int* __restrict__ done,
                                                               uint32_t line)
{
    constexpr uint32_t S = sizeof(T);
    constexpr int R = S == 8 ? 2 : kDiaRows;  // FP64: fewer rows per step keep it under 128 VGPRs
    __shared__ double sh[4];
    if (done != nullptr && *done != 0) return;
    const __amdgpu_buffer_rsrc_t rv =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(vals), 0, K * n * S, 0x00020000);
    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p), 0, n * S, 0x00020000);
    // Dot operand (p itself in CG); no dot when part is null.
    const __amdgpu_buffer_rsrc_t rw =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(part != nullptr ? w : p), 0, part != nullptr ? n * S : 0u,
                                          0x00020000);
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    double a = 0.0, a2 = 0.0;
    if (line == 0) {
        const uint32_t stride = gridDim.x * 256;
        for (uint32_t i0 = blk * 256 + threadIdx.x; i0 < n; i0 += R * stride) {
            uint32_t iu[R];
            bool live[R];
#pragma unroll
            for (int u = 0; u < R; ++u) {
                iu[u] = i0 + u * stride;
                live[u] = iu[u] < n;
            }
            dia_rows<K, SYM, T, R>(n, off, rv, rx, rw, y, iu, live, a, a2);
        }
    } else {
        const uint32_t threads = gridDim.x * 256;
        const LinePart lp = line_part(n, line, threads);
        for (uint32_t q = blk * 256 + threadIdx.x; q < lp.items; q += threads) {
            uint32_t c, k_beg, k_end;
            line_item(lp, q, c, k_beg, k_end);
            if constexpr (SYM && K >= 3) {
                if (k_beg < k_end) {
                    const uint32_t i = c + k_beg * line;
                    T pm = buf_load<T>(rx, (i - line) * S, 0);  // wraps below 0: 0
                    T p0 = buf_load<T>(rx, i * S, 0);
                    T f0 = buf_load<T>(rv, i < n ? i * S : 0xFFFFFFFFu, 0);
                    for (uint32_t k0 = k_beg; k0 < k_end; k0 += 2) {
                        const uint32_t i0 = c + k0 * line;
                        dia_line_pair<K, T>(n, off, rv, rx, rw, y, i0, line, i0 < n,
                                            k0 + 1 < k_end && i0 + line < n, pm, p0, f0, a, a2);
                    }
                }
            } else {
                for (uint32_t k0 = k_beg; k0 < k_end; k0 += R) {
                    uint32_t iu[R];
                    bool live[R];
#pragma unroll
                    for (int u = 0; u < R; ++u) {
                        iu[u] = c + (k0 + u) * line;
                        live[u] = k0 + u < k_end && iu[u] < n;
                        if (!live[u]) iu[u] = c;  // a row in range, computed and dropped
                    }
                    dia_rows<K, SYM, T, R>(n, off, rv, rx, rw, y, iu, live, a, a2);
                }
            }
        }
    }
    if (part == nullptr) return;
    a = block_sum256(a, sh);
    if (threadIdx.x == 0) part[blk] = a;
    if (part2 == nullptr) return;
    a2 = block_sum256(a2, sh);
    if (threadIdx.x == 0) part2[blk] = a2;
}

// *bad = 1 unless A(i, i + o) == A(i + o, i) for every upper diagonal o of
// the full diagonal storage (offsets already checked to be -o..o).
template <typename T>
__global__ __launch_bounds__(256) void dia_sym_check_kernel(uint32_t n, int K, DiaOffsets off,
                                                            const T* __restrict__ vals, int* __restrict__ bad)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int d = (K + 1) / 2; d < K; ++d) {
        const uint64_t j = static_cast<uint64_t>(i) + static_cast<uint32_t>(off.o[d]);
        if (j >= n) continue;
        if (!(vals[static_cast<size_t>(d) * n + i] == vals[static_cast<size_t>(K - 1 - d) * n + j])) {
            *bad = 1;
            return;
        }
    }
}

// dia[d n + i] = (TO) A(i, i + off[d]) for every stored entry; *bad = 1 if
// a row has an entry off the K diagonals or columns not strictly increasing
// (then the CSR SpMV stays). dia must be zero beforehand.
template <typename TA, typename TO>
__global__ __launch_bounds__(256) void dia_fill_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                       const uint32_t* __restrict__ ci,
                                                       const TA* __restrict__ v, int K,
                                                       DiaOffsets off, TO* __restrict__ dia,
                                                       int* __restrict__ bad)
{
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    int64_t prev = INT64_MIN;
    for (uint32_t j = rp[row]; j < rp[row + 1]; ++j) {
        const int64_t o = static_cast<int64_t>(ci[j]) - static_cast<int64_t>(row);
        int d = -1;
        for (int k = 0; k < K; ++k)
            if (off.o[k] == o) d = k;
        if (o <= prev || d < 0) {
            *bad = 1;
            return;
        }
        prev = o;
        dia[static_cast<size_t>(d) * n + row] = static_cast<TO>(v[j]);
    }
}

// Column - row offsets of kDiaSamples rows spread over the matrix: out[16 k]
// = entry count of sample row k (16: more than 15 entries or an offset
// outside int32), then its offsets.
constexpr int kDiaSamples = 64;
__global__ __launch_bounds__(64) void dia_sample_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ ci,
                                                        int32_t* __restrict__ out)
{
    const uint32_t k = threadIdx.x;
    if (k >= static_cast<uint32_t>(kDiaSamples)) return;
    // Golden-ratio sequence: no aliasing with the row length of a grid (a
    // regular stride such as n / 64 lands every sample on the same column of
    // a 160 x 160 x 160 grid, a boundary column).
    const double t = (k + 0.5) * 0.6180339887498949;
    const uint32_t row = min(n - 1, static_cast<uint32_t>((t - floor(t)) * n));
    int32_t* o = out + 16 * k;
    const uint32_t b = rp[row], e = rp[row + 1];
    if (e < b || e - b > 15) {
        o[0] = 16;
        return;
    }
    o[0] = static_cast<int32_t>(e - b);
    for (uint32_t j = b; j < e; ++j) {
        const int64_t d = static_cast<int64_t>(ci[j]) - static_cast<int64_t>(row);
        if (d < INT32_MIN || d > INT32_MAX) {
            o[0] = 16;
            return;
        }
        o[1 + j - b] = static_cast<int32_t>(d);
    }
}

// The line length of the line-order kernels: the largest |offset| when it is
// at least 4096 (a 3-D grid's n_x n_y) and the matrix holds at least 8 lines;
// 0 (grid-stride rows) otherwise or when CLSPARSE_AMD_DIA_LINES=0.
inline uint32_t dia_line(const DiaOffsets& off, int k, size_t n)
{
    const char* e = std::getenv("CLSPARSE_AMD_DIA_LINES");
    if (e && e[0] == '0') return 0;
    int64_t far = 0;
    for (int d = 0; d < k; ++d) far = std::max<int64_t>(far, off.o[d] < 0 ? -int64_t(off.o[d]) : off.o[d]);
    return far >= 4096 && static_cast<uint64_t>(far) * 8 <= n ? static_cast<uint32_t>(far) : 0u;
}

template <int K, bool SYM, typename T>
void launch_dia_spmv_dot(size_t n, uint32_t g, const DiaOffsets& off, const T* vals, const T* p, T* y,
                         const T* w, double* part, double* part2, const int* done, hipStream_t s)
{
    if (dia_buf_fits(K, n, sizeof(T))) {
        // Line order when the largest offset is long (a 3-D grid's n_x n_y)
        // and the matrix holds several lines (CLSPARSE_AMD_DIA_LINES=0: off).
        const uint32_t line = dia_line(off, K, n);
        hipLaunchKernelGGL((dia_spmv_dot_buf_kernel<K, SYM, T>), dim3(g), dim3(256), 0, s,
                           static_cast<uint32_t>(n), off, vals, p, y, w, part, part2, done, line);
    }
    else
        hipLaunchKernelGGL((dia_spmv_dot_kernel<K, SYM, T>), dim3(g), dim3(256), 0, s, n, off, vals, p, y, w,
                           part, part2, done);
}

// A stencil matrix in diagonal storage for one solve (ndiag = 0: none).
template <typename T>
struct DiaMatrix
{
    const T* vals = nullptr;
    int ndiag = 0;
    bool sym = false;
    DiaOffsets off{};
};

// y = A x from diagonal storage on g workgroups; with part, also the
// partials of <y, w>, and with part2 those of <y, y> (grid-stride order, one
// per workgroup).
template <typename T>
clsparseStatus launch_dia_spmv(const DiaMatrix<T>& A, size_t n, uint32_t g, const T* x, T* y, const T* w,
                               double* part, const int* done, hipStream_t s, double* part2 = nullptr)
{
#define CLS_DIA_SPMV(K, S) launch_dia_spmv_dot<K, S, T>(n, g, A.off, A.vals, x, y, w, part, part2, done, s)
    if (A.sym) {
        switch (A.ndiag) {
        case 3: CLS_DIA_SPMV(3, true); break;
        case 5: CLS_DIA_SPMV(5, true); break;
        case 7: CLS_DIA_SPMV(7, true); break;
        case 9: CLS_DIA_SPMV(9, true); break;
        default: return clsparseInvalidValue;
        }
    } else {
        switch (A.ndiag) {
        case 1: CLS_DIA_SPMV(1, false); break;
        case 2: CLS_DIA_SPMV(2, false); break;
        case 3: CLS_DIA_SPMV(3, false); break;
        case 4: CLS_DIA_SPMV(4, false); break;
        case 5: CLS_DIA_SPMV(5, false); break;
        case 6: CLS_DIA_SPMV(6, false); break;
        case 7: CLS_DIA_SPMV(7, false); break;
        case 8: CLS_DIA_SPMV(8, false); break;
        case 9: CLS_DIA_SPMV(9, false); break;
        default: return clsparseInvalidValue;
        }
    }
#undef CLS_DIA_SPMV
    CLS_HIP_TRY(hipGetLastError());
    return clsparseSuccess;
}

bool dia_enabled()
{
    const char* e = std::getenv("CLSPARSE_AMD_DIA");
    return !(e && e[0] == '0');
}

// Union of the column - row offsets of kDiaSamples rows spread over A,
// ascending, when it has at most kDiaMax entries and that many diagonals of
// length n would store at most 1.5 nnz values; *ndiag = 0 otherwise. The
// fill kernel checks every row against the result.
// honour_env: CLSPARSE_AMD_DIA=0 turns the detection off (the solvers' A);
// the learned factor's storage has its own switch at solve time.
clsparseStatus dia_candidates(const clsparseCsrMatrix* A, int32_t* scratch, hipStream_t s,
                              DiaOffsets* off, int* ndiag, bool honour_env = true)
{
    *ndiag = 0;
    const uint32_t n = A->num_rows;
    if (n < 8 || A->num_nonzeros == 0 || (honour_env && !dia_enabled())) return clsparseSuccess;
    hipLaunchKernelGGL(dia_sample_kernel, dim3(1), dim3(64), 0, s, n,
                       dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer),
                       dev_ptr<uint32_t>(A->col_indices, A->off_col_indices), scratch);
    CLS_HIP_TRY(hipGetLastError());
    std::vector<int32_t> h(16 * kDiaSamples);
    CLS_HIP_TRY(hipMemcpyAsync(h.data(), scratch, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CLS_HIP_TRY(hipStreamSynchronize(s));
    std::vector<int32_t> offs;
    for (int k = 0; k < kDiaSamples; ++k) {
        const int32_t cnt = h[16 * k];
        if (cnt > 15) return clsparseSuccess;
        for (int32_t j = 0; j < cnt; ++j) {
            const int32_t o = h[16 * k + 1 + j];
            if (std::find(offs.begin(), offs.end(), o) == offs.end()) offs.push_back(o);
        }
        if (offs.size() > static_cast<size_t>(kDiaMax)) return clsparseSuccess;
    }
    if (offs.empty()) return clsparseSuccess;
    if (static_cast<double>(offs.size()) * n > 1.5 * static_cast<double>(A->num_nonzeros))
        return clsparseSuccess;
    std::sort(offs.begin(), offs.end());
    for (size_t d = 0; d < offs.size(); ++d) off->o[d] = offs[d];
    *ndiag = static_cast<int>(offs.size());
    return clsparseSuccess;
}

// Fills `vals` (ndiag * n values, from the workspace) with A's values in
// diagonal storage; *ndiag = 0 when a row does not fit the candidates (the
// CSR SpMV stays).
template <typename TA, typename TO>
clsparseStatus dia_fill(const clsparseCsrMatrix* A, size_t n, int* ndiag, const DiaOffsets& off, TO* vals,
                        int* bad, hipStream_t s, bool* sym)
{
    *sym = false;
    hipLaunchKernelGGL((fill_kernel<TO>), dim3(ew_grid(*ndiag * n)), dim3(256), 0, s, *ndiag * n, vals,
                       TO(0));
    hipLaunchKernelGGL((fill_kernel<int>), dim3(1), dim3(256), 0, s, size_t(1), bad, 0);
    hipLaunchKernelGGL((dia_fill_kernel<TA, TO>), dim3((n + 255) / 256), dim3(256), 0, s,
                       static_cast<uint32_t>(n), dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer),
                       dev_ptr<uint32_t>(A->col_indices, A->off_col_indices),
                       dev_ptr<TA>(A->values, A->off_values), *ndiag, off, vals, bad);
    CLS_HIP_TRY(hipGetLastError());
    int h = 0;
    CLS_HIP_TRY(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    CLS_HIP_TRY(hipStreamSynchronize(s));
    if (h) {
        *ndiag = 0;
        return clsparseSuccess;
    }
    // Symmetric storage: offsets -o..o (odd count, 0 in the middle) and equal
    // mirrored values; then only the lower diagonals are read.
    const int K = *ndiag;
    bool offs_sym = (K % 2 == 1) && K >= 3 && off.o[K / 2] == 0;
    for (int d = 0; offs_sym && d < K; ++d) offs_sym = off.o[d] == -off.o[K - 1 - d];
    const char* e = std::getenv("CLSPARSE_AMD_DIA_SYM");
    if (!offs_sym || (e && e[0] == '0')) return clsparseSuccess;
    hipLaunchKernelGGL((dia_sym_check_kernel<TO>), dim3((n + 255) / 256), dim3(256), 0, s,
                       static_cast<uint32_t>(n), K, off, vals, bad);
    CLS_HIP_TRY(hipGetLastError());
    CLS_HIP_TRY(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    CLS_HIP_TRY(hipStreamSynchronize(s));
    *sym = (h == 0);
    return clsparseSuccess;
}


}  // namespace
