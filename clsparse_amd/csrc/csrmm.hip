// Sparse times dense: C = alpha * A * B + beta * C (clsparse[S/D]csrmm).
//
// Semantics of the reference (src/library/blas3/clsparse-csrmm.cpp L25-76,
// clsparse-csrmm.hpp L84-179, kernels/csrmm_general.cl L59-167): B and C are
// row-major with lead_dim as the row stride (the major field is not read),
// element (k, c) of B at off_values + k * lead_dim + c; alpha == 0 gives
// beta * C without reading B, beta == 0 never reads C, beta == 1 adds C.
// The reference runs one CSR-Vector SpMV per column of C (A re-read for every
// column). Here each lane owns V consecutive columns (V = 16 bytes of them
// when the shapes and addresses allow, else 1), W lanes (a power of two >= the
// column groups, up to 64) form a row slot, and a wave runs 64 / W rows of C
// at once: each slot loads W nonzeros of its row with one coalesced read,
// hands each nonzero's value and column to its lanes through lane permutes,
// and keeps 8 B row segments in flight per step, every one a coalesced read
// (16-byte accesses halve the load and permute instructions per byte, which
// is what bounds the kernel: the texture data path is busy for most of it). Waves
// walk consecutive rows and workgroups are remapped so each XCD covers one
// contiguous range of rows (neighbouring rows share B rows in its L2). Each
// element of C is summed over the row's nonzeros in storage order, so results
// are deterministic.
#include <algorithm>

#include "common.hpp"

namespace {

constexpr int kBInFlight = 8;     // B row segments in flight per lane (scalar columns)
constexpr int kBInFlightVec = 4;  // the same with 16-byte column groups (twice the bytes each)

// V consecutive columns of a row of B or C per lane: one 16-byte access
// when V * sizeof(T) == 16 (V = 2 doubles or 4 floats), else V scalars.
template <typename T, int V>
struct Cols
{
    T e[V];
};

template <typename T, int V>
__device__ __forceinline__ Cols<T, V> load_cols(const T* p)
{
    Cols<T, V> r;
    if constexpr (V * sizeof(T) == 16) {
        typedef T vec __attribute__((ext_vector_type(V)));
        const vec x = *reinterpret_cast<const vec*>(p);
#pragma unroll
        for (int k = 0; k < V; ++k) r.e[k] = x[k];
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k) r.e[k] = p[k];
    }
    return r;
}

template <typename T, int V>
__device__ __forceinline__ void store_cols(T* p, const Cols<T, V>& r)
{
    if constexpr (V * sizeof(T) == 16) {
        typedef T vec __attribute__((ext_vector_type(V)));
        vec x;
#pragma unroll
        for (int k = 0; k < V; ++k) x[k] = r.e[k];
        *reinterpret_cast<vec*>(p) = x;
    } else {
#pragma unroll
        for (int k = 0; k < V; ++k) p[k] = r.e[k];
    }
}

// W lanes per row slot, V columns per lane (the host picks V > 1 only when
// the columns, both leading dimensions and both base addresses are multiples
// of V elements / 16 bytes, so every access is an aligned 16-byte one).
template <typename T, int W, int V>
__global__ __launch_bounds__(256) void csrmm_kernel(uint32_t num_rows, uint32_t ncols,
                                                    const uint32_t* __restrict__ rp,
                                                    const uint32_t* __restrict__ ci,
                                                    const T* __restrict__ v, const T* __restrict__ B,
                                                    uint32_t ldb, T* __restrict__ C, uint32_t ldc,
                                                    const T* __restrict__ pAlpha,
                                                    const T* __restrict__ pBeta)
{
    constexpr uint32_t kSlots = 64 / W;
    constexpr int kIn = V > 1 ? kBInFlightVec : kBInFlight;
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
        for (uint32_t c0 = 0; c0 < ncols; c0 += W * V) {
            const uint32_t c = c0 + t * V;
            const bool colok = c < ncols;
            Cols<T, V> sum;
#pragma unroll
            for (int k = 0; k < V; ++k) sum.e[k] = T(0);
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
                    for (uint32_t q = 0; q < cnt; q += kIn) {
                        T a8[kIn];
                        Cols<T, V> b8[kIn];
#pragma unroll
                        for (int u = 0; u < kIn; ++u) {
                            const int src = base + static_cast<int>(min(q + u, static_cast<uint32_t>(W - 1)));
                            a8[u] = __shfl(av, src, 64);
                            const uint32_t col = __shfl(cc, src, 64);
#pragma unroll
                            for (int k = 0; k < V; ++k) b8[u].e[k] = T(0);
                            if (colok && q + u < cnt) b8[u] = load_cols<T, V>(B + static_cast<size_t>(col) * ldb + c);
                        }
#pragma unroll
                        for (int u = 0; u < kIn; ++u)
                            if (q + u < cnt)
#pragma unroll
                                for (int k = 0; k < V; ++k) sum.e[k] = fma(a8[u], b8[u].e[k], sum.e[k]);
                    }
                }
            }
            if (live && colok) {
                T* out = C + static_cast<size_t>(row) * ldc + c;
                if (beta != T(0)) {
                    const Cols<T, V> old = load_cols<T, V>(out);
#pragma unroll
                    for (int k = 0; k < V; ++k)
                        sum.e[k] = beta == T(1) ? sum.e[k] + old.e[k] : sum.e[k] + beta * old.e[k];
                }
                store_cols<T, V>(out, sum);
            }
        }
    }
}

// Launches csrmm_kernel<T, W, V> on at most as many workgroups as the chip
// holds at once: the kernel's resident workgroups per CU (queried once) times
// the CUs, a multiple of 8 (one equal share per XCD). The resident waves of an
// XCD then work on one compact window of its rows that advances together, so
// the B rows that neighbouring rows share stay in its L2.
template <typename T, int W, int V>
clsparseStatus launch_csrmm(uint32_t want, int num_cus, hipStream_t s, uint32_t n, uint32_t ncols,
                            const uint32_t* rp, const uint32_t* ci, const T* v, const T* B, uint32_t ldb, T* C,
                            uint32_t ldc, const T* pa, const T* pb)
{
    static const int per_cu = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, csrmm_kernel<T, W, V>, 256, 0) != hipSuccess) b = 0;
        return std::max(b, 1);
    }();
    const uint32_t resident = std::max<uint32_t>(8u, static_cast<uint32_t>(per_cu * std::max(num_cus, 1)) / 8 * 8);
    const uint32_t grid = std::min<uint32_t>(resident, (want + 7) / 8 * 8);
    hipLaunchKernelGGL((csrmm_kernel<T, W, V>), dim3(grid), dim3(256), 0, s, n, ncols, rp, ci, v, B, ldb, C, ldc,
                       pa, pb);
    CLS_HIP_TRY(hipGetLastError());
    return clsparseSuccess;
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
    const T* Bp = dev_ptr<T>(B->values, B->off_values);
    T* Cp = dev_ptr<T>(C->values, C->off_values);
    if (Bp == nullptr || Cp == nullptr || pa == nullptr || pb == nullptr) return clsparseInvalidMemObject;
    CLS_TRY(launch_csrmm_rows<T>(A, ncols, Bp, B->lead_dim, Cp, C->lead_dim, pa, pb, control->num_cus,
                                 control->stream));
    return finish_op(control);
}

}  // namespace

template <typename T>
clsparseStatus launch_csrmm_rows(const clsparseCsrMatrix* A, uint32_t ncols, const T* Bp, uint32_t ldb, T* Cp,
                                 uint32_t ldc, const T* pa, const T* pb, int num_cus, hipStream_t s)
{
    const uint32_t n = A->num_rows;
    if (n == 0 || ncols == 0) return clsparseSuccess;
    const uint32_t* rp = dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer);
    const uint32_t* ci = dev_ptr<uint32_t>(A->col_indices, A->off_col_indices);
    const T* v = dev_ptr<T>(A->values, A->off_values);
    if (rp == nullptr) return clsparseInvalidMemObject;
    // V columns per lane (one 16-byte access) when the columns, both leading
    // dimensions and both base addresses allow it; W lanes per row slot (a
    // power of two >= the column groups, at most 64).
    constexpr uint32_t kVec = 16 / sizeof(T);
    const bool vec = ncols % kVec == 0 && ldb % kVec == 0 && ldc % kVec == 0 &&
                     reinterpret_cast<uintptr_t>(Bp) % 16 == 0 && reinterpret_cast<uintptr_t>(Cp) % 16 == 0;
    const uint32_t groups = vec ? ncols / kVec : ncols;
    const uint32_t W = groups > 32 ? 64u : groups > 16 ? 32u : groups > 8 ? 16u : groups > 4 ? 8u
                                     : groups > 2 ? 4u : groups > 1 ? 2u : 1u;
    const uint32_t slots = 64 / W;
    const uint32_t want = (n + 4 * slots - 1) / (4 * slots);
#define CLS_CSRMM(WW, VV) \
    CLS_TRY((launch_csrmm<T, WW, VV>(want, num_cus, s, n, ncols, rp, ci, v, Bp, ldb, Cp, ldc, pa, pb)))
#define CLS_CSRMM_W(VV)                  \
    switch (W) {                         \
    case 64: CLS_CSRMM(64, VV); break;   \
    case 32: CLS_CSRMM(32, VV); break;   \
    case 16: CLS_CSRMM(16, VV); break;   \
    case 8: CLS_CSRMM(8, VV); break;     \
    case 4: CLS_CSRMM(4, VV); break;     \
    case 2: CLS_CSRMM(2, VV); break;     \
    default: CLS_CSRMM(1, VV); break;    \
    }
    if (vec) {
        CLS_CSRMM_W(kVec);
    } else {
        CLS_CSRMM_W(1);
    }
#undef CLS_CSRMM_W
#undef CLS_CSRMM
    CLS_HIP_TRY(hipGetLastError());
    return clsparseSuccess;
}

template clsparseStatus launch_csrmm_rows<float>(const clsparseCsrMatrix*, uint32_t, const float*, uint32_t, float*,
                                                 uint32_t, const float*, const float*, int, hipStream_t);
template clsparseStatus launch_csrmm_rows<double>(const clsparseCsrMatrix*, uint32_t, const double*, uint32_t,
                                                  double*, uint32_t, const double*, const double*, int, hipStream_t);

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
