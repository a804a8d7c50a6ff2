// Internal declarations shared by the library's translation units.
//
// Host objects (control, solver control, CSR-Adaptive metadata) and the small
// device helpers every kernel file uses. Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "clSPARSE.h"
#include "clSPARSE-amd-ext.h"

// Global init flag (src/library/clsparse-init.cpp L25).
extern int clsparseInitialized;

// CSR-Adaptive constants of src/library/include/clSPARSE-private.hpp L46-50.
constexpr int kWgBits = 24;
constexpr int kRowBits = 32;
constexpr int kBlkSize = 1024;
constexpr int kBlockMultiplier = 3;
constexpr int kRowsForVector = 1;
constexpr int kLongSeg = kBlkSize * kBlockMultiplier;  // nnz per long-row segment

// Deterministic two-stage reductions use at most this many first-stage blocks.
constexpr int kReducePartialsMax = 1024;

// ---------------------------------------------------------------------------
// Control object (reference: src/library/internal/clsparse-control.hpp L30-80).
// One HIP stream, the device it was created on, async / extended-precision
// flags, the event of the last operation and a grow-only device workspace.
struct _clsparseControl
{
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cus = 256;
    bool async = false;
    bool extended_precision = false;
    hipEvent_t event = nullptr;

    void* work = nullptr;        // device workspace for solvers (grow-only)
    size_t work_bytes = 0;
    double* red_partials = nullptr;  // per-block partials of the BLAS-1 reductions
    void* pinned = nullptr;      // small pinned host buffer for scalar readback
    size_t pinned_bytes = 0;
    hipStream_t solver_stream = nullptr;  // non-blocking stream the solver loop is captured on
};

// Solver control (reference: src/library/solvers/solver-control.hpp L37-128).
struct _solverControl
{
    cl_int nIters = 0;
    cl_int maxIters = 0;
    PRECONDITIONER preconditioner = NOPRECOND;
    cl_double relativeTolerance = 0.0;
    cl_double absoluteTolerance = 0.0;
    cl_double initialResidual = 0.0;
    cl_double currentResidual = 0.0;
    PRINT_MODE printMode = QUIET;

    // Extension state (no reference counterpart).
    clsparseAmdSolverStats stats{};

    // finished / converged: solver-control.hpp L47-64, same evaluation order.
    bool converged(double residuum)
    {
        currentResidual = residuum;
        return residuum <= relativeTolerance || residuum <= absoluteTolerance * initialResidual;
    }
    bool finished(double residuum) { return converged(residuum) || nIters >= maxIters; }
    void print() const;
    void printSummary(clsparseStatus status) const;
};

// What one row-block entry asks of the SpMV kernel, derived once from the
// entry pair and the row pointer when the metadata is created:
//   nrows >= 2: CSR-Stream block of rows [row, row + nrows), nnz [base, base + cnt)
//   nrows == 1: CSR-Vector row `row`, nnz [base, base + cnt)
//   nrows == 0: one CSR-LongRows segment of row `row`, nnz [base, base + cnt)
struct BlockDesc
{
    uint32_t row;
    uint32_t nrows;
    uint32_t base;
    uint32_t cnt;
};

// One long row split over several row-block entries (CSR-LongRows).
struct LongRow
{
    uint32_t row;
    uint32_t first_entry;  // row-block entry of segment 0
    uint32_t num_segments;
    uint32_t pad;
};

// CSR-Adaptive metadata owned by the library (reference matrix_meta,
// src/library/include/clSPARSE-1x.hpp L45-62, plus what this library derives
// from it once at creation time).
struct matrix_meta
{
    uint32_t magic = 0x4d455441u;  // "META"
    int device = 0;
    clsparseIdx_t rowBlockSize = 0;   // 2 * entries, as the reference stores it
    uint32_t entries = 0;             // row-block entries written
    uint32_t num_rows = 0;
    uint32_t num_nonzeros = 0;
    uint64_t* d_rowBlocks = nullptr;  // rowBlockSize words; entries filled, rest zero
    std::vector<uint64_t> h_rowBlocks;  // host copy of the filled entries
    BlockDesc* d_desc = nullptr;      // entries - 1 kernel descriptors
    // Stream blocks of 65..511 rows (the row-wise path) as a fraction of all
    // blocks: near 1 for stencil and finite-element matrices.
    double rowwise_fraction = 0.0;

    // Long-row combine tables (deterministic replacement of the reference's
    // atomic + spin-flag protocol, csrmv_adaptive.cl L564-686).
    uint32_t num_long_rows = 0;
    LongRow* d_long = nullptr;
    double* d_partials = nullptr;     // per row-block entry: segment sum, segment error

};

// ---------------------------------------------------------------------------
// Host helpers

#define CLS_HIP_TRY(...)                                                   \
    do {                                                                   \
        hipError_t _e = (__VA_ARGS__);                                     \
        if (_e != hipSuccess) {                                            \
            std::fprintf(stderr, "clSPARSE(hip) %s failed: %s (%s:%d)\n", #__VA_ARGS__, \
                         hipGetErrorString(_e), __FILE__, __LINE__);       \
            return clsparseInvalidKernelExecution;                         \
        }                                                                  \
    } while (0)

#define CLS_TRY(...)                           \
    do {                                       \
        clsparseStatus _s = (__VA_ARGS__);     \
        if (_s != clsparseSuccess) return _s;  \
    } while (0)

// Common argument checks of every reference entry point
// (e.g. src/library/blas2/clsparse-csrmv.cpp L29-38).
inline clsparseStatus check_init_control(const clsparseControl control)
{
    if (!clsparseInitialized) return clsparseNotInitialized;
    if (control == nullptr) return clsparseInvalidControlObject;
    return clsparseSuccess;
}

// Ends an operation: records the control's event and, unless the control is
// in async mode, waits for the stream (reference kernel-wrap.cpp L52-64).
clsparseStatus finish_op(clsparseControl control);

// Workspace of at least `bytes` on the control's device (grow-only; frees the
// old buffer after synchronising the stream).
clsparseStatus control_workspace(clsparseControl control, size_t bytes, void** out);
clsparseStatus control_pinned(clsparseControl control, size_t bytes, void** out);

template <typename T>
inline T* dev_ptr(void* base, clsparseIdx_t off)
{
    return static_cast<T*>(base) + off;
}
template <typename T>
inline const T* dev_ptr(const void* base, clsparseIdx_t off)
{
    return static_cast<const T*>(base) + off;
}

inline const matrix_meta* get_meta(const clsparseCsrMatrix* A)
{
    return static_cast<const matrix_meta*>(A->meta);
}

// SpMV launch used by the public csrmv entry points and by the solvers.
// done_flag (may be null): device int; the kernels return immediately when
// it is non-zero (used by the device-resident solver loop).
template <typename T>
clsparseStatus launch_csrmv(const T* alpha, const clsparseCsrMatrix* A, const T* x,
                            const T* beta, T* y, hipStream_t stream, bool extended,
                            const int* done_flag);

// Sparse times dense, row-major B (ldb) and C (ldc), ncols columns of C:
// C = alpha A B + beta C with alpha, beta on the device (csrmm.hip).
template <typename T>
clsparseStatus launch_csrmm_rows(const clsparseCsrMatrix* A, uint32_t ncols, const T* B, uint32_t ldb, T* C,
                                 uint32_t ldc, const T* alpha, const T* beta, int num_cus, hipStream_t stream);

// Learned FSAI factor (fsai.hip): G and G^T in FP32 / FP64 with metadata.
// Brings the factor's solver value copies up to date (after training steps).
clsparseStatus fsai_prepare(clsparseAmdFsai f, hipStream_t stream);
const clsparseCsrMatrix* fsai_G32(clsparseAmdFsai f);
const clsparseCsrMatrix* fsai_GT32(clsparseAmdFsai f);
const clsparseCsrMatrix* fsai_G64(clsparseAmdFsai f);
const clsparseCsrMatrix* fsai_GT64(clsparseAmdFsai f);
const clsparseCsrMatrix* fsai_G16(clsparseAmdFsai f);   // BF16 values, FP32 compute
const clsparseCsrMatrix* fsai_GT16(clsparseAmdFsai f);
bool fsai_valid(clsparseAmdFsai f);
// Pilot-race outcome stored with the factor (0 not run, 1 G, 2 Jacobi).
int fsai_race(clsparseAmdFsai f, double* rate_fsai, double* rate_jacobi);
void fsai_set_race(clsparseAmdFsai f, int outcome, double rate_fsai, double rate_jacobi);
uint32_t fsai_n(clsparseAmdFsai f);
// Creates the CSR metadata of G and G^T on first use (the solvers that
// apply the factor from diagonal storage never need it).
clsparseStatus fsai_ensure_meta(clsparseAmdFsai f, clsparseControl control);

// Diagonal storage of the BF16 factor, present when G's pattern lies on at
// most kFsaiMaxDiags diagonals (structured-grid stencils): vals[k * n + i] =
// G(i, i + off[k]), zero where the pattern has no entry; off ascending, all
// <= 0 (G is lower triangular), the last one 0.
constexpr int kFsaiMaxDiags = 8;
struct FsaiDia
{
    int ndiag;
    int32_t off[kFsaiMaxDiags];
    const uint16_t* vals;
};
bool fsai_dia(clsparseAmdFsai f, FsaiDia* out);

// FP32 CSR-Adaptive SpMV whose values are stored as BF16 bit patterns
// (metadata required).
clsparseStatus launch_csrmv_bf16(const float* alpha, const clsparseCsrMatrix* A, const float* x,
                                 const float* beta, float* y, hipStream_t stream,
                                 const int* done_flag);

// 1^T (A x) in storage order (clsparseAmd[S|D]csrSumAx): argument checks, the
// number of per-workgroup partials for nnz entries (at most kSumAxMaxGrid),
// and the launch of both kernels on a partials buffer of that many doubles.
constexpr uint32_t kSumAxMaxGrid = 4096;
uint32_t sum_ax_grid(uint64_t nnz);
template <typename T>
clsparseStatus sum_ax_check(const clsparseCsrMatrix* A, const cldenseVector* x, const clsparseScalar* sum,
                            clsparseControl control);
template <typename T>
clsparseStatus launch_sum_ax(const clsparseCsrMatrix* A, const cldenseVector* x, clsparseScalar* sum,
                             double* part, hipStream_t s);
// Takes the device address of the no-allocation SumAx's partials array on
// `device` (which must be current); once per device, at control creation.
clsparseStatus sum_ax_scratch_setup(int device);

// Deterministic reductions: result written to a device scalar of type R.
enum class RedOp { Sum, Abs, Sqr, Dot };
template <typename T, typename R>
clsparseStatus launch_reduce(RedOp op, const T* x, const T* y, size_t n, R* result,
                             bool sqrt_result, double* partials, hipStream_t stream);

// ---------------------------------------------------------------------------
// Device helpers (gfx950: wave64).

__device__ __forceinline__ void wave_sync_lds()
{
    // Lanes of one wave exchange data through LDS: order the LDS accesses of
    // the wave and stop the compiler from moving them across this point.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int mask)
{
    return __shfl_xor(v, mask, 64);
}

// DPP row-level butterflies for 32-bit lanes; 64-bit values move as two halves.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    int lo = __double2loint(v);
    int hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) { return dpp_f<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double dpp(double v) { return dpp_d<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ long long dpp(long long v)
{
    int lo = static_cast<int>(static_cast<unsigned long long>(v) & 0xffffffffull);
    int hi = static_cast<int>(static_cast<unsigned long long>(v) >> 32);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                  static_cast<unsigned>(lo));
}

// DPP controls (GCN/CDNA encoding).
constexpr int kDppQuadSwap1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

// Sum over aligned groups of G lanes (G a power of two, 1..64); every lane of
// a group ends with the group sum.
template <typename T>
__device__ __forceinline__ T group_sum(T v, unsigned G)
{
    if (G >= 2) v += dpp<kDppQuadSwap1>(v);
    if (G >= 4) v += dpp<kDppQuadSwap2>(v);
    if (G >= 8) v += dpp<kDppRowHalfMirror>(v);
    if (G >= 16) v += dpp<kDppRowMirror>(v);
    if (G >= 32) v += shfl_xor_t(v, 16);
    if (G >= 64) v += shfl_xor_t(v, 32);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    v += dpp<kDppQuadSwap1>(v);
    v += dpp<kDppQuadSwap2>(v);
    v += dpp<kDppRowHalfMirror>(v);
    v += dpp<kDppRowMirror>(v);
    v += shfl_xor_t(v, 16);
    v += shfl_xor_t(v, 32);
    return v;
}

// Sum over a 256-lane workgroup (4 waves), fixed order: DPP/permute inside
// each wave, then wave 0..3 in order. Every thread must call it; the result
// is valid in thread 0. `scratch` holds 4 values.
template <typename T>
__device__ __forceinline__ T block_sum256(T v, T* scratch)
{
    v = wave_sum(v);
    const unsigned wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scratch[wave] = v;
    __syncthreads();
    T r = T(0);
    if (threadIdx.x == 0) r = ((scratch[0] + scratch[1]) + scratch[2]) + scratch[3];
    __syncthreads();
    return r;
}

// Raw buffer load of a float or double at byte offset voff + soff of the
// buffer rs; offsets past the buffer's size read 0 (the hardware's range
// check), so a 32-bit offset replaces a 64-bit flat address.
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 4)
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
    else
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
}

// XCD-aware block order: hardware hands consecutive workgroups to the 8 XCDs
// round-robin; remap so that each XCD walks one contiguous range of logical
// blocks (keeps neighbouring rows -- and their x gathers -- in one L2).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblk)
{
    const uint32_t q = nblk >> 3, r = nblk & 7;
    const uint32_t xcd = bid & 7, pos = bid >> 3;
    return (xcd < r) ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}
