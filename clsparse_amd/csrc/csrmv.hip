// CSR SpMV for gfx950: y = alpha * A * x + beta * y.
//
// Two drivers, selected as the reference selects them
// (src/library/blas2/clsparse-csrmv.hpp L27-56):
//   * CSR-Adaptive (meta present): one wave64 workgroup per row-block entry of
//     the reference metadata (src/library/internal/data-types/csr-meta.hpp
//     L93-261); single-wave workgroups release their LDS as soon as their
//     block is done (4-wave groups measured lower occupancy). Each entry is
//     classified exactly as csrmv_adaptive.cl L349-359, L382, L515, L564
//     (resolved once into a BlockDesc when the metadata is created):
//       - stream   (>= 2 rows, <= 1024 nnz): coalesced val/col loads, products
//                   staged in LDS, per-row reduction by a power-of-two lane
//                   group (DPP butterflies) or one lane per row;
//       - vector   (1 row, segment id 0): lane-strided FMA + wave reduction;
//       - long row (segment id != 0 or empty range): one 3072-nnz segment per
//                   wave, partial sum to a per-entry slot; a second kernel
//                   adds the segments of each long row in segment order.
//     The long-row combine replaces the reference's float atomics and
//     inter-workgroup spin flag (csrmv_adaptive.cl L564-686): results are
//     deterministic and the metadata is never written.
//   * CSR-Vector (meta == NULL): one sub-wave of 2..64 lanes per row, the
//     sub-wave width chosen from nnz/row as csrmv-vector.hpp L35-49.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>

#include "accum.hpp"
#include "common.hpp"

namespace {

constexpr int kWavesPerWg = 1;  // one wave per workgroup: each wave frees its LDS on exit
constexpr int kStreamPerLane = kBlkSize / 64;  // 16 nnz per lane per stream block

// LDS of one stream block: products (or alpha * values on the row-wise path),
// the columns for the row-wise path, the product errors in extended mode.
template <typename T, bool EXT, bool ROWWISE>
struct StreamLds
{
    T p[kWavesPerWg][kBlkSize];
};
template <typename T>
struct StreamLds<T, false, true>
{
    T p[kWavesPerWg][kBlkSize];
    uint32_t c[kWavesPerWg][kBlkSize];
};
template <typename T, bool ROWWISE>
struct StreamLds<T, true, ROWWISE>
{
    T p[kWavesPerWg][kBlkSize];
    T e[kWavesPerWg][kBlkSize];
};

template <typename T, bool EXT>
__device__ __forceinline__ void write_row(T* __restrict__ y, uint32_t row, const Acc<T, EXT>& a,
                                          T beta)
{
    // beta == 0 never reads y (csrmv_adaptive.cl L481-483, L505-507, L557-559).
    // Non-temporal: y is not read again by the kernel (the lower triangle of
    // the 160^3 stencil in FP64: 70 -> 59 us; the others unchanged).
    __builtin_nontemporal_store((beta != T(0)) ? a.with_beta(beta, y[row]) : a.value(), y + row);
}

// x gathers. FP32: through a buffer resource with 32-bit byte offsets (less
// address traffic into the texture unit than 64-bit flat addresses; measured
// ~8 % faster on the 128^3 stencil) when ncols * 4 fits 32 bits. FP64: flat
// loads (the buffer form measured slower there).
template <typename T>
struct XGather;
template <>
struct XGather<double>
{
    const double* x;
    __device__ __forceinline__ XGather(const double* xp, uint32_t) : x(xp) {}
    __device__ __forceinline__ double operator()(uint32_t c) const { return x[c]; }
};
template <>
struct XGather<float>
{
    const float* x;
    __amdgpu_buffer_rsrc_t rs;
    bool buf;
    __device__ __forceinline__ XGather(const float* xp, uint32_t ncols)
        : x(xp),
          rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(xp), 0, ncols * 4u, 0x00020000)),
          buf(ncols < (1u << 30))
    {
    }
    __device__ __forceinline__ float operator()(uint32_t c) const
    {
        if (buf)
            return __int_as_float(static_cast<int>(__builtin_amdgcn_raw_buffer_load_b32(rs, c * 4u, 0, 0)));
        return x[c];
    }
};

// Stored value -> compute type. Values are stored as T, or (FP32 compute
// only) as BF16 bit patterns: the preconditioner factors of the learned FSAI.
template <typename T, typename TV>
__device__ __forceinline__ T val_to(TV v)
{
    if constexpr (sizeof(TV) == 2)
        return __uint_as_float(static_cast<uint32_t>(v) << 16);
    else
        return v;
}

// Sum over [beg, end) of (alpha * val) * x[col], lanes strided by 64, four
// loads in flight per lane per step.
template <typename T, typename TV, bool EXT>
__device__ __forceinline__ Acc<T, EXT> strided_dot(const TV* __restrict__ vals,
                                                   const uint32_t* __restrict__ cols,
                                                   const XGather<T>& xg, T alpha,
                                                   uint32_t beg, uint32_t end, uint32_t lane)
{
    Acc<T, EXT> acc;
    uint32_t j = beg + lane;
    for (; j + 192 < end; j += 256) {
        TV v[4];
        uint32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u] = __builtin_nontemporal_load(vals + j + 64 * u);
            c[u] = __builtin_nontemporal_load(cols + j + 64 * u);
        }
        T xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xg(c[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc.mul_add(alpha * val_to<T>(v[u]), xv[u]);
    }
    for (; j < end; j += 64) {
        const TV v = __builtin_nontemporal_load(vals + j);
        const uint32_t c = __builtin_nontemporal_load(cols + j);
        acc.mul_add(alpha * val_to<T>(v), xg(c));
    }
    return acc;
}

// a * b rounded on its own: never contracted into a following add.
template <typename T>
__device__ __forceinline__ T mul_rounded(T a, T b)
{
#pragma clang fp contract(off)
    return a * b;
}

// One row's sum from the block's staged alpha*values and columns: W gathers
// in flight per step, products added in row order. Indices past the row stay
// inside the tile; their gathers are masked off.
template <int W, typename T, bool EXT, typename XG>
__device__ __forceinline__ Acc<T, EXT> row_from_lds(const T* __restrict__ buf,
                                                    const uint32_t* __restrict__ cols, uint32_t s,
                                                    uint32_t e, const XG& xg)
{
    Acc<T, EXT> acc;
    for (uint32_t j = s; j < e; j += W) {
        T pr[W];
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const uint32_t idx = min(j + u, uint32_t(kBlkSize - 1));
            const T a = buf[idx];
            const uint32_t col = cols[idx];
            T xv = T(0);
            if (j + u < e) xv = xg(col);
            pr[u] = mul_rounded(a, xv);
        }
#pragma unroll
        for (int u = 0; u < W; ++u)
            if (j + u < e) acc.add(pr[u]);
    }
    return acc;
}

// Row offsets (relative to the block's first nnz) of rows row .. row+511 held
// in registers: lane l keeps row l + 64k in slot k. Loaded together with the
// block's values so the reduction needs no dependent global loads.
constexpr int kRowSlots = 8;

// ROWWISE: compile the row-wise path (its column staging costs LDS and with it
// occupancy, which the gather-bound matrices with irregular rows need more).
template <typename T, typename TV, bool EXT, bool ROWWISE>
__global__ __launch_bounds__(64 * kWavesPerWg) void csrmv_adaptive_kernel(
    const TV* __restrict__ vals, const uint32_t* __restrict__ cols,
    const uint32_t* __restrict__ rowPtr, const T* __restrict__ x, T* __restrict__ y,
    const BlockDesc* __restrict__ desc, uint32_t numBlocks, uint32_t bpw,
    const T* __restrict__ pAlpha, const T* __restrict__ pBeta, double* __restrict__ partials,
    const int* __restrict__ done, uint32_t ncols)
{
    __shared__ StreamLds<T, EXT, ROWWISE> lds;

    if (done != nullptr && *done != 0) return;

    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wg = xcd_remap(blockIdx.x, gridDim.x);
    const XGather<T> xg(x, ncols);
    // Each wave walks bpw consecutive entries; every branch below is wave-uniform
    // and no workgroup barrier is used.
    const uint32_t first = __builtin_amdgcn_readfirstlane((wg * kWavesPerWg + wave) * bpw);
    for (uint32_t q = 0; q < bpw; ++q) {
    const uint32_t blk = first + q;
    if (blk >= numBlocks) break;
    wave_sync_lds();  // the previous block's LDS reads are done before new writes

    const BlockDesc d = desc[blk];
    const uint32_t row = d.row;
    const uint32_t nrows = d.nrows;
    const uint32_t base = d.base;
    const uint32_t cnt = d.cnt;
    const T alpha = *pAlpha;
    const T beta = *pBeta;

    if (nrows > static_cast<uint32_t>(kRowsForVector)) {
        // ---- CSR-Stream -------------------------------------------------
        const bool in_regs = nrows < 64 * kRowSlots;
        uint32_t ro[kRowSlots];
        if (in_regs) {
#pragma unroll
            for (int k = 0; k < kRowSlots; ++k) {
                const uint32_t r = lane + 64 * k;
                ro[k] = (r <= nrows) ? rowPtr[row + r] - base : 0u;
            }
        }
        T* buf = lds.p[wave];
        // Blocks of more than 64 short rows (stencils): each lane walks whole
        // rows, so one gather instruction reads x at the same offset from 64
        // consecutive rows -- a few cache lines instead of one per row run.
        // The block's alpha*values and columns are staged in LDS for that.
        const bool rowwise = ROWWISE && !EXT && in_regs && nrows > 64;
        {
            // 16-byte loads: lane owns elements 4*lane + 256*k .. +3.
            TV v[kStreamPerLane];
            uint32_t c[kStreamPerLane];
#pragma unroll
            for (int k = 0; k < kStreamPerLane / 4; ++k) {
                const uint32_t i0 = 256 * k + 4 * lane;
                if (!ROWWISE && i0 + 3 < cnt) {
                    // Read once: non-temporal, so the stream leaves the
                    // caches to the x gathers (the row-wise variant, for
                    // stencils, is faster with ordinary loads).
                    typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));
                    const u32x4a cc = __builtin_nontemporal_load(reinterpret_cast<const u32x4a*>(cols + base + i0));
                    for (int u = 0; u < 4; ++u) c[4 * k + u] = cc[u];
                    if constexpr (sizeof(TV) == 8) {
                        const u32x4a v0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4a*>(vals + base + i0));
                        const u32x4a v1 = __builtin_nontemporal_load(reinterpret_cast<const u32x4a*>(vals + base + i0 + 2));
                        __builtin_memcpy(&v[4 * k], &v0, 16);
                        __builtin_memcpy(&v[4 * k + 2], &v1, 16);
                    } else {
                        __builtin_memcpy(&v[4 * k], vals + base + i0, 4 * sizeof(TV));
                    }
                } else if (i0 + 3 < cnt) {
                    __builtin_memcpy(&v[4 * k], vals + base + i0, 4 * sizeof(TV));
                    __builtin_memcpy(&c[4 * k], cols + base + i0, 4 * sizeof(uint32_t));
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (i0 + u < cnt) {
                            v[4 * k + u] = vals[base + i0 + u];
                            c[4 * k + u] = cols[base + i0 + u];
                        }
                }
            }
            if (rowwise) {
                if constexpr (!EXT && ROWWISE) {
#pragma unroll
                    for (int k = 0; k < kStreamPerLane; ++k) {
                        const uint32_t i = 256 * (k / 4) + 4 * lane + (k % 4);
                        if (i < cnt) {
                            buf[i] = alpha * val_to<T>(v[k]);
                            lds.c[wave][i] = c[k];
                        }
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < kStreamPerLane; ++k) {
                    const uint32_t i = 256 * (k / 4) + 4 * lane + (k % 4);
                    if (i < cnt) {
                        const T a = alpha * val_to<T>(v[k]);
                        const T xv = xg(c[k]);
                        const T p = a * xv;
                        buf[i] = p;
                        if constexpr (EXT) lds.e[wave][i] = fma(a, xv, -p);
                    }
                }
            }
        }
        wave_sync_lds();

        if (nrows <= 64) {
            // Lane groups of 64 / 2^ceil(log2(nrows)) lanes per row.
            const uint32_t lg = 32 - __clz(nrows - 1);
            const uint32_t tpr = 64u >> lg;
            const uint32_t lr = lane >> (6 - lg);
            const uint32_t t = lane & (tpr - 1);
            const uint32_t s = __shfl(ro[0], static_cast<int>(lr & 63), 64);
            const uint32_t e_in = __shfl(ro[0], static_cast<int>((lr + 1) & 63), 64);
            const uint32_t e_wrap = __shfl(ro[1], 0, 64);
            const uint32_t e = (lr + 1 < 64) ? e_in : e_wrap;
            Acc<T, EXT> acc;
            if (lr < nrows) {
                for (uint32_t j = s + t; j < e; j += tpr) {
                    acc.add(buf[j]);
                    if constexpr (EXT) acc.c += lds.e[wave][j];
                }
            }
            acc = group_reduce(acc, tpr);
            if (t == 0 && lr < nrows) write_row<T, EXT>(y, row + lr, acc, beta);
        } else if (rowwise) {
            if constexpr (!EXT && ROWWISE) {
                // Each lane walks whole rows, 4 or 8 gathers in flight (by
                // the block's mean row length), each product rounded on its
                // own and added in row order (the same sums as the
                // product-staging path).
                const bool narrow = cnt <= 4u * nrows;
#pragma unroll
                for (int k = 0; k < kRowSlots; ++k) {
                    const uint32_t nxt = __shfl(ro[k], static_cast<int>((lane + 1) & 63), 64);
                    const uint32_t wrap = (k + 1 < kRowSlots) ? __shfl(ro[k + 1 < kRowSlots ? k + 1 : k], 0, 64) : 0u;
                    const uint32_t lr = lane + 64 * k;
                    if (lr < nrows) {
                        const uint32_t s = ro[k];
                        const uint32_t e = (lane < 63) ? nxt : wrap;
                        const Acc<T, EXT> acc = narrow
                                                    ? row_from_lds<4, T, EXT>(buf, lds.c[wave], s, e, xg)
                                                    : row_from_lds<8, T, EXT>(buf, lds.c[wave], s, e, xg);
                        write_row<T, EXT>(y, row + lr, acc, beta);
                    }
                }
            }
        } else if (in_regs) {
            // More rows than lanes: each lane reduces whole rows from LDS.
#pragma unroll
            for (int k = 0; k < kRowSlots; ++k) {
                const uint32_t nxt = __shfl(ro[k], static_cast<int>((lane + 1) & 63), 64);
                const uint32_t wrap = (k + 1 < kRowSlots) ? __shfl(ro[k + 1 < kRowSlots ? k + 1 : k], 0, 64) : 0u;
                const uint32_t lr = lane + 64 * k;
                if (lr < nrows) {
                    const uint32_t s = ro[k];
                    const uint32_t e = (lane < 63) ? nxt : wrap;
                    Acc<T, EXT> acc;
                    uint32_t j = s;
                    if constexpr (!EXT) {
                        // The first 8 terms: independent LDS reads (index
                        // clamped inside the tile), then adds in row order.
                        T t8[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) t8[u] = buf[min(s + u, uint32_t(kBlkSize - 1))];
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (s + u < e) acc.add(t8[u]);
                        j = s + 8;
                    }
                    for (; j < e; ++j) {
                        acc.add(buf[j]);
                        if constexpr (EXT) acc.c += lds.e[wave][j];
                    }
                    write_row<T, EXT>(y, row + lr, acc, beta);
                }
            }
        } else {
            // Hundreds of (mostly empty) rows: offsets straight from memory.
            for (uint32_t lr = lane; lr < nrows; lr += 64) {
                const uint32_t s = rowPtr[row + lr] - base;
                const uint32_t e = rowPtr[row + lr + 1] - base;
                Acc<T, EXT> acc;
                for (uint32_t j = s; j < e; ++j) {
                    acc.add(buf[j]);
                    if constexpr (EXT) acc.c += lds.e[wave][j];
                }
                write_row<T, EXT>(y, row + lr, acc, beta);
            }
        }
    } else if (nrows == 1) {
        // ---- CSR-Vector: one row, <= 3072 nnz --------------------------
        Acc<T, EXT> acc = strided_dot<T, TV, EXT>(vals, cols, xg, alpha, base, base + cnt, lane);
        acc = wave_reduce(acc);
        if (lane == 0) write_row<T, EXT>(y, row, acc, beta);
    } else {
        // ---- CSR-LongRows segment --------------------------------------
        Acc<T, EXT> acc = strided_dot<T, TV, EXT>(vals, cols, xg, alpha, base, base + cnt, lane);
        acc = wave_reduce(acc);
        if (lane == 0) {
            // Segment partial (sum, and its running error in extended mode).
            partials[2 * blk] = static_cast<double>(acc.s);
            if constexpr (EXT) partials[2 * blk + 1] = static_cast<double>(acc.c);
        }
    }
    }
}

// y[row] = beta * y[row] + sum of the row's segment partials, in segment
// order. One wave per long row: the lanes load 64 partials at a time into
// LDS (loads in parallel, not one dependent load per segment), lane 0 adds
// them in segment order.
template <typename T, bool EXT>
__global__ __launch_bounds__(64) void csrmv_longrow_combine(const LongRow* __restrict__ lr,
                                                           uint32_t numLong,
                                                           const double* __restrict__ partials,
                                                           T* __restrict__ y,
                                                           const T* __restrict__ pBeta,
                                                           const int* __restrict__ done)
{
    __shared__ double ps[64];
    __shared__ double pe[EXT ? 64 : 1];
    if (done != nullptr && *done != 0) return;
    const uint32_t i = blockIdx.x;
    if (i >= numLong) return;
    const LongRow L = lr[i];
    const uint32_t lane = threadIdx.x;
    Acc<double, EXT> acc;
    for (uint32_t w0 = 0; w0 < L.num_segments; w0 += 64) {
        const uint32_t w = w0 + lane;
        if (w < L.num_segments) {
            ps[lane] = partials[2 * (static_cast<size_t>(L.first_entry) + w)];
            if constexpr (EXT) pe[lane] = partials[2 * (static_cast<size_t>(L.first_entry) + w) + 1];
        }
        __syncthreads();
        if (lane == 0) {
            const uint32_t cnt = min(64u, L.num_segments - w0);
            for (uint32_t k = 0; k < cnt; ++k) {
                acc.add(ps[k]);
                if constexpr (EXT) acc.c += pe[k];
            }
        }
        __syncthreads();
    }
    if (lane != 0) return;
    const T beta = *pBeta;
    if (beta == T(0)) {
        y[L.row] = static_cast<T>(acc.value());
    } else if constexpr (EXT) {
        acc.mul_add(static_cast<double>(beta), static_cast<double>(y[L.row]));
        y[L.row] = static_cast<T>(acc.value());
    } else {
        y[L.row] = static_cast<T>(fma(static_cast<double>(beta), static_cast<double>(y[L.row]), acc.value()));
    }
}

// CSR-Vector fallback (csrmv_general.cl L189-259): SW lanes per row.
template <typename T, bool EXT, unsigned SW>
__global__ __launch_bounds__(256) void csrmv_vector_kernel(
    uint32_t numRows, const T* __restrict__ pAlpha, const uint32_t* __restrict__ rowPtr,
    const uint32_t* __restrict__ cols, const T* __restrict__ vals, const T* __restrict__ x,
    const T* __restrict__ pBeta, T* __restrict__ y, const int* __restrict__ done)
{
    if (done != nullptr && *done != 0) return;
    const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & (SW - 1);
    const uint64_t nvec = static_cast<uint64_t>(gridDim.x) * blockDim.x / SW;
    const T alpha = *pAlpha;
    const T beta = *pBeta;
    // Every lane of a wave runs the same number of trips so the cross-lane
    // reduction always sees a full wave.
    const uint64_t first = gid / SW;
    const uint64_t wave_first = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u)) / SW;
    for (uint64_t base = wave_first; base < numRows; base += nvec) {
        const uint64_t row = base + (first - wave_first);
        Acc<T, EXT> acc;
        if (row < numRows) {
            const uint32_t s = rowPtr[row];
            const uint32_t e = rowPtr[row + 1];
            for (uint32_t j = s + lane; j < e; j += SW) acc.mul_add(alpha * vals[j], x[cols[j]]);
        }
        acc = group_reduce(acc, SW);
        if (lane == 0 && row < numRows) write_row<T, EXT>(y, static_cast<uint32_t>(row), acc, beta);
    }
}

template <typename T, bool EXT, unsigned SW>
void launch_vector_sw(uint32_t n, const T* alpha, const uint32_t* rp, const uint32_t* ci,
                      const T* v, const T* x, const T* beta, T* y, const int* done,
                      hipStream_t s)
{
    const uint64_t threads = static_cast<uint64_t>(n) * SW;
    uint64_t wgs = (threads + 255) / 256;
    if (wgs > (1u << 20)) wgs = (1u << 20);
    if (wgs == 0) wgs = 1;
    hipLaunchKernelGGL((csrmv_vector_kernel<T, EXT, SW>), dim3(static_cast<uint32_t>(wgs)),
                       dim3(256), 0, s, n, alpha, rp, ci, v, x, beta, y, done);
}

template <typename T, bool EXT>
void launch_vector(uint32_t n, uint32_t nnz, const T* alpha, const uint32_t* rp,
                   const uint32_t* ci, const T* v, const T* x, const T* beta, T* y,
                   const int* done, hipStream_t s)
{
    // Sub-wave width from the average row length (csrmv-vector.hpp L35-49, wave 64).
    const uint32_t npr = nnz / n;
    if (npr < 4) launch_vector_sw<T, EXT, 2>(n, alpha, rp, ci, v, x, beta, y, done, s);
    else if (npr < 8) launch_vector_sw<T, EXT, 4>(n, alpha, rp, ci, v, x, beta, y, done, s);
    else if (npr < 16) launch_vector_sw<T, EXT, 8>(n, alpha, rp, ci, v, x, beta, y, done, s);
    else if (npr < 32) launch_vector_sw<T, EXT, 16>(n, alpha, rp, ci, v, x, beta, y, done, s);
    else if (npr < 64) launch_vector_sw<T, EXT, 32>(n, alpha, rp, ci, v, x, beta, y, done, s);
    else launch_vector_sw<T, EXT, 64>(n, alpha, rp, ci, v, x, beta, y, done, s);
}

}  // namespace

template <typename T>
clsparseStatus launch_csrmv(const T* alpha, const clsparseCsrMatrix* A, const T* x,
                            const T* beta, T* y, hipStream_t s, bool ext, const int* done)
{
    const T* vals = dev_ptr<T>(A->values, A->off_values);
    const uint32_t* cols = dev_ptr<uint32_t>(A->col_indices, A->off_col_indices);
    const uint32_t* rp = dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer);

    if (A->num_rows == 0) return clsparseSuccess;

    if (A->meta == nullptr) {
        if (ext)
            launch_vector<T, true>(A->num_rows, A->num_nonzeros, alpha, rp, cols, vals, x, beta,
                                   y, done, s);
        else
            launch_vector<T, false>(A->num_rows, A->num_nonzeros, alpha, rp, cols, vals, x,
                                    beta, y, done, s);
        CLS_HIP_TRY(hipGetLastError());
        return clsparseSuccess;
    }

    const matrix_meta* m = get_meta(A);
    if (m->magic != 0x4d455441u || m->rowBlockSize == 0) return clsparseStructInvalid;
    // The metadata describes one sparsity pattern; refuse a matrix of another
    // shape rather than index past its arrays.
    if (m->num_rows != A->num_rows || m->num_nonzeros != A->num_nonzeros)
        return clsparseStructInvalid;

    const uint32_t numBlocks = m->entries - 1;
    if (numBlocks > 0) {
        uint32_t bpw = 1;
        if (const char* e = std::getenv("CLSPARSE_AMD_BPW")) bpw = std::max(1, std::atoi(e));
        const uint32_t per_wg = kWavesPerWg * bpw;
        const uint32_t wgs = (numBlocks + per_wg - 1) / per_wg;
        // FP32 always compiles the row-wise path (its LDS still fits five waves
        // per SIMD); FP64 only where nearly every block takes it.
        const bool rowwise = sizeof(T) == 4 || m->rowwise_fraction >= 0.9;
        if (ext)
            hipLaunchKernelGGL((csrmv_adaptive_kernel<T, T, true, false>), dim3(wgs), dim3(64 * kWavesPerWg), 0,
                               s, vals, cols, rp, x, y, m->d_desc, numBlocks, bpw, alpha, beta,
                               m->d_partials, done, A->num_cols);
        else if (rowwise)
            hipLaunchKernelGGL((csrmv_adaptive_kernel<T, T, false, true>), dim3(wgs), dim3(64 * kWavesPerWg), 0,
                               s, vals, cols, rp, x, y, m->d_desc, numBlocks, bpw, alpha, beta,
                               m->d_partials, done, A->num_cols);
        else
            hipLaunchKernelGGL((csrmv_adaptive_kernel<T, T, false, false>), dim3(wgs), dim3(64 * kWavesPerWg), 0,
                               s, vals, cols, rp, x, y, m->d_desc, numBlocks, bpw, alpha, beta,
                               m->d_partials, done, A->num_cols);
        CLS_HIP_TRY(hipGetLastError());
    }
    if (m->num_long_rows > 0) {
        const uint32_t g = m->num_long_rows;  // one wave per long row
        if (ext)
            hipLaunchKernelGGL((csrmv_longrow_combine<T, true>), dim3(g), dim3(64), 0, s,
                               m->d_long, m->num_long_rows, m->d_partials, y, beta, done);
        else
            hipLaunchKernelGGL((csrmv_longrow_combine<T, false>), dim3(g), dim3(64), 0, s,
                               m->d_long, m->num_long_rows, m->d_partials, y, beta, done);
        CLS_HIP_TRY(hipGetLastError());
    }
    return clsparseSuccess;
}

// FP32 SpMV whose matrix values are BF16 bit patterns (CSR-Adaptive only).
clsparseStatus launch_csrmv_bf16(const float* alpha, const clsparseCsrMatrix* A, const float* x,
                                 const float* beta, float* y, hipStream_t s, const int* done)
{
    if (A->num_rows == 0) return clsparseSuccess;
    const matrix_meta* m = get_meta(A);
    if (m == nullptr || m->magic != 0x4d455441u || m->rowBlockSize == 0 || m->num_rows != A->num_rows ||
        m->num_nonzeros != A->num_nonzeros)
        return clsparseStructInvalid;
    const uint16_t* vals = dev_ptr<uint16_t>(A->values, A->off_values);
    const uint32_t* cols = dev_ptr<uint32_t>(A->col_indices, A->off_col_indices);
    const uint32_t* rp = dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer);
    const uint32_t numBlocks = m->entries - 1;
    if (numBlocks > 0) {
        const uint32_t wgs = (numBlocks + kWavesPerWg - 1) / kWavesPerWg;
        hipLaunchKernelGGL((csrmv_adaptive_kernel<float, uint16_t, false, true>), dim3(wgs), dim3(64 * kWavesPerWg), 0,
                           s, vals, cols, rp, x, y, m->d_desc, numBlocks, 1u, alpha, beta, m->d_partials,
                           done, A->num_cols);
        CLS_HIP_TRY(hipGetLastError());
    }
    if (m->num_long_rows > 0) {
        const uint32_t g = m->num_long_rows;  // one wave per long row
        hipLaunchKernelGGL((csrmv_longrow_combine<float, false>), dim3(g), dim3(64), 0, s, m->d_long,
                           m->num_long_rows, m->d_partials, y, beta, done);
        CLS_HIP_TRY(hipGetLastError());
    }
    return clsparseSuccess;
}

template clsparseStatus launch_csrmv<float>(const float*, const clsparseCsrMatrix*, const float*,
                                            const float*, float*, hipStream_t, bool, const int*);
template clsparseStatus launch_csrmv<double>(const double*, const clsparseCsrMatrix*,
                                             const double*, const double*, double*, hipStream_t,
                                             bool, const int*);

namespace {

template <typename T>
clsparseStatus csrmv_entry(const clsparseScalar* alpha, const clsparseCsrMatrix* matx,
                           const cldenseVector* x, const clsparseScalar* beta, cldenseVector* y,
                           const clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (alpha == nullptr || matx == nullptr || x == nullptr || beta == nullptr || y == nullptr)
        return clsparseInvalidValue;
    if (matx->num_rows > 0 && (y->num_values < matx->num_rows || x->num_values < matx->num_cols))
        return clsparseInvalidSize;
    if (matx->num_rows > 0 &&
        (matx->row_pointer == nullptr || alpha->value == nullptr || beta->value == nullptr ||
         x->values == nullptr || y->values == nullptr ||
         (matx->num_nonzeros > 0 && (matx->values == nullptr || matx->col_indices == nullptr))))
        return clsparseInvalidMemObject;
    CLS_TRY(launch_csrmv<T>(dev_ptr<T>(alpha->value, alpha->off_value), matx,
                            dev_ptr<T>(x->values, x->off_values),
                            dev_ptr<T>(beta->value, beta->off_value),
                            dev_ptr<T>(y->values, y->off_values), control->stream,
                            control->extended_precision, nullptr));
    return finish_op(control);
}

}  // namespace

// src/library/blas2/clsparse-csrmv.cpp L21-48
extern "C" clsparseStatus clsparseScsrmv(const clsparseScalar* alpha,
                                         const clsparseCsrMatrix* matx, const cldenseVector* x,
                                         const clsparseScalar* beta, cldenseVector* y,
                                         const clsparseControl control)
{
    return csrmv_entry<float>(alpha, matx, x, beta, y, control);
}

// src/library/blas2/clsparse-csrmv.cpp L50-76
extern "C" clsparseStatus clsparseDcsrmv(const clsparseScalar* alpha,
                                         const clsparseCsrMatrix* matx, const cldenseVector* x,
                                         const clsparseScalar* beta, cldenseVector* y,
                                         const clsparseControl control)
{
    return csrmv_entry<double>(alpha, matx, x, beta, y, control);
}

// ---- sum of A x in storage order ---------------------------------------------
// sum = sum_j vals[j] * x[cols[j]] over all stored entries, ignoring the rows:
// 1^T (A x), the checksum of an SpMV, from exactly the loads an SpMV of A
// cannot avoid (values, columns, x gathers) -- its time is the floor for any
// CSR SpMV that gathers x directly. FP64 accumulation, fixed order.

namespace {

constexpr int kSumAxUnroll = 8;

template <typename T>
__global__ __launch_bounds__(256) void sum_ax_kernel(uint64_t nnz, const T* __restrict__ vals,
                                                     const uint32_t* __restrict__ cols,
                                                     const T* __restrict__ x, double* __restrict__ part)
{
    // Each lane reads groups of 4 consecutive values and columns with 16-byte
    // non-temporal loads (as the SpMV's stream path does), kSumAxUnroll / 4
    // groups in flight, then gathers their 4 x values.
    typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));
    constexpr int G = kSumAxUnroll / 4;
    __shared__ double sh[4];
    const uint64_t ngroups = nnz / 4;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
    double acc = 0.0;
    for (uint64_t g0 = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; g0 < ngroups; g0 += G * stride) {
        T v[G][4];
        uint32_t c[G][4];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const uint64_t g = g0 + u * stride;
            if (g < ngroups) {
                const u32x4a cc = __builtin_nontemporal_load(reinterpret_cast<const u32x4a*>(cols + 4 * g));
                for (int e = 0; e < 4; ++e) c[u][e] = cc[e];
                u32x4a vv[sizeof(T) / 4];
                for (int h = 0; h < static_cast<int>(sizeof(T) / 4); ++h)
                    vv[h] = __builtin_nontemporal_load(reinterpret_cast<const u32x4a*>(vals + 4 * g) + h);
                __builtin_memcpy(&v[u][0], &vv[0], 4 * sizeof(T));
            } else {
                for (int e = 0; e < 4; ++e) {
                    c[u][e] = 0u;
                    v[u][e] = T(0);
                }
            }
        }
        T xv[G][4];
#pragma unroll
        for (int u = 0; u < G; ++u)
            for (int e = 0; e < 4; ++e) xv[u][e] = (g0 + u * stride < ngroups) ? x[c[u][e]] : T(0);
#pragma unroll
        for (int u = 0; u < G; ++u)
            for (int e = 0; e < 4; ++e) acc += static_cast<double>(v[u][e]) * static_cast<double>(xv[u][e]);
    }
    // The last nnz % 4 entries.
    const uint64_t t = 4 * ngroups + static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t < nnz) acc += static_cast<double>(vals[t]) * static_cast<double>(x[cols[t]]);
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void sum_ax_finish(const double* __restrict__ part, uint32_t g,
                                                     T* __restrict__ out)
{
    __shared__ double sh[4];
    double a = 0.0;
    for (uint32_t i = threadIdx.x; i < g; i += 256) a += part[i];
    a = block_sum256(a, sh);
    if (threadIdx.x == 0) *out = static_cast<T>(a);
}

template <typename T>
clsparseStatus sum_ax_entry(const clsparseCsrMatrix* A, const cldenseVector* x, clsparseScalar* sum,
                            clsparseControl control)
{
    CLS_TRY(sum_ax_check<T>(A, x, sum, control));
    const uint32_t g = sum_ax_grid(A->num_nonzeros);
    void* work = nullptr;
    CLS_TRY(control_workspace(control, g * sizeof(double), &work));
    CLS_TRY(launch_sum_ax<T>(A, x, sum, static_cast<double*>(work), control->stream));
    return finish_op(control);
}

}  // namespace

uint32_t sum_ax_grid(uint64_t nnz)
{
    return static_cast<uint32_t>(std::max<uint64_t>(
        1, std::min<uint64_t>(kSumAxMaxGrid, (nnz + 256 * kSumAxUnroll - 1) / (256 * kSumAxUnroll))));
}

template <typename T>
clsparseStatus sum_ax_check(const clsparseCsrMatrix* A, const cldenseVector* x, const clsparseScalar* sum,
                            clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (A == nullptr || x == nullptr || sum == nullptr) return clsparseInvalidValue;
    if (x->num_values < A->num_cols) return clsparseInvalidSize;
    if (sum->value == nullptr || (A->num_nonzeros > 0 &&
                                  (A->values == nullptr || A->col_indices == nullptr || x->values == nullptr)))
        return clsparseInvalidMemObject;
    return clsparseSuccess;
}

// Both kernels of 1^T (A x) on `part`, which holds sum_ax_grid(nnz) doubles:
// every workgroup writes its partial before the finishing kernel reads them.
template <typename T>
clsparseStatus launch_sum_ax(const clsparseCsrMatrix* A, const cldenseVector* x, clsparseScalar* sum,
                             double* part, hipStream_t s)
{
    const uint64_t nnz = A->num_nonzeros;
    const uint32_t g = sum_ax_grid(nnz);
    hipLaunchKernelGGL((sum_ax_kernel<T>), dim3(g), dim3(256), 0, s, nnz,
                       dev_ptr<T>(A->values, A->off_values),
                       dev_ptr<uint32_t>(A->col_indices, A->off_col_indices),
                       dev_ptr<T>(x->values, x->off_values), part);
    hipLaunchKernelGGL((sum_ax_finish<T>), dim3(1), dim3(256), 0, s, part, g,
                       dev_ptr<T>(sum->value, sum->off_value));
    CLS_HIP_TRY(hipGetLastError());
    return clsparseSuccess;
}

template clsparseStatus sum_ax_check<float>(const clsparseCsrMatrix*, const cldenseVector*,
                                            const clsparseScalar*, clsparseControl);
template clsparseStatus sum_ax_check<double>(const clsparseCsrMatrix*, const cldenseVector*,
                                             const clsparseScalar*, clsparseControl);
template clsparseStatus launch_sum_ax<float>(const clsparseCsrMatrix*, const cldenseVector*, clsparseScalar*,
                                             double*, hipStream_t);
template clsparseStatus launch_sum_ax<double>(const clsparseCsrMatrix*, const cldenseVector*, clsparseScalar*,
                                              double*, hipStream_t);

extern "C" clsparseStatus clsparseAmdScsrSumAx(const clsparseCsrMatrix* A, const cldenseVector* x,
                                               clsparseScalar* sum, const clsparseControl control)
{
    return sum_ax_entry<float>(A, x, sum, control);
}

// ---- random-gather floor ------------------------------------------------------
// sum of x[h(seed, i) mod window] over i < count, window a power of two: the
// rate of independent 8-byte (or 4-byte) gathers into a window of x, the part
// of an SpMV with scattered columns that no row structure can avoid. h is a
// 32-bit integer hash (xor-shift-multiply, the finaliser of MurmurHash3 with
// Wang's constants) of i + seed.

namespace {

__device__ __forceinline__ uint32_t gather_hash(uint32_t a)
{
    a ^= a >> 16;
    a *= 0x7feb352du;
    a ^= a >> 15;
    a *= 0x846ca68bu;
    a ^= a >> 16;
    return a;
}

constexpr int kGatherUnroll = 8;

template <typename T>
__global__ __launch_bounds__(256) void gather_sum_kernel(uint64_t count, const T* __restrict__ x, uint32_t mask,
                                                         uint32_t seed, double* __restrict__ part)
{
    __shared__ double sh[4];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
    double acc = 0.0;
    for (uint64_t j0 = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; j0 < count;
         j0 += kGatherUnroll * stride) {
        T v[kGatherUnroll];
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) {
            const uint64_t j = j0 + u * stride;
            v[u] = j < count ? x[gather_hash(static_cast<uint32_t>(j) + seed) & mask] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kGatherUnroll; ++u) acc += static_cast<double>(v[u]);
    }
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <typename T>
clsparseStatus gather_sum_entry(const cldenseVector* x, cl_uint window, cl_ulong count, cl_uint seed,
                                clsparseScalar* sum, clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (x == nullptr || sum == nullptr) return clsparseInvalidValue;
    if (window == 0 || (window & (window - 1)) != 0 || window > x->num_values || count >= (1ull << 32))
        return clsparseInvalidSize;
    if (x->values == nullptr || sum->value == nullptr) return clsparseInvalidMemObject;
    const uint32_t g = static_cast<uint32_t>(std::max<uint64_t>(
        1, std::min<uint64_t>(kSumAxMaxGrid, (count + 256 * kGatherUnroll - 1) / (256 * kGatherUnroll))));
    void* work = nullptr;
    CLS_TRY(control_workspace(control, g * sizeof(double), &work));
    double* part = static_cast<double*>(work);
    hipLaunchKernelGGL((gather_sum_kernel<T>), dim3(g), dim3(256), 0, control->stream, count,
                       dev_ptr<T>(x->values, x->off_values), window - 1, seed, part);
    hipLaunchKernelGGL((sum_ax_finish<T>), dim3(1), dim3(256), 0, control->stream, part, g,
                       dev_ptr<T>(sum->value, sum->off_value));
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}

}  // namespace

extern "C" clsparseStatus clsparseAmdSgatherSum(const cldenseVector* x, cl_uint window, cl_ulong count, cl_uint seed,
                                                clsparseScalar* sum, const clsparseControl control)
{
    return gather_sum_entry<float>(x, window, count, seed, sum, control);
}

extern "C" clsparseStatus clsparseAmdDgatherSum(const cldenseVector* x, cl_uint window, cl_ulong count, cl_uint seed,
                                                clsparseScalar* sum, const clsparseControl control)
{
    return gather_sum_entry<double>(x, window, count, seed, sum, control);
}

// ---- 1^T (A x) without allocation ----------------------------------------------
// The FP64 variant's per-workgroup partials live in a file-scope device array
// (kSumAxMaxGrid doubles, 32 KiB; one copy per GPU, held until the process
// ends). Its name is a host-side stand-in: host code uses only the device
// address that hipGetSymbolAddress returns, taken once per device while a
// control object is created on it (the device is current then).

namespace {

constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[kSumAxMaxGrid];
std::atomic<double*> g_scratch_ptr[kScratchDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

clsparseStatus sum_ax_scratch_setup(int device)
{
    if (device < 0 || device >= kScratchDevices) return clsparseInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device].load() != nullptr) return clsparseSuccess;
    void* p = nullptr;
    CLS_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device].store(static_cast<double*>(p));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdDcsrSumAxNoAlloc(const clsparseCsrMatrix* A, const cldenseVector* x,
                                                      clsparseScalar* sum, const clsparseControl control)
{
    CLS_TRY(sum_ax_check<double>(A, x, sum, control));
    // The grid is capped at kSumAxMaxGrid workgroups, so every nnz fits the
    // scratch; the check guards that invariant.
    if (sum_ax_grid(A->num_nonzeros) > kSumAxMaxGrid) {
        std::fprintf(stderr, "clSPARSE: clsparseAmdDcsrSumAxNoAlloc needs %u partials, scratch holds %u\n",
                     sum_ax_grid(A->num_nonzeros), kSumAxMaxGrid);
        return clsparseInvalidSize;
    }
    if (control->device < 0 || control->device >= kScratchDevices) return clsparseInvalidDevice;
    double* part = g_scratch_ptr[control->device].load();
    if (part == nullptr) return clsparseInvalidDevice;
    // Every workgroup writes its partial before the finishing kernel reads
    // them, so nothing in the scratch has to start at zero.
    CLS_TRY(launch_sum_ax<double>(A, x, sum, part, control->stream));
    return finish_op(control);
}

extern "C" clsparseStatus clsparseAmdDcsrSumAx(const clsparseCsrMatrix* A, const cldenseVector* x,
                                               clsparseScalar* sum, const clsparseControl control)
{
    return sum_ax_entry<double>(A, x, sum, control);
}
