// Helpers shared by the device-resident solvers (cg.hip, bicgstab.hip):
// fixed-order partial sums, the Jacobi diagonal, elementwise casts and fills,
// grid sizes, workspace carving, and the ordering of the private solver stream
// against the caller's stream.
#pragma once

#include <cstdlib>

#include "common.hpp"

namespace {

// Fixed-order sum of g partials; every thread of the block gets the result.
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, uint32_t g,
                                               double* sh)
{
    double a = 0.0;
    for (uint32_t i = threadIdx.x; i < g; i += 256) a += part[i];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    const double r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return r;
}

template <typename TA, typename TO>
__global__ __launch_bounds__(256) void extract_inv_diag(uint32_t n, const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ ci,
                                                        const TA* __restrict__ v,
                                                        TO* __restrict__ invD)
{
    const uint32_t row = blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    // Rows without a stored diagonal keep 1 (no scaling); the reference leaves
    // them uninitialised (src/library/solvers/preconditioners/diagonal.hpp L51).
    TO d = TO(1);
    for (uint32_t j = rp[row]; j < rp[row + 1]; ++j)
        if (ci[j] == row) {
            d = static_cast<TO>(TA(1) / v[j]);
            break;
        }
    invD[row] = d;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void cast_kernel(size_t n, const TI* __restrict__ in,
                                                   TO* __restrict__ out)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
        out[i] = static_cast<TO>(in[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(size_t n, T* __restrict__ out, T v)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
        out[i] = v;
}

uint32_t grid_for(size_t n)
{
    size_t g = (n + 2047) / 2048;
    if (g < 1) g = 1;
    if (g > static_cast<size_t>(kReducePartialsMax)) g = kReducePartialsMax;
    return static_cast<uint32_t>(g);
}

uint32_t ew_grid(size_t n)
{
    size_t g = (n + 1023) / 1024;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return static_cast<uint32_t>(g);
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

// Carves named regions out of one device allocation.
struct Carve
{
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t count)
    {
        T* p = reinterpret_cast<T*>(base + off);
        off += align_up(count * sizeof(T));
        return p;
    }
};

bool graphs_enabled()
{
    const char* e = std::getenv("CLSPARSE_AMD_NO_GRAPH");
    return !(e && e[0] == '1');
}

// Orders the control's solver stream after the caller's stream.
clsparseStatus enter_solver_stream(clsparseControl control)
{
    CLS_HIP_TRY(hipEventRecord(control->event, control->stream));
    CLS_HIP_TRY(hipStreamWaitEvent(control->solver_stream, control->event, 0));
    return clsparseSuccess;
}

// Orders the caller's stream after the solver stream and records the event.
clsparseStatus leave_solver_stream(clsparseControl control)
{
    CLS_HIP_TRY(hipEventRecord(control->event, control->solver_stream));
    CLS_HIP_TRY(hipStreamWaitEvent(control->stream, control->event, 0));
    CLS_HIP_TRY(hipStreamSynchronize(control->solver_stream));
    return clsparseSuccess;
}

}  // namespace
