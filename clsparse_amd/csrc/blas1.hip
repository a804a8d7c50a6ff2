// Dense BLAS-1 for gfx950: deterministic reductions and element-wise kernels.
//
// Reductions (dot, nrm1, nrm2, sum) run in two fixed-shape stages:
//   stage 1: G <= 1024 workgroups, grid-stride over the vector, FP64 (int64
//            for integers) accumulation per lane, workgroup sum -> partial[g];
//   stage 2: one workgroup adds the G partials in a fixed order, applies the
//            square root for nrm2 once, and stores the scalar.
// G depends only on n, so a result is bit-reproducible run to run. This
// replaces the reference's 256x256 inner-product kernel followed by a CAS
// float-atomic final stage whose order varies between runs
// (src/library/kernels/dot.cl L51-88, reduce.cl L66-103,
// atomic_reduce.cl L96-113).
//
// Element-wise kernels read alpha/beta from device memory and apply the
// reference's special cases (src/library/blas1/cldense-axpby.cpp L27-84,
// cldense-axpy.cpp L51-63, cldense-scale.cpp L44-55) on the device, so no
// host read-back is needed.
#include "common.hpp"

namespace {

template <typename T>
struct AccOf
{
    using type = double;
};
template <>
struct AccOf<int32_t>
{
    using type = long long;
};

template <typename T, typename A>
__device__ __forceinline__ A red_term(RedOp op, T a, T b)
{
    switch (op) {
        case RedOp::Sum: return static_cast<A>(a);
        case RedOp::Abs: return static_cast<A>(a < T(0) ? -a : a);
        case RedOp::Sqr: return static_cast<A>(a) * static_cast<A>(a);
        default: return static_cast<A>(a) * static_cast<A>(b);
    }
}

// The grid-stride loop loads U elements (i, i + stride, ...) of each input
// before adding any, so a lane keeps U loads per input in flight; each lane
// still adds its elements in ascending order, so the sums are the same for
// every U. On 2^26 doubles: one input (nrm1, nrm2, reduce) U = 4 at 5.8 TB/s
// against 3.4 one element at a time; two inputs (dot) U = 2 at 6.0 TB/s
// (U = 4: 5.5).
template <typename T, typename A, int U>
__global__ __launch_bounds__(256) void reduce_stage1(RedOp op, const T* __restrict__ x,
                                                     const T* __restrict__ y, size_t n,
                                                     A* __restrict__ partials)
{
    __shared__ A scratch[4];
    A acc = A(0);
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i0 = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i0 < n; i0 += U * stride) {
        T xv[U], yv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + u * stride;
            xv[u] = i < n ? x[i] : T(0);
            yv[u] = (y != nullptr && i < n) ? y[i] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i0 + u * stride < n) acc += red_term<T, A>(op, xv[u], yv[u]);
    }
    acc = block_sum256(acc, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

template <typename A, typename R>
__global__ __launch_bounds__(256) void reduce_stage2(const A* __restrict__ partials,
                                                     uint32_t g, R* __restrict__ result,
                                                     bool take_sqrt)
{
    __shared__ A scratch[4];
    A acc = A(0);
    for (uint32_t i = threadIdx.x; i < g; i += 256) acc += partials[i];
    acc = block_sum256(acc, scratch);
    if (threadIdx.x == 0) {
        if (take_sqrt) acc = static_cast<A>(sqrt(static_cast<double>(acc)));
        *result = static_cast<R>(acc);
    }
}

uint32_t reduce_blocks(size_t n)
{
    size_t g = (n + 2047) / 2048;  // >= 8 elements per lane
    if (g < 1) g = 1;
    if (g > static_cast<size_t>(kReducePartialsMax)) g = kReducePartialsMax;
    return static_cast<uint32_t>(g);
}

enum class Ew { Add, Sub, Mul, Div };

// r[i] = f(i) for i < n, where f reads element i of its inputs through `at`:
// each lane takes 4 consecutive elements per step, read and written with
// 16-byte accesses (two for FP64), so a wave moves 1-2 KB per instruction
// instead of 256-512 B; the n % 4 tail element-wise.
template <typename T>
struct Quad
{
    T e[4];
};

template <typename T>
__device__ __forceinline__ Quad<T> load4(const T* __restrict__ p)
{
    typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));
    u32x4a w[sizeof(T) / 4];
#pragma unroll
    for (int h = 0; h < static_cast<int>(sizeof(T) / 4); ++h) w[h] = reinterpret_cast<const u32x4a*>(p)[h];
    Quad<T> q;
    __builtin_memcpy(&q, w, sizeof(q));
    return q;
}

template <typename T>
__device__ __forceinline__ void store4(T* __restrict__ p, const Quad<T>& q)
{
    typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));
    u32x4a w[sizeof(T) / 4];
    __builtin_memcpy(w, &q, sizeof(q));
#pragma unroll
    for (int h = 0; h < static_cast<int>(sizeof(T) / 4); ++h) reinterpret_cast<u32x4a*>(p)[h] = w[h];
}

// r = f(x, y) element by element (y may be null when f ignores it).
template <typename T, typename F>
__device__ __forceinline__ void ew_map(size_t n, T* __restrict__ r, const T* __restrict__ x,
                                       const T* __restrict__ y, F f)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    const size_t nq = n / 4;
    for (size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; k < nq; k += stride) {
        const Quad<T> a = x ? load4(x + 4 * k) : Quad<T>{};
        const Quad<T> b = y ? load4(y + 4 * k) : Quad<T>{};
        Quad<T> v;
#pragma unroll
        for (int e = 0; e < 4; ++e) v.e[e] = f(a.e[e], b.e[e]);
        store4(r + 4 * k, v);
    }
    const size_t i = 4 * nq + static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (blockIdx.x == 0 && i < n) r[i] = f(x ? x[i] : T(0), y ? y[i] : T(0));
}

template <typename T>
__global__ __launch_bounds__(256) void ew_kernel(Ew op, size_t n, T* __restrict__ r,
                                                 const T* __restrict__ x,
                                                 const T* __restrict__ y)
{
    switch (op) {
        case Ew::Add: ew_map(n, r, x, y, [](T a, T b) { return a + b; }); break;
        case Ew::Sub: ew_map(n, r, x, y, [](T a, T b) { return a - b; }); break;
        case Ew::Mul: ew_map(n, r, x, y, [](T a, T b) { return a * b; }); break;
        default: ew_map(n, r, x, y, [](T a, T b) { return a / b; }); break;
    }
}

// r = alpha * y, alpha == 0 writes zeros (cldense-scale.cpp L44-55). One
// input: element-wise loads were faster here than ew_map's 16-byte ones
// (198-200 against 208 us for 2^26 doubles).
template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(size_t n, T* __restrict__ r,
                                                    const T* __restrict__ pAlpha,
                                                    const T* __restrict__ y)
{
    const T alpha = *pAlpha;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
        r[i] = (alpha == T(0)) ? T(0) : y[i] * alpha;
}

// r = alpha * x + y; alpha == 0 copies y (cldense-axpy.cpp L51-63).
template <typename T>
__global__ __launch_bounds__(256) void axpy_kernel(size_t n, T* __restrict__ r,
                                                   const T* __restrict__ pAlpha,
                                                   const T* __restrict__ x,
                                                   const T* __restrict__ y)
{
    const T alpha = *pAlpha;
    if (alpha == T(0))
        ew_map(n, r, y, static_cast<const T*>(nullptr), [](T b, T) { return b; });
    else
        ew_map(n, r, x, y, [alpha](T a, T b) { return b + alpha * a; });
}

// r = alpha * x + beta * y with the special cases of cldense-axpby.cpp L27-84:
// both zero -> 0; alpha 0 -> beta*y; beta 0 -> alpha*x; alpha 1 -> x + beta*y;
// beta 1 -> alpha*x + y.
template <typename T>
__global__ __launch_bounds__(256) void axpby_kernel(size_t n, T* __restrict__ r,
                                                    const T* __restrict__ pAlpha,
                                                    const T* __restrict__ x,
                                                    const T* __restrict__ pBeta,
                                                    const T* __restrict__ y)
{
    const T alpha = *pAlpha;
    const T beta = *pBeta;
    const T* none = nullptr;
    if (alpha == T(0)) {
        if (beta == T(0))
            ew_map(n, r, none, none, [](T, T) { return T(0); });
        else
            ew_map(n, r, none, y, [beta](T, T b) { return b * beta; });
    } else if (beta == T(0)) {
        ew_map(n, r, x, none, [alpha](T a, T) { return a * alpha; });
    } else if (alpha == T(1)) {
        ew_map(n, r, x, y, [beta](T a, T b) { return a + beta * b; });
    } else if (beta == T(1)) {
        ew_map(n, r, x, y, [alpha](T a, T b) { return alpha * a + b; });
    } else {
        ew_map(n, r, x, y, [alpha, beta](T a, T b) { return alpha * a + beta * b; });
    }
}

uint32_t ew_blocks(size_t n)
{
    size_t g = (n + 1023) / 1024;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return static_cast<uint32_t>(g);
}

template <typename T>
clsparseStatus reduce_entry(RedOp op, clsparseScalar* s, const cldenseVector* x,
                            const cldenseVector* y, bool take_sqrt, clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (s == nullptr || x == nullptr || s->value == nullptr) return clsparseInvalidValue;
    size_t n = x->num_values;
    const T* yp = nullptr;
    if (y != nullptr) {
        if (y->num_values < n) n = y->num_values;
        yp = dev_ptr<T>(y->values, y->off_values);
    }
    if (n > 0 && (x->values == nullptr || (y != nullptr && y->values == nullptr)))
        return clsparseInvalidMemObject;
    CLS_TRY(launch_reduce<T, T>(op, dev_ptr<T>(x->values, x->off_values), yp, n,
                                dev_ptr<T>(s->value, s->off_value), take_sqrt,
                                control->red_partials, control->stream));
    return finish_op(control);
}

template <typename T>
clsparseStatus ew_entry(Ew op, cldenseVector* r, const cldenseVector* x,
                        const cldenseVector* y, clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (r == nullptr || x == nullptr || y == nullptr) return clsparseInvalidValue;
    const size_t n = r->num_values;
    if (n == 0) return clsparseSuccess;
    if (x->num_values < n || y->num_values < n) return clsparseInvalidSize;
    hipLaunchKernelGGL((ew_kernel<T>), dim3(ew_blocks(n)), dim3(256), 0, control->stream, op, n,
                       dev_ptr<T>(r->values, r->off_values), dev_ptr<T>(x->values, x->off_values),
                       dev_ptr<T>(y->values, y->off_values));
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}

template <typename T>
clsparseStatus scale_entry(cldenseVector* r, const clsparseScalar* alpha, const cldenseVector* y,
                           clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (r == nullptr || alpha == nullptr || y == nullptr) return clsparseInvalidValue;
    const size_t n = r->num_values;
    if (n == 0) return clsparseSuccess;
    if (y->num_values < n) return clsparseInvalidSize;
    hipLaunchKernelGGL((scale_kernel<T>), dim3(ew_blocks(n)), dim3(256), 0, control->stream, n,
                       dev_ptr<T>(r->values, r->off_values),
                       dev_ptr<T>(alpha->value, alpha->off_value),
                       dev_ptr<T>(y->values, y->off_values));
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}

template <typename T>
clsparseStatus axpy_entry(cldenseVector* r, const clsparseScalar* alpha, const cldenseVector* x,
                          const cldenseVector* y, clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (r == nullptr || alpha == nullptr || x == nullptr || y == nullptr)
        return clsparseInvalidValue;
    const size_t n = r->num_values;
    if (n == 0) return clsparseSuccess;
    if (x->num_values < n || y->num_values < n) return clsparseInvalidSize;
    hipLaunchKernelGGL((axpy_kernel<T>), dim3(ew_blocks(n)), dim3(256), 0, control->stream, n,
                       dev_ptr<T>(r->values, r->off_values),
                       dev_ptr<T>(alpha->value, alpha->off_value),
                       dev_ptr<T>(x->values, x->off_values), dev_ptr<T>(y->values, y->off_values));
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}

template <typename T>
clsparseStatus axpby_entry(cldenseVector* r, const clsparseScalar* alpha, const cldenseVector* x,
                           const clsparseScalar* beta, const cldenseVector* y,
                           clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (r == nullptr || alpha == nullptr || x == nullptr || beta == nullptr || y == nullptr)
        return clsparseInvalidValue;
    const size_t n = r->num_values;
    if (n == 0) return clsparseSuccess;
    if (x->num_values < n || y->num_values < n) return clsparseInvalidSize;
    hipLaunchKernelGGL((axpby_kernel<T>), dim3(ew_blocks(n)), dim3(256), 0, control->stream, n,
                       dev_ptr<T>(r->values, r->off_values),
                       dev_ptr<T>(alpha->value, alpha->off_value),
                       dev_ptr<T>(x->values, x->off_values),
                       dev_ptr<T>(beta->value, beta->off_value),
                       dev_ptr<T>(y->values, y->off_values));
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}

}  // namespace

template <typename T, typename R>
clsparseStatus launch_reduce(RedOp op, const T* x, const T* y, size_t n, R* result,
                             bool sqrt_result, double* partials_buf, hipStream_t stream)
{
    using A = typename AccOf<T>::type;
    const uint32_t g = reduce_blocks(n);
    A* partials = reinterpret_cast<A*>(partials_buf);
    if (y == nullptr)
        hipLaunchKernelGGL((reduce_stage1<T, A, 4>), dim3(g), dim3(256), 0, stream, op, x, y, n, partials);
    else
        hipLaunchKernelGGL((reduce_stage1<T, A, 2>), dim3(g), dim3(256), 0, stream, op, x, y, n, partials);
    CLS_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((reduce_stage2<A, R>), dim3(1), dim3(256), 0, stream, partials, g, result,
                       sqrt_result);
    CLS_HIP_TRY(hipGetLastError());
    return clsparseSuccess;
}

template clsparseStatus launch_reduce<float, float>(RedOp, const float*, const float*, size_t,
                                                    float*, bool, double*, hipStream_t);
template clsparseStatus launch_reduce<double, double>(RedOp, const double*, const double*, size_t,
                                                      double*, bool, double*, hipStream_t);
template clsparseStatus launch_reduce<float, double>(RedOp, const float*, const float*, size_t,
                                                     double*, bool, double*, hipStream_t);

// ---- C ABI -------------------------------------------------------------------
// src/library/blas1/cldense-scale.cpp L20-96
extern "C" clsparseStatus cldenseSscale(cldenseVector* r, const clsparseScalar* alpha,
                                        const cldenseVector* y, const clsparseControl control)
{
    return scale_entry<float>(r, alpha, y, control);
}
extern "C" clsparseStatus cldenseDscale(cldenseVector* r, const clsparseScalar* alpha,
                                        const cldenseVector* y, const clsparseControl control)
{
    return scale_entry<double>(r, alpha, y, control);
}
// src/library/blas1/cldense-axpy.cpp L23-111
extern "C" clsparseStatus cldenseSaxpy(cldenseVector* r, const clsparseScalar* alpha,
                                       const cldenseVector* x, const cldenseVector* y,
                                       const clsparseControl control)
{
    return axpy_entry<float>(r, alpha, x, y, control);
}
extern "C" clsparseStatus cldenseDaxpy(cldenseVector* r, const clsparseScalar* alpha,
                                       const cldenseVector* x, const cldenseVector* y,
                                       const clsparseControl control)
{
    return axpy_entry<double>(r, alpha, x, y, control);
}
// src/library/blas1/cldense-axpby.cpp L88-152
extern "C" clsparseStatus cldenseSaxpby(cldenseVector* r, const clsparseScalar* alpha,
                                        const cldenseVector* x, const clsparseScalar* beta,
                                        const cldenseVector* y, const clsparseControl control)
{
    return axpby_entry<float>(r, alpha, x, beta, y, control);
}
extern "C" clsparseStatus cldenseDaxpby(cldenseVector* r, const clsparseScalar* alpha,
                                        const cldenseVector* x, const clsparseScalar* beta,
                                        const cldenseVector* y, const clsparseControl control)
{
    return axpby_entry<double>(r, alpha, x, beta, y, control);
}
// src/library/blas1/cldense-reduce.cpp L23-95 (integer sum wraps modulo 2^32)
extern "C" clsparseStatus cldenseIreduce(clsparseScalar* s, const cldenseVector* x,
                                         const clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (s == nullptr || x == nullptr || s->value == nullptr) return clsparseInvalidValue;
    const size_t n = x->num_values;
    if (n > 0 && x->values == nullptr) return clsparseInvalidMemObject;
    const uint32_t g = reduce_blocks(n);
    long long* partials = reinterpret_cast<long long*>(control->red_partials);
    hipLaunchKernelGGL((reduce_stage1<int32_t, long long, 4>), dim3(g), dim3(256), 0,
                       control->stream, RedOp::Sum, dev_ptr<int32_t>(x->values, x->off_values),
                       static_cast<const int32_t*>(nullptr), n, partials);
    CLS_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((reduce_stage2<long long, int32_t>), dim3(1), dim3(256), 0,
                       control->stream, partials, g, dev_ptr<int32_t>(s->value, s->off_value),
                       false);
    CLS_HIP_TRY(hipGetLastError());
    return finish_op(control);
}
extern "C" clsparseStatus cldenseSreduce(clsparseScalar* s, const cldenseVector* x,
                                         const clsparseControl control)
{
    return reduce_entry<float>(RedOp::Sum, s, x, nullptr, false, control);
}
extern "C" clsparseStatus cldenseDreduce(clsparseScalar* s, const cldenseVector* x,
                                         const clsparseControl control)
{
    return reduce_entry<double>(RedOp::Sum, s, x, nullptr, false, control);
}
// src/library/blas1/cldense-nrm1.cpp L20-62
extern "C" clsparseStatus cldenseSnrm1(clsparseScalar* s, const cldenseVector* x,
                                       const clsparseControl control)
{
    return reduce_entry<float>(RedOp::Abs, s, x, nullptr, false, control);
}
extern "C" clsparseStatus cldenseDnrm1(clsparseScalar* s, const cldenseVector* x,
                                       const clsparseControl control)
{
    return reduce_entry<double>(RedOp::Abs, s, x, nullptr, false, control);
}
// src/library/blas1/cldense-nrm2.cpp (sqrt applied once, after the sum)
extern "C" clsparseStatus cldenseSnrm2(clsparseScalar* s, const cldenseVector* x,
                                       const clsparseControl control)
{
    return reduce_entry<float>(RedOp::Sqr, s, x, nullptr, true, control);
}
extern "C" clsparseStatus cldenseDnrm2(clsparseScalar* s, const cldenseVector* x,
                                       const clsparseControl control)
{
    return reduce_entry<double>(RedOp::Sqr, s, x, nullptr, true, control);
}
// src/library/blas1/cldense-dot.cpp L20-70
extern "C" clsparseStatus cldenseSdot(clsparseScalar* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return reduce_entry<float>(RedOp::Dot, r, x, y, false, control);
}
extern "C" clsparseStatus cldenseDdot(clsparseScalar* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return reduce_entry<double>(RedOp::Dot, r, x, y, false, control);
}
// src/library/blas1/elementwise-transform.cpp L20-180
extern "C" clsparseStatus cldenseSadd(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<float>(Ew::Add, r, x, y, control);
}
extern "C" clsparseStatus cldenseDadd(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<double>(Ew::Add, r, x, y, control);
}
extern "C" clsparseStatus cldenseSsub(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<float>(Ew::Sub, r, x, y, control);
}
extern "C" clsparseStatus cldenseDsub(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<double>(Ew::Sub, r, x, y, control);
}
extern "C" clsparseStatus cldenseSmul(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<float>(Ew::Mul, r, x, y, control);
}
extern "C" clsparseStatus cldenseDmul(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<double>(Ew::Mul, r, x, y, control);
}
extern "C" clsparseStatus cldenseSdiv(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<float>(Ew::Div, r, x, y, control);
}
extern "C" clsparseStatus cldenseDdiv(cldenseVector* r, const cldenseVector* x,
                                      const cldenseVector* y, const clsparseControl control)
{
    return ew_entry<double>(Ew::Div, r, x, y, control);
}
