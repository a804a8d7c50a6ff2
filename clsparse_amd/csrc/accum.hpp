// Per-lane accumulators for the SpMV kernels.
//
// Acc<T, false>: plain FMA accumulation (the default precision mode).
// Acc<T, true>:  compensated accumulation for clsparseEnableExtendedPrecision:
//   error-free TwoProduct (fma residual) and branch-free TwoSum (Knuth), with
//   the running error kept in a second register -- SumK with K = 2 of Ogita,
//   Rump and Oishi, "Accurate sum and dot product", SISC 26(6), 2005. The
//   reference's extended mode (src/library/kernels/csrmv_adaptive.cl L167-308)
//   drops the product error; this one keeps it.
#pragma once

#include "common.hpp"

template <typename T, bool EXT>
struct Acc;

template <typename T>
struct Acc<T, false>
{
    T s = T(0);
    __device__ __forceinline__ void add(T v) { s += v; }
    __device__ __forceinline__ void mul_add(T a, T b) { s = fma(a, b, s); }
    __device__ __forceinline__ void merge(const Acc& o) { s += o.s; }
    __device__ __forceinline__ T value() const { return s; }
    template <int CTRL>
    __device__ __forceinline__ Acc dpp_move() const
    {
        Acc r;
        r.s = dpp<CTRL>(s);
        return r;
    }
    __device__ __forceinline__ Acc xor_move(int m) const
    {
        Acc r;
        r.s = shfl_xor_t(s, m);
        return r;
    }
    // y_out = beta * y_old + sum (fma, as csrmv_adaptive.cl L481-483)
    __device__ __forceinline__ T with_beta(T beta, T y_old) const { return fma(beta, y_old, s); }
};

// FP32 values: the sum is kept in FP64 (a product of two FP32 values is exact
// in FP64), so a row's result carries one FP32 rounding of an FP64 sum: per-row
// error <= 2^-24 sum|a_ij x_j| + 2^-53 (len) sum|a_ij x_j|, within 1e-6 for
// any row length, where FP32 accumulation is bounded only by ~len 2^-24.
template <>
struct Acc<float, false>
{
    double s = 0.0;
    __device__ __forceinline__ void add(float v) { s += static_cast<double>(v); }
    __device__ __forceinline__ void mul_add(float a, float b)
    {
        s = fma(static_cast<double>(a), static_cast<double>(b), s);
    }
    __device__ __forceinline__ void merge(const Acc& o) { s += o.s; }
    __device__ __forceinline__ float value() const { return static_cast<float>(s); }
    template <int CTRL>
    __device__ __forceinline__ Acc dpp_move() const
    {
        Acc r;
        r.s = dpp<CTRL>(s);
        return r;
    }
    __device__ __forceinline__ Acc xor_move(int m) const
    {
        Acc r;
        r.s = shfl_xor_t(s, m);
        return r;
    }
    __device__ __forceinline__ float with_beta(float beta, float y_old) const
    {
        return static_cast<float>(fma(static_cast<double>(beta), static_cast<double>(y_old), s));
    }
};

template <typename T>
struct Acc<T, true>
{
    T s = T(0);
    T c = T(0);
    // Error-free transformations: no FMA contraction inside them.
    __device__ __forceinline__ void add(T v)
    {
#pragma clang fp contract(off)
        const T t = s + v;
        const T bp = t - s;
        c += (s - (t - bp)) + (v - bp);
        s = t;
    }
    __device__ __forceinline__ void mul_add(T a, T b)
    {
#pragma clang fp contract(off)
        const T p = a * b;
        const T e = fma(a, b, -p);
        add(p);
        c += e;
    }
    __device__ __forceinline__ void merge(const Acc& o)
    {
#pragma clang fp contract(off)
        add(o.s);
        c += o.c;
    }
    __device__ __forceinline__ T value() const { return s + c; }
    template <int CTRL>
    __device__ __forceinline__ Acc dpp_move() const
    {
        Acc r;
        r.s = dpp<CTRL>(s);
        r.c = dpp<CTRL>(c);
        return r;
    }
    __device__ __forceinline__ Acc xor_move(int m) const
    {
        Acc r;
        r.s = shfl_xor_t(s, m);
        r.c = shfl_xor_t(c, m);
        return r;
    }
    __device__ __forceinline__ T with_beta(T beta, T y_old) const
    {
        Acc t = *this;
        t.mul_add(beta, y_old);
        return t.value();
    }
};

// Butterfly reduction over aligned groups of G lanes (G = 1..64, power of
// two): DPP inside a row of 16 lanes, cross-lane permutes above that.
template <typename A>
__device__ __forceinline__ A group_reduce(A a, unsigned G)
{
    if (G >= 2) a.merge(a.template dpp_move<kDppQuadSwap1>());
    if (G >= 4) a.merge(a.template dpp_move<kDppQuadSwap2>());
    if (G >= 8) a.merge(a.template dpp_move<kDppRowHalfMirror>());
    if (G >= 16) a.merge(a.template dpp_move<kDppRowMirror>());
    if (G >= 32) a.merge(a.xor_move(16));
    if (G >= 64) a.merge(a.xor_move(32));
    return a;
}

template <typename A>
__device__ __forceinline__ A wave_reduce(A a)
{
    a.merge(a.template dpp_move<kDppQuadSwap1>());
    a.merge(a.template dpp_move<kDppQuadSwap2>());
    a.merge(a.template dpp_move<kDppRowHalfMirror>());
    a.merge(a.template dpp_move<kDppRowMirror>());
    a.merge(a.xor_move(16));
    a.merge(a.xor_move(32));
    return a;
}
