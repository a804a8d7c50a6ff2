// Learned factorized sparse approximate inverse (no reference counterpart).
//
// G is lower triangular on the pattern of tril(A) and parametrised as
//     G_ij = (delta_ij + o_ij) / sqrt(a_ii)
// where o_ij is the output of a per-nonzero MLP (16 features -> 64 -> 64 -> 1,
// ReLU, BF16 weights and activations on MFMA with FP32 accumulation). A zero
// network output is exactly the Jacobi preconditioner. The preconditioner is
// M = G G^T (z = G (G^T r): two CSR-Adaptive SpMVs with the CSR of G and of
// G^T), trained so that G^T A G ~ I with the Hutchinson estimate
//     L = 1/(k n) sum_z ||(G^T A G - I) z||^2,   z Rademacher,
// whose gradient w.r.t. G_ij is 2/(k n) sum_z (u_i v_j + w_i z_j) with
// t = G z, u = A t, v = G^T u - z, w = A G v.
//
// MLP kernels compute every layer transposed, Z^T = W H^T (hidden x nnz):
// the 16x16 BF16 MFMAs write their result in exactly the lane layout they
// read their B operand in, so each layer's activations feed the next MFMA
// straight from registers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "dia.hpp"

namespace {

constexpr int kFeat = 16;
constexpr int kHid = 64;
// Parameter vector layout (FP32 master copy).
constexpr int kOffW1 = 0;                        // W1[64][16]
constexpr int kOffB1 = kOffW1 + kHid * kFeat;    // b1[64]
constexpr int kOffW2 = kOffB1 + kHid;            // W2[64][64] (row = output unit)
constexpr int kOffB2 = kOffW2 + kHid * kHid;     // b2[64]
constexpr int kOffW3 = kOffB2 + kHid;            // w3[64]
constexpr int kOffB3 = kOffW3 + kHid;            // b3
constexpr int kNumParams = kOffB3 + 1;           // 5313
constexpr int kMlpWaves = 4;
constexpr int kMlpGrid = 512;                    // workgroups of the backward kernel

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
// FP32 -> BF16, round to nearest even, in integer arithmetic. (A native
// __bf16 conversion of a float that was just rounded from a double may be
// compiled as one rounding of the double, which differs from the two the
// features are defined with.)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
// Two network activations (FP32 MFMA results) -> BF16 with gfx950's
// v_cvt_pk_bf16_f32: the same rounding as f2bf for every non-NaN float
// (checked bit for bit on 4 M values: ties, denormals, random bit patterns),
// one instruction for two values where f2bf takes four per value -- the
// activation code is VALU-bound.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi)
{
    const bf16x2_t v = __builtin_convertvector(f32x2_t{lo, hi}, bf16x2_t);
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf2f(uint16_t h)
{
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 load_bf16x4(const uint16_t* p)
{
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = static_cast<short>(v.x & 0xFFFFu);
    r[1] = static_cast<short>(v.x >> 16);
    r[2] = static_cast<short>(v.y & 0xFFFFu);
    r[3] = static_cast<short>(v.y >> 16);
    return r;
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bf16x4 pack_bf16x4(f32x4 v)
{
    return __builtin_bit_cast(bf16x4, u32x2{f2bf2(v[0], v[1]), f2bf2(v[2], v[3])});
}

// ReLU of a pair of BF16 values: a BF16 value is negative exactly when its
// bits, read as an int16, are negative, so the 16-bit integer max with 0 is
// the ReLU (-0 becomes +0). Rounding keeps the sign, so relu(round(a)) =
// round(relu(a)).
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t p)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, p), i16x2{0, 0}));
}

// round_bf16(relu(v)) for 4 activations.
__device__ __forceinline__ bf16x4 pack_relu_bf16x4(f32x4 v)
{
    return __builtin_bit_cast(bf16x4, u32x2{relu_bf16x2(f2bf2(v[0], v[1])), relu_bf16x2(f2bf2(v[2], v[3]))});
}

// Weights as the MFMA reads them (BF16). W1: [64][16]; W2: [64][64] row =
// output unit; W2T: its transpose (row = input unit) for the backward pass.
struct MlpDev
{
    const uint16_t* W1;
    const uint16_t* W2;
    const uint16_t* W2T;
    const float* p;  // FP32 master parameters (biases, w3, b3 are read from here)
};

// The 16x16x32 BF16 MFMA: lane group q = l/16 holds 8 K values of row (A) or
// column (B) l%16. Which 8 does not matter as long as A and B agree, so a K=32
// operand is two K=16 fragments side by side (units of blocks 2j and 2j+1).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

__device__ __forceinline__ s16x8 cat8(bf16x4 lo, bf16x4 hi)
{
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Piece `which` (0, 1, 2) of the exact split v = hi + mid + lo into BF16 values:
// each piece takes the next 8 significant bits (truncation), so both
// remainders are exact in FP32 and the last one fits a BF16 exactly.
__device__ __forceinline__ short bf16_piece(float v, uint32_t which)
{
    const float hi = __uint_as_float(__float_as_uint(v) & 0xFFFF0000u);
    const float r = v - hi;
    const float mid = __uint_as_float(__float_as_uint(r) & 0xFFFF0000u);
    const float lo = r - mid;
    const float piece = which == 0 ? hi : (which == 1 ? mid : lo);
    return static_cast<short>(__float_as_uint(piece) >> 16);
}

// Per-lane register copy of the network for the forward kernel, loaded once
// per wave. Lane l = 16q + c: the MFMA A fragments of W1 (rows 16m + c,
// features 4q..4q+3) and W2 (rows 16m + c, K pair j), the output weights as
// an A operand whose rows 0, 1, 2 hold the three BF16 pieces of w3 (so the MFMA
// forms w3 . h2 with w3 exact), and b1, b2 in the accumulator layout (units
// 16m + 4q + e), the MFMA chains' initial values.
struct NetRegs
{
    bf16x4 w1[4];
    s16x8 w2[4][2];  // [m][j]
    s16x8 w3[2];     // [j]
    f32x4 b1[4], b2[4];
    float b3;
};

__device__ __forceinline__ void load_net(const MlpDev& net, uint32_t lane, NetRegs& w)
{
    const uint32_t q = lane >> 4, c16 = lane & 15;
    for (int m = 0; m < 4; ++m) {
        w.w1[m] = load_bf16x4(net.W1 + (16 * m + c16) * kFeat + 4 * q);
        for (int j = 0; j < 2; ++j)
            w.w2[m][j] = cat8(load_bf16x4(net.W2 + (16 * m + c16) * kHid + 16 * (2 * j) + 4 * q),
                              load_bf16x4(net.W2 + (16 * m + c16) * kHid + 16 * (2 * j + 1) + 4 * q));
        for (int e = 0; e < 4; ++e) {
            w.b1[m][e] = net.p[kOffB1 + 16 * m + 4 * q + e];
            w.b2[m][e] = net.p[kOffB2 + 16 * m + 4 * q + e];
        }
    }
    for (int j = 0; j < 2; ++j)
        for (int t = 0; t < 8; ++t) {
            const int u = 16 * (2 * j + t / 4) + 4 * q + t % 4;
            w.w3[j][t] = c16 < 3 ? bf16_piece(net.p[kOffW3 + u], c16) : short(0);
        }
    w.b3 = net.p[kOffB3];
}

// Features of nonzeros r0 .. r0+15 as the layer-1 B operand: lane l holds
// features 4(l/16)..+3 of nonzero r0 + l%16 (zeros past nnz).
// Features of nonzeros r0 .. r0+15 as the layer-1 B operand: lane l holds
// features 4(l/16)..+3 of nonzero r0 + l%16. A buffer load over the tile's
// own rows: past nnz the hardware's range check reads zeros, so the load has
// no branch (and the loop's wait for it can stay a tile behind).
__device__ __forceinline__ bf16x4 load_tile_x(const uint16_t* __restrict__ X, uint32_t r0, uint32_t nnz, uint32_t lane)
{
    const uint32_t rows = std::min(nnz - r0, 16u);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t*>(X + static_cast<size_t>(r0) * kFeat), 0, rows * kFeat * 2, 0x00020000);
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (lane & 15) * kFeat * 2 + (lane >> 4) * 8, 0, 0);
    return __builtin_bit_cast(bf16x4, v);
}

// out[r0 + l] = o for lanes l < 16 and r0 + l < nnz (a buffer store: the
// other lanes' offsets are past the tile's range and the hardware drops them).
__device__ __forceinline__ void store_tile_out(float* __restrict__ out, uint32_t r0, uint32_t nnz, uint32_t lane,
                                               float o)
{
    const uint32_t rows = std::min(nnz - r0, 16u);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(out + r0, 0, rows * 4, 0x00020000);
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(o), rs, lane * 4, 0, 0);
}

// One tile of 16 nonzeros, every layer transposed (hidden x nnz): lane l works
// on nonzero l % 16. Returns o = w3 . round_bf16(relu(a2)) + b3 in lanes 0..15.
__device__ __forceinline__ float tile_forward(const NetRegs& w, bf16x4 x)
{
    // Layer 1: A1^T (64 x 16nnz) = W1 (64 x 16) . X^T (16 x 16nnz) + b1
    bf16x4 h1[4];
    for (int m = 0; m < 4; ++m) h1[m] = pack_relu_bf16x4(mfma16(w.w1[m], x, w.b1[m]));
    const s16x8 h1p[2] = {cat8(h1[0], h1[1]), cat8(h1[2], h1[3])};
    // Layer 2: A2^T (64 x 16) = W2 (64 x 64) . H1^T (64 x 16) + b2
    bf16x4 h2[4];
    for (int m = 0; m < 4; ++m) {
        f32x4 c = w.b2[m];
        for (int j = 0; j < 2; ++j) c = mfma32(w.w2[m][j], h1p[j], c);
        h2[m] = pack_relu_bf16x4(c);
    }
    // Output: rows 0, 1, 2 of W3pieces . H2^T; b3 rides in row 0's initial value.
    f32x4 s = {w.b3, 0.f, 0.f, 0.f};
    s = mfma32(w.w3[0], cat8(h2[0], h2[1]), s);
    s = mfma32(w.w3[1], cat8(h2[2], h2[3]), s);
    return (s[0] + s[1]) + s[2];
}

// Persistent waves (the grid is what the GPU holds at once): each wave loads
// the network once and walks tiles with the next tile's features in flight.
__global__ __launch_bounds__(64 * kMlpWaves) void mlp_forward_kernel(MlpDev net,
                                                                      const uint16_t* __restrict__ X,
                                                                      uint32_t nnz,
                                                                      float* __restrict__ out)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar loop control
    const uint32_t tiles = (nnz + 15) / 16;
    const uint32_t stride = gridDim.x * kMlpWaves;
    uint32_t tile = blockIdx.x * kMlpWaves + wave;
    if (tile >= tiles) return;
    NetRegs w;
    load_net(net, lane, w);
    // Two feature buffers in turn, each refilled (from a clamped tile index,
    // so the load is always issued) right after its tile has used it: the
    // loads run two tiles ahead.
    bf16x4 xa = load_tile_x(X, tile * 16, nnz, lane);
    bf16x4 xb = load_tile_x(X, std::min(tile + stride, tiles - 1) * 16, nnz, lane);
    while (true) {
        float o = tile_forward(w, xa);
        xa = load_tile_x(X, std::min(tile + 2 * stride, tiles - 1) * 16, nnz, lane);
        store_tile_out(out, tile * 16, nnz, lane, o);
        tile += stride;
        if (tile >= tiles) break;
        o = tile_forward(w, xb);
        xb = load_tile_x(X, std::min(tile + 2 * stride, tiles - 1) * 16, nnz, lane);
        store_tile_out(out, tile * 16, nnz, lane, o);
        tile += stride;
        if (tile >= tiles) break;
    }
}

// Backward. Two lane layouts of a 16-nonzero tile: the forward kernel's
// unit x nnz layout ("T": lane (q, c) holds units 16m + 4q + e of nonzero c)
// and the nnz x unit layout ("N": lane (q, c) holds nonzeros 4q + e of unit
// 16n + c). A T-layout fragment read as an MFMA A operand and multiplied by the
// 16x16 identity comes out in N layout, and an N-layout one in T layout, each
// value exactly (one product by 1 plus zeros). So per tile:
//   H1^T = relu(W1 X^T + b1)           T    (forward kernel's layer 1)
//   H1   = transpose(H1^T)             N
//   A2   = H1 W2^T + b2                N    A = H1^T in T, B = W2 fragments
//   D2   = relu'(A2) g w3 (FP32)       N    -> db2; BF16 -> dW2 operand
//   D2^T = transpose(D2)               T
//   D1   = relu'(A1) (D2 W2)           N    A = D2 in T, B = W2^T fragments
//   dW2 += D2^T H1                     A = D2 in N, B = H1 in N
//   dW1 += D1^T X                      A = D1 in N, B = X in N (loaded so)
// relu'(A1) is read off H1 (a nonzero BF16 activation). In N each lane holds
// one unit per block for 4 nonzeros, so the bias-like sums (db1, db2, dw3) take
// 4 accumulators per lane, reduced over the 4 lane groups at the end. The
// products over the 64 hidden units run on the K=32 MFMA; the biases are the
// MFMA chains' initial values. The W2 and W2^T fragments are read from LDS for
// every tile, so a wave holds the 80 gradient accumulators, the next tile's
// inputs and the tile's activations in registers for two waves per SIMD. The
// 4 waves of a workgroup add their gradients into one LDS vector in wave
// order, and each workgroup writes one partial gradient vector (reduced later
// in a fixed order).
__global__ __launch_bounds__(64 * kMlpWaves) __attribute__((amdgpu_waves_per_eu(2, 2))) void mlp_backward_kernel(MlpDev net,
                                                                       const uint16_t* __restrict__ X,
                                                                       const float* __restrict__ gout,
                                                                       uint32_t nnz,
                                                                       float* __restrict__ partials)
{
    __shared__ float sh_b1[kHid];         // b1 for the T-layout reads
    __shared__ f32x4 sh_b2[kHid];         // b2, each value 4 times (N-layout initial values)
    __shared__ s16x8 sh_w2[4][2][64];     // W2 fragments [block][K pair][lane]
    __shared__ s16x8 sh_w2t[4][2][64];    // W2^T fragments
    __shared__ float sh_g[kNumParams];    // the workgroup's gradient
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar loop control
    const uint32_t q = lane >> 4, c16 = lane & 15;

    for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(kHid); i += 64 * kMlpWaves) {
        sh_b1[i] = net.p[kOffB1 + i];
        const float b2 = net.p[kOffB2 + i];
        sh_b2[i] = f32x4{b2, b2, b2, b2};
    }
    // The W2 / W2^T fragments of K pair j (blocks 2j, 2j+1) in lane order, so
    // each read is one conflict-free 16-byte load per lane.
    for (uint32_t i = threadIdx.x; i < 4 * 2 * 64; i += 64 * kMlpWaves) {
        const uint32_t m = i / 128, j = (i / 64) % 2, l = i % 64;
        const uint32_t row = 16 * m + (l & 15), col = 32 * j + 4 * (l >> 4);
        sh_w2[m][j][l] = cat8(load_bf16x4(net.W2 + row * kHid + col), load_bf16x4(net.W2 + row * kHid + col + 16));
        sh_w2t[m][j][l] = cat8(load_bf16x4(net.W2T + row * kHid + col), load_bf16x4(net.W2T + row * kHid + col + 16));
    }
    bf16x4 w1[4];
    float w3n[4];  // unit 16n + c16
    for (int m = 0; m < 4; ++m) {
        w1[m] = load_bf16x4(net.W1 + (16 * m + c16) * kFeat + 4 * q);
        w3n[m] = net.p[kOffW3 + 16 * m + c16];
    }
    // The identity as a B operand: lane (q, c) holds rows 4q..4q+3 of column c.
    bf16x4 ident;
    for (int e = 0; e < 4; ++e) ident[e] = (4 * q + e == c16) ? short(0x3F80) : short(0);
    const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
    __syncthreads();

    f32x4 dW2[4][4];  // block (m = output unit, n = input unit)
    f32x4 dW1[4];     // block m (rows = hidden), 16 features
    float db1[4], db2[4], dw3[4];  // per lane: unit 16n + c16
    float db3 = 0.f;
    for (int m = 0; m < 4; ++m) {
        dW1[m] = zero4;
        for (int n = 0; n < 4; ++n) dW2[m][n] = zero4;
        db1[m] = db2[m] = dw3[m] = 0.f;
    }

    // A tile's inputs, loaded one tile ahead: X as the T layout's B operand
    // (features 4q..4q+3 of nonzero c16) and as the dW1 B operand (feature c16
    // of nonzeros 4q + e), gout in both layouts.
    struct TileIn
    {
        bf16x4 x, xn;
        float gN[4];
        float gT;
    };
    // Buffer loads over the tile's own rows: past nnz the hardware's range
    // check reads zeros (zero features and zero gout add nothing), so the
    // loads have no branch and the loop's waits for them stay a tile behind.
    auto load_in = [&](uint32_t tile, TileIn& t) {
        const uint32_t r0 = tile * 16;
        const uint32_t rows = std::min(nnz - r0, 16u);
        const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint16_t*>(X + static_cast<size_t>(r0) * kFeat), 0, rows * kFeat * 2, 0x00020000);
        const __amdgpu_buffer_rsrc_t rg =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(gout + r0), 0, rows * 4, 0x00020000);
        t.x = __builtin_bit_cast(bf16x4, __builtin_amdgcn_raw_buffer_load_b64(rx, c16 * kFeat * 2 + q * 8, 0, 0));
        for (int e = 0; e < 4; ++e) {
            t.xn[e] = static_cast<short>(__builtin_amdgcn_raw_buffer_load_b16(rx, (4 * q + e) * kFeat * 2 + c16 * 2, 0, 0));
            t.gN[e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rg, (4 * q + e) * 4, 0, 0));
        }
        t.gT = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rg, c16 * 4, 0, 0));
    };
    const uint32_t tiles = (nnz + 15) / 16;
    const uint32_t stride = gridDim.x * kMlpWaves;
    auto process = [&](const TileIn& in) {
        // Keeps the W2 fragment reads with the tile (hoisted out of the loop,
        // they would hold 64 registers across it).
        asm volatile("" ::: "memory");
        const float* gN = in.gN;
        if (q == 0) db3 += in.gT;

        // H1^T (T), then H1 (N) by the identity product.
        bf16x4 h1T[4], h1n[4];
        for (int m = 0; m < 4; ++m)
            h1T[m] = pack_relu_bf16x4(mfma16(w1[m], in.x, *reinterpret_cast<const f32x4*>(&sh_b1[16 * m + 4 * q])));
        for (int n = 0; n < 4; ++n) h1n[n] = pack_bf16x4(mfma16(h1T[n], ident, zero4));
        const s16x8 h1Tp[2] = {cat8(h1T[0], h1T[1]), cat8(h1T[2], h1T[3])};
        // A2 and D2 in N; db2, dw3.
        bf16x4 d2n[4];
        for (int n = 0; n < 4; ++n) {
            f32x4 a = sh_b2[16 * n + c16];
            for (int j = 0; j < 2; ++j) a = mfma32(h1Tp[j], sh_w2[n][j][lane], a);
            f32x4 d;
            for (int e = 0; e < 4; ++e) {
                d[e] = a[e] > 0.f ? gN[e] * w3n[n] : 0.f;
                db2[n] += d[e];
            }
            d2n[n] = pack_bf16x4(d);
            // dw3 += g * round_bf16(relu(a2))
            const uint32_t lo = relu_bf16x2(f2bf2(a[0], a[1])), hi = relu_bf16x2(f2bf2(a[2], a[3]));
            dw3[n] += gN[0] * __uint_as_float(lo << 16);
            dw3[n] += gN[1] * __uint_as_float(lo & 0xFFFF0000u);
            dw3[n] += gN[2] * __uint_as_float(hi << 16);
            dw3[n] += gN[3] * __uint_as_float(hi & 0xFFFF0000u);
        }
        // D2 in T by the identity product; D1 = relu'(A1) (D2 W2) in N.
        bf16x4 d2T[4];
        for (int m = 0; m < 4; ++m) d2T[m] = pack_bf16x4(mfma16(d2n[m], ident, zero4));
        const s16x8 d2Tp[2] = {cat8(d2T[0], d2T[1]), cat8(d2T[2], d2T[3])};
        bf16x4 d1n[4];
        for (int n = 0; n < 4; ++n) {
            f32x4 a = zero4;
            for (int j = 0; j < 2; ++j) a = mfma32(d2Tp[j], sh_w2t[n][j][lane], a);
            f32x4 d;
            for (int e = 0; e < 4; ++e) {
                d[e] = h1n[n][e] != 0 ? a[e] : 0.f;
                db1[n] += d[e];
            }
            d1n[n] = pack_bf16x4(d);
        }
        // Weight gradients, the nonzeros in the K slot.
        for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 4; ++n) dW2[m][n] = mfma16(d2n[m], h1n[n], dW2[m][n]);
        for (int m = 0; m < 4; ++m) dW1[m] = mfma16(d1n[m], in.xn, dW1[m]);
    };
    // Two input buffers in turn: the next tile's loads are in flight while a
    // tile is processed.
    // (The index of the next tile is clamped, so a load is always issued.)
    TileIn in0, in1;
    uint32_t tile = blockIdx.x * kMlpWaves + wave;
    if (tile < tiles) load_in(tile, in0);
    while (tile < tiles) {
        load_in(std::min(tile + stride, tiles - 1), in1);
        process(in0);
        tile += stride;
        if (tile >= tiles) break;
        load_in(std::min(tile + stride, tiles - 1), in0);
        process(in1);
        tile += stride;
    }

    // Bias-like sums: the 4 lane groups hold the same units for different
    // nonzeros.
    for (int n = 0; n < 4; ++n) {
        db1[n] += __shfl_xor(db1[n], 16, 64);
        db1[n] += __shfl_xor(db1[n], 32, 64);
        db2[n] += __shfl_xor(db2[n], 16, 64);
        db2[n] += __shfl_xor(db2[n], 32, 64);
        dw3[n] += __shfl_xor(dw3[n], 16, 64);
        dw3[n] += __shfl_xor(dw3[n], 32, 64);
    }
    db3 = wave_sum(db3);

    // Waves 0..3 in order: the first stores, the others add.
    for (uint32_t wv = 0; wv < static_cast<uint32_t>(kMlpWaves); ++wv) {
        if (wave == wv) {
            const bool first = (wv == 0);
            auto put = [&](int idx, float v) { sh_g[idx] = first ? v : sh_g[idx] + v; };
            // dW2 block (m, n): lane l holds rows 16m + 4q + e, column 16n + c16.
            for (int m = 0; m < 4; ++m)
                for (int n = 0; n < 4; ++n)
                    for (int e = 0; e < 4; ++e)
                        put(kOffW2 + (16 * m + 4 * q + e) * kHid + 16 * n + c16, dW2[m][n][e]);
            for (int m = 0; m < 4; ++m)
                for (int e = 0; e < 4; ++e) put(kOffW1 + (16 * m + 4 * q + e) * kFeat + c16, dW1[m][e]);
            if (q == 0)
                for (int n = 0; n < 4; ++n) {
                    put(kOffB1 + 16 * n + c16, db1[n]);
                    put(kOffB2 + 16 * n + c16, db2[n]);
                    put(kOffW3 + 16 * n + c16, dw3[n]);
                }
            if (lane == 0) put(kOffB3, db3);
        }
        __syncthreads();
    }
    float* out = partials + static_cast<size_t>(blockIdx.x) * kNumParams;
    for (uint32_t i = threadIdx.x; i < static_cast<uint32_t>(kNumParams); i += 64 * kMlpWaves)
        out[i] = sh_g[i];
}

// Two-stage fixed-order sum of `count` partial gradient vectors:
// stage 1: block (x, c) sums vectors [c*per, (c+1)*per) for 256 parameters;
// stage 2: the chunk sums in chunk order.
constexpr uint32_t kGradChunks = 32;

__global__ __launch_bounds__(256) void grad_reduce_stage1(const float* __restrict__ partials,
                                                          uint32_t count, double* __restrict__ chunk)
{
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= static_cast<uint32_t>(kNumParams)) return;
    const uint32_t per = (count + kGradChunks - 1) / kGradChunks;
    const uint32_t k0 = blockIdx.y * per;
    const uint32_t k1 = min(count, k0 + per);
    double s = 0.0;
    for (uint32_t k = k0; k < k1; ++k) s += partials[static_cast<size_t>(k) * kNumParams + p];
    chunk[static_cast<size_t>(blockIdx.y) * kNumParams + p] = s;
}

__global__ __launch_bounds__(256) void grad_reduce_stage2(const double* __restrict__ chunk,
                                                          float* __restrict__ grad)
{
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= static_cast<uint32_t>(kNumParams)) return;
    double s = 0.0;
    for (uint32_t c = 0; c < kGradChunks; ++c) s += chunk[static_cast<size_t>(c) * kNumParams + p];
    grad[p] = static_cast<float>(s);
}

// Adam step and refresh of the BF16 weight copies.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, float* __restrict__ m1,
                                                   float* __restrict__ m2,
                                                   const float* __restrict__ g, float lr,
                                                   float b1t, float b2t, uint16_t* __restrict__ W1,
                                                   uint16_t* __restrict__ W2,
                                                   uint16_t* __restrict__ W2T)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kNumParams) return;
    const float gi = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gi;
    const float b = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float mh = a / (1.f - b1t), vh = b / (1.f - b2t);
    const float v = p[i] - lr * mh / (sqrtf(vh) + 1e-8f);
    p[i] = v;
    if (i < kOffB1) {
        W1[i] = f2bf(v);
    } else if (i >= kOffW2 && i < kOffB2) {
        const int k = i - kOffW2, r = k / kHid, c = k % kHid;
        W2[k] = f2bf(v);
        W2T[c * kHid + r] = f2bf(v);
    }
}

__global__ __launch_bounds__(256) void refresh_bf16_kernel(const float* __restrict__ p,
                                                           uint16_t* __restrict__ W1,
                                                           uint16_t* __restrict__ W2,
                                                           uint16_t* __restrict__ W2T)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kHid * kFeat) W1[i] = f2bf(p[kOffW1 + i]);
    if (i < kHid * kHid) {
        const int r = i / kHid, c = i % kHid;
        W2[i] = f2bf(p[kOffW2 + i]);
        W2T[c * kHid + r] = f2bf(p[kOffW2 + i]);
    }
}

// ---- features and G assembly --------------------------------------------------

// Per row of A: d = a_ii (1 if no stored diagonal), then with dhat = 1/sqrt(d):
// r = sum_{k != i} |a_ik| dhat_i dhat_k, s = sum_{k != i} a_ik dhat_i dhat_k.
__global__ __launch_bounds__(256) void row_diag_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                       const uint32_t* __restrict__ ci,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ dhat)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double d = 1.0;
    for (uint32_t j = rp[i]; j < rp[i + 1]; ++j)
        if (ci[j] == i) {
            d = v[j];
            break;
        }
    dhat[i] = d > 0 ? 1.0 / sqrt(d) : 1.0;
}

__global__ __launch_bounds__(256) void row_stats_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ ci,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ dhat,
                                                        float* __restrict__ stats)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double r = 0.0, s = 0.0;
    for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) {
        const uint32_t c = ci[j];
        if (c == i || c >= n) continue;
        const double a = v[j] * dhat[i] * dhat[c];
        r += fabs(a);
        s += a;
    }
    stats[3 * i + 0] = static_cast<float>(r);
    stats[3 * i + 1] = static_cast<float>(s);
    stats[3 * i + 2] = static_cast<float>(rp[i + 1] - rp[i]) * 0.0625f;
}

// Features of G's nonzero k = (i, j), j <= i, taken from A's nonzero src[k]:
// [ahat_ij, i == j, r_i, r_j, s_i, s_j, ahat_ij^2, nnz_i / 16, 0 ...].
__global__ __launch_bounds__(256) void feature_kernel(uint32_t nnzG, const uint32_t* __restrict__ grow,
                                                      const uint32_t* __restrict__ gcol,
                                                      const uint32_t* __restrict__ src,
                                                      const double* __restrict__ v,
                                                      const double* __restrict__ dhat,
                                                      const float* __restrict__ stats,
                                                      uint16_t* __restrict__ X)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnzG) return;
    const uint32_t i = grow[k], j = gcol[k];
    const float a = static_cast<float>(v[src[k]] * dhat[i] * dhat[j]);
    float f[kFeat] = {a, i == j ? 1.f : 0.f, stats[3 * i], stats[3 * j], stats[3 * i + 1],
                      stats[3 * j + 1], a * a, stats[3 * i + 2], 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t w[kFeat / 2];
    for (int e = 0; e < kFeat / 2; ++e)
        w[e] = static_cast<uint32_t>(f2bf(f[2 * e])) | (static_cast<uint32_t>(f2bf(f[2 * e + 1])) << 16);
    uint4* dst = reinterpret_cast<uint4*>(X + static_cast<size_t>(k) * kFeat);
    dst[0] = uint4{w[0], w[1], w[2], w[3]};
    dst[1] = uint4{w[4], w[5], w[6], w[7]};
}

// G_k = (delta + o_k) dhat_i, scattered to G (CSR) and G^T (CSR) in both
// precisions; the FP32 assembly also writes BF16 copies (g16, gt16).
template <typename T>
__global__ __launch_bounds__(256) void assemble_kernel(uint32_t nnzG, const uint32_t* __restrict__ grow,
                                                       const uint32_t* __restrict__ gcol,
                                                       const uint32_t* __restrict__ tpos,
                                                       const float* __restrict__ o,
                                                       const double* __restrict__ dhat,
                                                       T* __restrict__ gval, T* __restrict__ gtval,
                                                       uint16_t* __restrict__ g16,
                                                       uint16_t* __restrict__ gt16,
                                                       const uint32_t* __restrict__ dia_pos,
                                                       uint16_t* __restrict__ dia16,
                                                       int* __restrict__ bad)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnzG) return;
    const uint32_t i = grow[k], j = gcol[k];
    const double g = ((i == j ? 1.0 : 0.0) + static_cast<double>(o ? o[k] : 0.f)) * dhat[i];
    if (!isfinite(g)) *bad = 1;
    gval[k] = static_cast<T>(g);
    gtval[tpos[k]] = static_cast<T>(g);
    if (g16) {
        const uint16_t h = f2bf(static_cast<float>(g));
        g16[k] = h;
        gt16[tpos[k]] = h;
        if (dia16) dia16[dia_pos[k]] = h;
    }
}

// Hutchinson gradient for one probe: gG[k] += c (u_i v_j + w_i z_j).
// dL/dG_k = sum over the probes p, in order, of c (u_ip v_jp + w_ip z_jp) for
// the k-th nonzero (i, j) of G; the probe blocks are row-major n x np.
__global__ __launch_bounds__(256) void hutch_grad_kernel(uint32_t nnzG, uint32_t np,
                                                         const uint32_t* __restrict__ grow,
                                                         const uint32_t* __restrict__ gcol,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ z, double c,
                                                         double* __restrict__ gG)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnzG) return;
    const size_t i = static_cast<size_t>(grow[k]) * np, j = static_cast<size_t>(gcol[k]) * np;
    double g = 0.0;
    uint32_t p = 0;
    if ((np & 1) == 0) {  // rows of the blocks are 16-byte aligned: two probes per load
        typedef double d2 __attribute__((ext_vector_type(2)));
        for (; p < np; p += 2) {
            const d2 ui = *reinterpret_cast<const d2*>(u + i + p), vj = *reinterpret_cast<const d2*>(v + j + p);
            const d2 wi = *reinterpret_cast<const d2*>(w + i + p), zj = *reinterpret_cast<const d2*>(z + j + p);
            g += c * (ui[0] * vj[0] + wi[0] * zj[0]);
            g += c * (ui[1] * vj[1] + wi[1] * zj[1]);
        }
    }
    for (; p < np; ++p) g += c * (u[i + p] * v[j + p] + w[i + p] * z[j + p]);
    gG[k] = g;
}

// dL/do_k = dL/dG_k * dhat_i
__global__ __launch_bounds__(256) void chain_kernel(uint32_t nnzG, const uint32_t* __restrict__ grow,
                                                    const double* __restrict__ gG,
                                                    const double* __restrict__ dhat,
                                                    float* __restrict__ gout)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnzG) return;
    gout[k] = static_cast<float>(gG[k] * dhat[grow[k]]);
}

// Rademacher probes from a counter-based hash of (seed_p, i): z[i * np + p] with
// seed_p = seed * 0x100000001B3 + p.
__global__ __launch_bounds__(256) void rademacher_kernel(uint32_t n, uint32_t np, uint64_t seed,
                                                         double* __restrict__ z)
{
    const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= static_cast<size_t>(n) * np) return;
    const uint64_t i = e / np, p = e % np;
    uint64_t x = (seed * 0x100000001B3ull + p) ^ (0x9E3779B97F4A7C15ull * (i + 1));
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    z[e] = (x & 1) ? 1.0 : -1.0;
}

__global__ __launch_bounds__(256) void sub_kernel(size_t n, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ out)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

template <typename T>
__global__ __launch_bounds__(256) void set_scalar(T* p, T v)
{
    if (threadIdx.x == 0) *p = v;
}

uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + 255) / 256); }

}  // namespace

// ---- host objects -------------------------------------------------------------

struct _clsparseAmdMlp
{
    int device = 0;
    float* p = nullptr;   // kNumParams FP32 master
    float* m1 = nullptr;  // Adam moments
    float* m2 = nullptr;
    float* grad = nullptr;
    float* partials = nullptr;  // kMlpGrid partial gradients (one per workgroup)
    double* chunk = nullptr;    // kGradChunks stage-1 sums
    uint16_t* W1 = nullptr;
    uint16_t* W2 = nullptr;
    uint16_t* W2T = nullptr;
    int step = 0;
    MlpDev dev() const { return MlpDev{W1, W2, W2T, p}; }
};

struct _clsparseAmdFsai
{
    int device = 0;
    uint32_t n = 0, nnzG = 0, nnzA = 0;
    bool full_diag = true;  // every row of A stores its diagonal (else G is singular)
    // Outcome of the solver's pilot race for the current G (0 not run, 1 G,
    // 2 Jacobi) and the two measured rates; reset whenever G changes.
    int race = 0;
    double race_rate_fsai = 0.0, race_rate_jacobi = 0.0;
    // Pattern of G (CSR, lower) and of G^T (CSR, upper); per-G-nonzero row,
    // column, source index in A and position in G^T.
    uint32_t *g_rp = nullptr, *g_ci = nullptr, *gt_rp = nullptr, *gt_ci = nullptr;
    uint32_t *g_row = nullptr, *g_src = nullptr, *g_tpos = nullptr;
    double* dhat = nullptr;
    float* stats = nullptr;
    uint16_t* X = nullptr;  // features, BF16 [nnzG][16]
    float* o = nullptr;     // network output per G nonzero
    double *g64 = nullptr, *gt64 = nullptr;
    float *g32 = nullptr, *gt32 = nullptr;
    uint16_t *g16 = nullptr, *gt16 = nullptr;  // BF16 copies of the FP32 values
    // Diagonal storage of the BF16 values (see FsaiDia): ndiag = 0 when G's
    // pattern needs more than kFsaiMaxDiags diagonals or too much padding.
    int ndiag = 0;
    int32_t dia_off[kFsaiMaxDiags] = {};
    uint16_t* dia16 = nullptr;   // [ndiag][n]
    uint32_t* dia_pos = nullptr; // per G nonzero: its index in dia16
    int* bad = nullptr;
    clsparseCsrMatrix G64{}, GT64{}, G32{}, GT32{};
    // The CSR metadata of G and G^T (G32 / G16 share G64's, GT32 / GT16
    // GT64's) is created on first use: solves from diagonal storage never
    // need it.
    bool meta_ready = false;
    clsparseCsrMatrix G16{}, GT16{};  // share G32's / GT32's metadata (not owned)
    bool valid = true;
    // A training step assembles only the FP64 values (no host sync): the
    // solver's FP32 / BF16 / diagonal copies and `valid` are then brought up
    // to date from the stored network output before the factor is used.
    bool solver_stale = false;
    bool g64_stale = false;  // the FP64 copies (g64, gt64) are not assembled yet
    bool have_o = false;  // o holds the network output G was last assembled from
    // Device memory of the arrays above: two allocations (the n-sized
    // arrays, then the nnz(G)-sized ones), carved in 256-byte pieces.
    void* arena_n = nullptr;
    void* arena_g = nullptr;
    // Training scratch: the probe blocks Z, T, U, S, V, T2, W (n x probes,
    // row-major, one allocation for up to `probe_cap` probes), the G gradient.
    void* probe_buf = nullptr;
    uint32_t probe_cap = 0;
    double *gG = nullptr, *sc = nullptr;
    float* gout = nullptr;
};

namespace {

template <typename T>
hipError_t dmalloc(T** p, size_t count)
{
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

enum class Assemble { Full, Training, Refresh };
clsparseStatus fsai_assemble(_clsparseAmdFsai* f, clsparseAmdMlp net, hipStream_t st, int num_cus = 0,
                             Assemble mode = Assemble::Full);
clsparseStatus defer_jacobi(_clsparseAmdFsai* f);

// Diagonal scaling, row statistics and features from A's values.
clsparseStatus compute_features(_clsparseAmdFsai* f, const clsparseCsrMatrix* A, hipStream_t st)
{
    const uint32_t n = f->n;
    const uint32_t* arp = dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer);
    const uint32_t* aci = dev_ptr<uint32_t>(A->col_indices, A->off_col_indices);
    const double* av = dev_ptr<double>(A->values, A->off_values);
    if (n == 0) return clsparseSuccess;
    hipLaunchKernelGGL(row_diag_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, arp, aci, av, f->dhat);
    hipLaunchKernelGGL(row_stats_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, arp, aci, av,
                       f->dhat, f->stats);
    if (f->nnzG)
        hipLaunchKernelGGL(feature_kernel, dim3(blocks_for(f->nnzG)), dim3(256), 0, st, f->nnzG,
                           f->g_row, f->g_ci, f->g_src, av, f->dhat, f->stats, f->X);
    CLS_HIP_TRY(hipGetLastError());
    CLS_HIP_TRY(hipStreamSynchronize(st));
    return clsparseSuccess;
}

void free_fsai(_clsparseAmdFsai* f)
{
    if (!f) return;
    clsparseCsrMetaDelete(&f->G64);  // G32 / G16 share it
    clsparseCsrMetaDelete(&f->GT64);
    void* ptrs[] = {f->arena_n, f->arena_g, f->probe_buf, f->gG, f->sc, f->gout};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete f;
}

}  // namespace

extern "C" clsparseStatus clsparseAmdMlpCreate(clsparseAmdMlp* out, cl_ulong seed)
{
    if (out == nullptr) return clsparseInvalidValue;
    _clsparseAmdMlp* m = new (std::nothrow) _clsparseAmdMlp;
    if (!m) return clsparseOutOfHostMemory;
    (void)hipGetDevice(&m->device);
    // He-normal W1, W2; w3 = 0 and b3 = 0 so the initial G is Jacobi exactly.
    std::vector<float> h(kNumParams, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 12345;
    auto nrm = [&]() {
        auto nx = [&]() {
            st += 0x9E3779B97F4A7C15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return static_cast<double>((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
        };
        double u1 = nx(), u2 = nx();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < kHid * kFeat; ++i) h[kOffW1 + i] = static_cast<float>(nrm() * std::sqrt(2.0 / 8));
    for (int i = 0; i < kHid * kHid; ++i) h[kOffW2 + i] = static_cast<float>(nrm() * std::sqrt(2.0 / kHid));
    hipError_t e = dmalloc(&m->p, kNumParams);
    if (e == hipSuccess) e = dmalloc(&m->m1, kNumParams);
    if (e == hipSuccess) e = dmalloc(&m->m2, kNumParams);
    if (e == hipSuccess) e = dmalloc(&m->grad, kNumParams);
    if (e == hipSuccess) e = dmalloc(&m->partials, static_cast<size_t>(kMlpGrid) * kNumParams);
    if (e == hipSuccess) e = dmalloc(&m->chunk, static_cast<size_t>(kGradChunks) * kNumParams);
    if (e == hipSuccess) e = dmalloc(&m->W1, kHid * kFeat);
    if (e == hipSuccess) e = dmalloc(&m->W2, kHid * kHid);
    if (e == hipSuccess) e = dmalloc(&m->W2T, kHid * kHid);
    if (e == hipSuccess) e = hipMemcpy(m->p, h.data(), kNumParams * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m1, 0, kNumParams * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->m2, 0, kNumParams * sizeof(float));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(refresh_bf16_kernel, dim3(16), dim3(256), 0, 0, m->p, m->W1, m->W2, m->W2T);
        e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) {
        clsparseAmdMlpDestroy(m);
        return clsparseOutOfResources;
    }
    *out = m;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdMlpDestroy(clsparseAmdMlp m)
{
    if (m == nullptr) return clsparseInvalidValue;
    void* ptrs[] = {m->p, m->m1, m->m2, m->grad, m->partials, m->chunk, m->W1, m->W2, m->W2T};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
    return clsparseSuccess;
}

extern "C" cl_uint clsparseAmdMlpNumParams(void) { return kNumParams; }

extern "C" clsparseStatus clsparseAmdMlpGetParams(clsparseAmdMlp m, cl_float* host)
{
    if (m == nullptr || host == nullptr) return clsparseInvalidValue;
    CLS_HIP_TRY(hipMemcpy(host, m->p, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdMlpSetParams(clsparseAmdMlp m, const cl_float* host)
{
    if (m == nullptr || host == nullptr) return clsparseInvalidValue;
    CLS_HIP_TRY(hipMemcpy(m->p, host, kNumParams * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(refresh_bf16_kernel, dim3(16), dim3(256), 0, 0, m->p, m->W1, m->W2, m->W2T);
    CLS_HIP_TRY(hipDeviceSynchronize());
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdMlpGetGrad(clsparseAmdMlp m, cl_float* host)
{
    if (m == nullptr || host == nullptr) return clsparseInvalidValue;
    CLS_HIP_TRY(hipMemcpy(host, m->grad, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
    return clsparseSuccess;
}

// Pattern of G = tril(A) and of G^T, built on the host from A's pattern.
namespace {

// ---- pattern of G = tril(A) and of G^T, built on the device ---------------------

// Row r of A: entries with column <= r (G's row length); *bad on a row
// pointer that decreases or leaves [0, nnzA], or a column >= n; *nodiag when
// the row has no diagonal entry.
__global__ __launch_bounds__(256) void tril_count_kernel(uint32_t n, uint32_t nnzA, const uint32_t* __restrict__ rp,
                                                         const uint32_t* __restrict__ ci, uint32_t* __restrict__ cnt,
                                                         int* __restrict__ bad, int* __restrict__ nodiag)
{
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const uint32_t b = rp[r], e = rp[r + 1];
    if (e < b || e > nnzA) {
        cnt[r] = 0;
        *bad = 1;
        return;
    }
    uint32_t c0 = 0;
    bool diag = false;
    for (uint32_t k = b; k < e; ++k) {
        const uint32_t c = ci[k];
        if (c >= n) {
            *bad = 1;
            continue;
        }
        c0 += c <= r;
        diag = diag || c == r;
    }
    cnt[r] = c0;
    if (!diag) *nodiag = 1;
}

// G's row r: column, row and source index in A of each entry with column <= r.
__global__ __launch_bounds__(256) void tril_fill_kernel(uint32_t n, const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ ci,
                                                        const uint32_t* __restrict__ grp, uint32_t* __restrict__ gci,
                                                        uint32_t* __restrict__ grow, uint32_t* __restrict__ gsrc,
                                                        uint64_t* __restrict__ key, uint32_t* __restrict__ idx)
{
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    uint32_t pos = grp[r];
    for (uint32_t k = rp[r]; k < rp[r + 1]; ++k) {
        const uint32_t c = ci[k];
        if (c > r) continue;
        gci[pos] = c;
        grow[pos] = r;
        gsrc[pos] = k;
        key[pos] = (static_cast<uint64_t>(c) << 32) | r;  // G^T order: by column, then row
        idx[pos] = pos;
        ++pos;
    }
}

// Sorted (column, row) keys -> G^T's columns and the position of every G entry in G^T.
// Thread p of 0 .. nnz: G^T's column and G's position of sorted entry p, and
// the starts of the G^T rows c in (column of p - 1, column of p] (every row
// from the first to n is written once; empty rows start where the next
// non-empty one does).
__global__ __launch_bounds__(256) void transpose_fill_kernel(uint32_t nnz, uint32_t n, const uint64_t* __restrict__ key,
                                                             const uint32_t* __restrict__ idx,
                                                             uint32_t* __restrict__ tci, uint32_t* __restrict__ tpos,
                                                             uint32_t* __restrict__ trp)
{
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p > nnz) return;
    const int64_t prev = p == 0 ? -1 : static_cast<int64_t>(key[p - 1] >> 32);
    const int64_t cur = p == nnz ? static_cast<int64_t>(n) : static_cast<int64_t>(key[p] >> 32);
    for (int64_t c = prev + 1; c <= cur; ++c) trp[c] = p;
    if (p == nnz) return;
    tci[p] = static_cast<uint32_t>(key[p] & 0xFFFFFFFFull);
    tpos[idx[p]] = p;
}

// Position of every G entry in the diagonal storage; *bad when an entry lies
// off the K candidate diagonals.
__global__ __launch_bounds__(256) void dia_position_kernel(uint32_t nnz, uint32_t n, const uint32_t* __restrict__ gci,
                                                           const uint32_t* __restrict__ grow, int K, DiaOffsets off,
                                                           uint32_t* __restrict__ pos, int* __restrict__ bad)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nnz) return;
    const int64_t o = static_cast<int64_t>(gci[k]) - static_cast<int64_t>(grow[k]);
    int d = -1;
    for (int j = 0; j < K; ++j)
        if (off.o[j] == o) d = j;
    if (d < 0) {
        *bad = 1;
        return;
    }
    pos[k] = static_cast<uint32_t>(d) * n + grow[k];
}

__global__ __launch_bounds__(256) void row_pointer_check_kernel(uint32_t n, uint32_t nnz, const uint32_t* __restrict__ rp,
                                                                int* __restrict__ bad)
{
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r < n && (rp[r + 1] < rp[r] || rp[r + 1] > nnz)) *bad = 1;
}

uint32_t blocks256(uint64_t n) { return static_cast<uint32_t>(std::max<uint64_t>(1, (n + 255) / 256)); }

// clsparseInvalidValue unless rp[0..n] never decreases and stays within nnz.
clsparseStatus check_row_pointer(const uint32_t* rp, uint32_t n, uint32_t nnz, int* flag, hipStream_t st)
{
    CLS_HIP_TRY(hipMemsetAsync(flag, 0, sizeof(int), st));
    hipLaunchKernelGGL(row_pointer_check_kernel, dim3(blocks256(n)), dim3(256), 0, st, n, nnz, rp, flag);
    CLS_HIP_TRY(hipGetLastError());
    int h = 0;
    CLS_HIP_TRY(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    CLS_HIP_TRY(hipStreamSynchronize(st));
    return h ? clsparseInvalidValue : clsparseSuccess;
}

// Byte offsets of arrays to be carved from one device allocation, each
// 256-byte aligned (planned before the allocation is made).
struct ByteLayout
{
    size_t bytes = 0;
    template <typename T>
    size_t add(size_t count)
    {
        const size_t at = bytes;
        bytes += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~static_cast<size_t>(255);
        return at;
    }
};
template <typename T>
T* at(void* base, size_t offset)
{
    return reinterpret_cast<T*>(static_cast<char*>(base) + offset);
}

// G = tril(A) (CSR, with each entry's row and source index in A), G^T (CSR,
// rows of G^T in ascending order), and the diagonal storage of G's BF16
// values when its pattern lies on at most kFsaiMaxDiags diagonals (found from
// sampled rows, checked on every entry). All on the device; the host reads
// back only counts and flags.
clsparseStatus build_pattern(_clsparseAmdFsai* f, const uint32_t* arp, const uint32_t* aci, uint32_t n,
                             uint32_t nnzA, clsparseControl control)
{
    hipStream_t st = control->stream;
    // The n-sized arrays of the factor in one allocation.
    {
        ByteLayout c;
        const size_t o_grp = c.add<uint32_t>(static_cast<size_t>(n) + 1), o_gtrp = c.add<uint32_t>(static_cast<size_t>(n) + 1);
        const size_t o_dhat = c.add<double>(n), o_stats = c.add<float>(3 * static_cast<size_t>(n));
        if (hipMalloc(&f->arena_n, c.bytes) != hipSuccess) {
            f->arena_n = nullptr;
            return clsparseOutOfResources;
        }
        f->g_rp = at<uint32_t>(f->arena_n, o_grp);
        f->gt_rp = at<uint32_t>(f->arena_n, o_gtrp);
        f->dhat = at<double>(f->arena_n, o_dhat);
        f->stats = at<float>(f->arena_n, o_stats);
    }
    // Build scratch in the control's grow-only workspace: the row counts, the
    // flags and the scan's temporary storage first.
    size_t tb = 0;
    CLS_HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tb, static_cast<uint32_t*>(nullptr), f->g_rp + 1,
                                                 static_cast<int>(n), st));
    uint32_t* cnt = nullptr;
    int* flags = nullptr;  // [0] bad, [1] no diagonal
    void* tstore = nullptr;
    {
        ByteLayout c;
        const size_t o_cnt = c.add<uint32_t>(n), o_flags = c.add<int>(4), o_t = c.add<char>(tb);
        void* work = nullptr;
        CLS_TRY(control_workspace(control, c.bytes, &work));
        cnt = at<uint32_t>(work, o_cnt);
        flags = at<int>(work, o_flags);
        tstore = at<void>(work, o_t);
    }
    CLS_HIP_TRY(hipMemsetAsync(flags, 0, 4 * sizeof(int), st));
    hipLaunchKernelGGL(tril_count_kernel, dim3(blocks256(n)), dim3(256), 0, st, n, nnzA, arp, aci, cnt, flags,
                       flags + 1);
    CLS_HIP_TRY(hipGetLastError());
    CLS_HIP_TRY(hipMemsetAsync(f->g_rp, 0, sizeof(uint32_t), st));
    CLS_HIP_TRY(hipcub::DeviceScan::InclusiveSum(tstore, tb, cnt, f->g_rp + 1, static_cast<int>(n), st));
    int hflags[4] = {0, 0, 0, 0};
    uint32_t nnzG = 0;
    CLS_HIP_TRY(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    CLS_HIP_TRY(hipMemcpyAsync(&nnzG, f->g_rp + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CLS_HIP_TRY(hipStreamSynchronize(st));
    if (hflags[0]) return clsparseInvalidValue;
    f->nnzG = nnzG;
    f->full_diag = hflags[1] == 0;

    // The nnz(G)-sized arrays of the factor in one allocation (the diagonal
    // storage sized for the most diagonals it may hold).
    {
        ByteLayout c;
        const size_t o_ci = c.add<uint32_t>(nnzG), o_row = c.add<uint32_t>(nnzG), o_src = c.add<uint32_t>(nnzG);
        const size_t o_gtci = c.add<uint32_t>(nnzG), o_tpos = c.add<uint32_t>(nnzG);
        const size_t o_X = c.add<uint16_t>(static_cast<size_t>(nnzG) * kFeat), o_o = c.add<float>(nnzG);
        const size_t o_g64 = c.add<double>(nnzG), o_gt64 = c.add<double>(nnzG);
        const size_t o_g32 = c.add<float>(nnzG), o_gt32 = c.add<float>(nnzG);
        const size_t o_g16 = c.add<uint16_t>(nnzG), o_gt16 = c.add<uint16_t>(nnzG);
        const size_t o_dia = c.add<uint16_t>(static_cast<size_t>(kFsaiMaxDiags) * n), o_dpos = c.add<uint32_t>(nnzG);
        const size_t o_bad = c.add<int>(1);
        if (hipMalloc(&f->arena_g, c.bytes) != hipSuccess) {
            f->arena_g = nullptr;
            return clsparseOutOfResources;
        }
        void* b = f->arena_g;
        f->g_ci = at<uint32_t>(b, o_ci);
        f->g_row = at<uint32_t>(b, o_row);
        f->g_src = at<uint32_t>(b, o_src);
        f->gt_ci = at<uint32_t>(b, o_gtci);
        f->g_tpos = at<uint32_t>(b, o_tpos);
        f->X = at<uint16_t>(b, o_X);
        f->o = at<float>(b, o_o);
        f->g64 = at<double>(b, o_g64);
        f->gt64 = at<double>(b, o_gt64);
        f->g32 = at<float>(b, o_g32);
        f->gt32 = at<float>(b, o_gt32);
        f->g16 = at<uint16_t>(b, o_g16);
        f->gt16 = at<uint16_t>(b, o_gt16);
        f->dia16 = at<uint16_t>(b, o_dia);
        f->dia_pos = at<uint32_t>(b, o_dpos);
        f->bad = at<int>(b, o_bad);
    }
    // Scratch for G^T: the sort's keys and indices, its temporary storage,
    // flags (the workspace is taken again: nothing of the first scratch is
    // needed any more). The sort orders the column bits of the (column, row)
    // keys only: it is stable and G's entries come in row order, so the rows
    // within a column stay ascending (3 passes of 8 bits at n = 4.1 M, not 7).
    int end_bit = 33;
    while (end_bit < 64 && (uint64_t(1) << (end_bit - 32)) < n) ++end_bit;
    uint64_t *key = nullptr, *key2 = nullptr;
    uint32_t *idx = nullptr, *idx2 = nullptr;
    size_t tb3 = 0;
    CLS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, static_cast<uint64_t*>(nullptr),
                                                   static_cast<uint64_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                                   static_cast<uint32_t*>(nullptr), static_cast<int>(nnzG), 32,
                                                   end_bit, st));
    void* tstore2 = nullptr;
    {
        ByteLayout c;
        const size_t o_k = c.add<uint64_t>(nnzG), o_k2 = c.add<uint64_t>(nnzG);
        const size_t o_i = c.add<uint32_t>(nnzG), o_i2 = c.add<uint32_t>(nnzG);
        const size_t o_t = c.add<char>(tb3), o_f = c.add<int>(4);
        void* work = nullptr;
        CLS_TRY(control_workspace(control, c.bytes, &work));
        key = at<uint64_t>(work, o_k);
        key2 = at<uint64_t>(work, o_k2);
        idx = at<uint32_t>(work, o_i);
        idx2 = at<uint32_t>(work, o_i2);
        tstore2 = at<void>(work, o_t);
        flags = at<int>(work, o_f);
    }
    CLS_HIP_TRY(hipMemsetAsync(flags, 0, 4 * sizeof(int), st));
    hipLaunchKernelGGL(tril_fill_kernel, dim3(blocks256(n)), dim3(256), 0, st, n, arp, aci, f->g_rp, f->g_ci,
                       f->g_row, f->g_src, key, idx);
    // G^T: a radix sort of the (column, row) keys (rows ascend within each
    // column); its columns, G's positions and its row pointer from the
    // sorted keys.
    if (nnzG > 0)
        CLS_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tstore2, tb3, key, key2, idx, idx2, static_cast<int>(nnzG), 32,
                                                       end_bit, st));
    hipLaunchKernelGGL(transpose_fill_kernel, dim3(blocks256(static_cast<uint64_t>(nnzG) + 1)), dim3(256), 0, st,
                       nnzG, n, key2, idx2, f->gt_ci, f->g_tpos, f->gt_rp);
    CLS_HIP_TRY(hipGetLastError());

    // Diagonal storage of the BF16 values.
    clsparseCsrMatrix Gm;
    clsparseInitCsrMatrix(&Gm);
    Gm.num_rows = Gm.num_cols = n;
    Gm.num_nonzeros = nnzG;
    Gm.row_pointer = f->g_rp;
    Gm.col_indices = f->g_ci;
    DiaOffsets off{};
    int K = 0;
    CLS_TRY(dia_candidates(&Gm, reinterpret_cast<int32_t*>(control->red_partials), st, &off, &K, false));
    if (K > kFsaiMaxDiags || static_cast<uint64_t>(K) * n >= (uint64_t(1) << 32) || (K > 0 && off.o[K - 1] != 0))
        K = 0;
    if (K > 0) {
        CLS_HIP_TRY(hipMemsetAsync(f->dia16, 0, static_cast<size_t>(K) * n * sizeof(uint16_t), st));  // padding stays 0
        hipLaunchKernelGGL(dia_position_kernel, dim3(blocks256(nnzG)), dim3(256), 0, st, nnzG, n, f->g_ci, f->g_row, K,
                           off, f->dia_pos, flags + 2);
        CLS_HIP_TRY(hipGetLastError());
        int off_diag = 0;
        CLS_HIP_TRY(hipMemcpyAsync(&off_diag, flags + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        CLS_HIP_TRY(hipStreamSynchronize(st));
        if (off_diag) K = 0;
    }
    if (K == 0) {
        f->dia16 = nullptr;
        f->dia_pos = nullptr;
    }
    f->ndiag = K;
    for (int d = 0; d < K; ++d) f->dia_off[d] = off.o[d];
    CLS_HIP_TRY(hipStreamSynchronize(st));  // before the scratch buffers go
    return clsparseSuccess;
}

}  // namespace

// CSR metadata of G and G^T on first use (training, the FP64 solver, and the
// FP32 solver when the factor is not in diagonal storage).
clsparseStatus fsai_ensure_meta(clsparseAmdFsai f, clsparseControl control)
{
    if (f->meta_ready) return clsparseSuccess;
    CLS_TRY(clsparseCsrMetaCreate(&f->G64, control));
    CLS_TRY(clsparseCsrMetaCreate(&f->GT64, control));
    f->G32.meta = f->G16.meta = f->G64.meta;
    f->GT32.meta = f->GT16.meta = f->GT64.meta;
    f->meta_ready = true;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdFsaiCreate(clsparseAmdFsai* out, const clsparseCsrMatrix* A,
                                                clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (out == nullptr || A == nullptr) return clsparseInvalidValue;
    if (A->num_rows != A->num_cols) return clsparseInvalidSystemSize;
    if (A->num_rows == 0) return clsparseInvalidSize;
    // The device scans and sorts of the pattern count items in int.
    if (A->num_rows > 0x7FFFFFFFu || A->num_nonzeros > 0x7FFFFFFFu) return clsparseInvalidSize;
    const uint32_t n = A->num_rows, nnzA = A->num_nonzeros;
    hipStream_t st = control->stream;
    const uint32_t* arp = dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer);
    const uint32_t* aci = dev_ptr<uint32_t>(A->col_indices, A->off_col_indices);
    if (arp == nullptr || (nnzA > 0 && aci == nullptr)) return clsparseInvalidMemObject;

    _clsparseAmdFsai* f = new (std::nothrow) _clsparseAmdFsai;
    if (!f) return clsparseOutOfHostMemory;
    f->device = control->device;
    f->n = n;
    f->nnzA = nnzA;
    clsparseStatus s = build_pattern(f, arp, aci, n, nnzA, control);
    if (s != clsparseSuccess) {
        free_fsai(f);
        return s;
    }
    const uint32_t nnzG = f->nnzG;
    auto csr = [&](clsparseCsrMatrix& M, uint32_t* prp, uint32_t* pci, void* vals) {
        clsparseInitCsrMatrix(&M);
        M.num_rows = M.num_cols = n;
        M.num_nonzeros = nnzG;
        M.row_pointer = prp;
        M.col_indices = pci;
        M.values = vals;
    };
    csr(f->G64, f->g_rp, f->g_ci, f->g64);
    csr(f->GT64, f->gt_rp, f->gt_ci, f->gt64);
    csr(f->G32, f->g_rp, f->g_ci, f->g32);
    csr(f->GT32, f->gt_rp, f->gt_ci, f->gt32);
    f->G16 = f->G32;
    f->G16.values = f->g16;
    f->GT16 = f->GT32;
    f->GT16.values = f->gt16;

    s = compute_features(f, A, st);
    if (s == clsparseSuccess) s = defer_jacobi(f);
    if (s != clsparseSuccess) {
        free_fsai(f);
        return s;
    }
    *out = f;
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdFsaiUpdateValues(clsparseAmdFsai f, const clsparseCsrMatrix* A,
                                                      clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (f == nullptr || A == nullptr) return clsparseInvalidValue;
    if (A->num_rows != f->n || A->num_cols != f->n || A->num_nonzeros != f->nnzA)
        return clsparseInvalidSystemSize;
    // The kernels walk A's rows: the row pointer must stay inside the arrays.
    CLS_TRY(check_row_pointer(dev_ptr<uint32_t>(A->row_pointer, A->off_row_pointer), f->n, f->nnzA, f->bad,
                              control->stream));
    CLS_TRY(compute_features(f, A, control->stream));
    return defer_jacobi(f);
}

extern "C" clsparseStatus clsparseAmdFsaiDestroy(clsparseAmdFsai f)
{
    if (f == nullptr) return clsparseInvalidValue;
    free_fsai(f);
    return clsparseSuccess;
}

namespace {

// num_cus: the device's compute units (sizes the forward kernel's grid).
// Full: the network's output (or the Jacobi factor when net is null), every
// value copy and `valid` (one host sync). Training: the network's output and
// the FP64 values only (the training step needs nothing else). Refresh: the
// value copies and `valid` from the stored output, after training steps.
clsparseStatus fsai_assemble(_clsparseAmdFsai* f, clsparseAmdMlp net, hipStream_t st, int num_cus, Assemble mode)
{
    if (f->nnzG == 0) {
        f->valid = false;
        f->solver_stale = false;
        return clsparseSuccess;
    }
    if (net && mode != Assemble::Refresh) {
        // As many workgroups as the GPU holds at once (queried once), or one
        // per 4 tiles if fewer.
        static const int per_cu = [] {
            int b = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, mlp_forward_kernel, 64 * kMlpWaves, 0) != hipSuccess)
                b = 0;
            return std::max(b, 1);
        }();
        const uint32_t tiles = (f->nnzG + 15) / 16;
        const uint32_t resident = static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(std::max(num_cus, 1));
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((tiles + kMlpWaves - 1) / kMlpWaves, resident));
        hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid), dim3(64 * kMlpWaves), 0, st, net->dev(), f->X,
                           f->nnzG, f->o);
    }
    if (mode != Assemble::Refresh) f->have_o = net != nullptr;
    const float* o = f->have_o ? f->o : nullptr;
    if (mode == Assemble::Training) {
        hipLaunchKernelGGL((assemble_kernel<double>), dim3(blocks_for(f->nnzG)), dim3(256), 0, st, f->nnzG,
                           f->g_row, f->g_ci, f->g_tpos, o, f->dhat, f->g64, f->gt64,
                           static_cast<uint16_t*>(nullptr), static_cast<uint16_t*>(nullptr),
                           static_cast<const uint32_t*>(nullptr), static_cast<uint16_t*>(nullptr), f->bad);
        CLS_HIP_TRY(hipGetLastError());
        f->solver_stale = true;
        f->g64_stale = false;
        fsai_set_race(f, 0, 0.0, 0.0);  // a new factor races (or is judged) afresh
        return clsparseSuccess;
    }
    CLS_HIP_TRY(hipMemsetAsync(f->bad, 0, sizeof(int), st));
    hipLaunchKernelGGL((assemble_kernel<double>), dim3(blocks_for(f->nnzG)), dim3(256), 0, st, f->nnzG,
                       f->g_row, f->g_ci, f->g_tpos, o, f->dhat, f->g64, f->gt64,
                       static_cast<uint16_t*>(nullptr), static_cast<uint16_t*>(nullptr),
                       static_cast<const uint32_t*>(nullptr), static_cast<uint16_t*>(nullptr), f->bad);
    hipLaunchKernelGGL((assemble_kernel<float>), dim3(blocks_for(f->nnzG)), dim3(256), 0, st, f->nnzG,
                       f->g_row, f->g_ci, f->g_tpos, o, f->dhat, f->g32, f->gt32, f->g16, f->gt16,
                       f->dia_pos, f->dia16, f->bad);
    CLS_HIP_TRY(hipGetLastError());
    int bad = 0;
    CLS_HIP_TRY(hipMemcpyAsync(&bad, f->bad, sizeof(int), hipMemcpyDeviceToHost, st));
    CLS_HIP_TRY(hipStreamSynchronize(st));
    f->valid = (bad == 0) && f->full_diag;
    fsai_set_race(f, 0, 0.0, 0.0);
    f->solver_stale = false;
    f->g64_stale = false;
    return clsparseSuccess;
}

// G = D^-1/2 (no network output) left to be assembled where it is first
// read (the solver's refresh, clsparseAmdFsaiGetValues): a caller usually
// applies a network next, which assembles G itself. D^-1/2 is finite by
// construction (row_diag_kernel), so the factor is valid with a full diagonal.
clsparseStatus defer_jacobi(_clsparseAmdFsai* f)
{
    CLS_HIP_TRY(hipGetLastError());
    f->have_o = false;
    f->valid = f->nnzG > 0 && f->full_diag;
    f->solver_stale = f->nnzG > 0;
    f->g64_stale = f->nnzG > 0;
    fsai_set_race(f, 0, 0.0, 0.0);
    return clsparseSuccess;
}

clsparseStatus ensure_train_scratch(_clsparseAmdFsai* f, uint32_t probes)
{
    if (!f->gG) {
        hipError_t e = dmalloc(&f->gG, f->nnzG);
        if (e == hipSuccess) e = dmalloc(&f->sc, 8);
        if (e == hipSuccess) e = dmalloc(&f->gout, f->nnzG);
        const double constants[2] = {1.0, 0.0};
        if (e == hipSuccess) e = hipMemcpy(f->sc, constants, sizeof(constants), hipMemcpyHostToDevice);
        if (e != hipSuccess) return clsparseOutOfResources;
    }
    if (probes > f->probe_cap) {
        if (f->probe_buf) (void)hipFree(f->probe_buf);
        f->probe_buf = nullptr;
        f->probe_cap = 0;
        const size_t block = (static_cast<size_t>(f->n) * probes * sizeof(double) + 255) / 256 * 256;
        if (hipMalloc(&f->probe_buf, std::max<size_t>(7 * block, 1)) != hipSuccess) return clsparseOutOfResources;
        f->probe_cap = probes;
    }
    return clsparseSuccess;
}

}  // namespace

extern "C" clsparseStatus clsparseAmdFsaiApplyNet(clsparseAmdFsai f, clsparseAmdMlp net,
                                                  clsparseControl control)
{
    CLS_TRY(check_init_control(control));
    if (f == nullptr) return clsparseInvalidValue;
    return fsai_assemble(f, net, control->stream, control->num_cus);
}

extern "C" clsparseStatus clsparseAmdFsaiGetValues(clsparseAmdFsai f, cl_uint* g_row_pointer,
                                                   cl_uint* g_col_indices, cl_double* g_values,
                                                   cl_uint* nnz)
{
    if (f == nullptr) return clsparseInvalidValue;
    if (nnz) *nnz = f->nnzG;
    if (g_row_pointer)
        CLS_HIP_TRY(hipMemcpy(g_row_pointer, f->g_rp, (f->n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (g_col_indices)
        CLS_HIP_TRY(hipMemcpy(g_col_indices, f->g_ci, f->nnzG * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (g_values) {
        if (f->g64_stale) CLS_TRY(fsai_assemble(f, nullptr, nullptr, 0, Assemble::Refresh));
        CLS_HIP_TRY(hipMemcpy(g_values, f->g64, f->nnzG * sizeof(double), hipMemcpyDeviceToHost));
    }
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdFsaiGetFeatures(clsparseAmdFsai f, uint16_t* host_bf16)
{
    if (f == nullptr || host_bf16 == nullptr) return clsparseInvalidValue;
    CLS_HIP_TRY(hipMemcpy(host_bf16, f->X, static_cast<size_t>(f->nnzG) * kFeat * sizeof(uint16_t),
                          hipMemcpyDeviceToHost));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdFsaiGetOutput(clsparseAmdFsai f, cl_float* host)
{
    if (f == nullptr || host == nullptr) return clsparseInvalidValue;
    CLS_HIP_TRY(hipMemcpy(host, f->o, f->nnzG * sizeof(float), hipMemcpyDeviceToHost));
    return clsparseSuccess;
}

extern "C" clsparseStatus clsparseAmdFsaiGetDiagonals(clsparseAmdFsai f, cl_int* ndiag, cl_int* offsets)
{
    if (f == nullptr || ndiag == nullptr) return clsparseInvalidValue;
    *ndiag = f->ndiag;
    if (offsets)
        for (int d = 0; d < kFsaiMaxDiags; ++d) offsets[d] = d < f->ndiag ? f->dia_off[d] : 0;
    return clsparseSuccess;
}

// Loss and parameter gradient for one matrix; with lr > 0 also an Adam step.
// probes: Rademacher vectors seeded by (seed, probe index).
extern "C" clsparseStatus clsparseAmdFsaiTrainStep(clsparseAmdFsai f, clsparseAmdMlp net,
                                                   const clsparseCsrMatrix* A, cl_uint probes,
                                                   cl_ulong seed, cl_float lr,
                                                   clsparseControl control, cl_double* loss,
                                                   cl_double* grad_g_host)
{
    CLS_TRY(check_init_control(control));
    if (f == nullptr || net == nullptr || A == nullptr || probes == 0) return clsparseInvalidValue;
    if (A->num_rows != f->n || A->num_cols != f->n) return clsparseInvalidSystemSize;
    if (f->nnzG == 0) return clsparseInvalidValue;
    CLS_TRY(ensure_train_scratch(f, probes));
    hipStream_t st = control->stream;
    CLS_TRY(fsai_assemble(f, net, st, control->num_cus, Assemble::Training));
    const uint32_t n = f->n, nnzG = f->nnzG;
    const double* one = f->sc;  // constants 1 and 0, written with the scratch
    const double* zero = f->sc + 1;
    // All probes at once: each product is one sparse x dense (n x probes)
    // product, reading G, G^T and A once for every probe.
    const size_t np = probes, elems = static_cast<size_t>(n) * np;
    const size_t block = (elems * sizeof(double) + 255) / 256 * 256;
    auto pb = [&](int k) { return reinterpret_cast<double*>(static_cast<char*>(f->probe_buf) + k * block); };
    double *Z = pb(0), *T = pb(1), *U = pb(2), *S = pb(3), *V = pb(4), *T2 = pb(5), *W = pb(6);
    const int cus = control->num_cus;
    hipLaunchKernelGGL(rademacher_kernel, dim3(blocks_for(elems)), dim3(256), 0, st, n, probes, seed, Z);
    CLS_TRY(launch_csrmm_rows<double>(&f->G64, probes, Z, probes, T, probes, one, zero, cus, st));   // T = G Z
    CLS_TRY(launch_csrmm_rows<double>(A, probes, T, probes, U, probes, one, zero, cus, st));         // U = A T
    CLS_TRY(launch_csrmm_rows<double>(&f->GT64, probes, U, probes, S, probes, one, zero, cus, st));  // S = G^T U
    hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(elems)), dim3(256), 0, st, elems, S, Z, V);       // V = S - Z
    CLS_TRY(launch_csrmm_rows<double>(&f->G64, probes, V, probes, T2, probes, one, zero, cus, st));  // T2 = G V
    CLS_TRY(launch_csrmm_rows<double>(A, probes, T2, probes, W, probes, one, zero, cus, st));        // W = A T2
    const double c = 2.0 / (static_cast<double>(probes) * n);
    hipLaunchKernelGGL(hutch_grad_kernel, dim3(blocks_for(nnzG)), dim3(256), 0, st, nnzG, probes, f->g_row,
                       f->g_ci, U, V, W, Z, c, f->gG);
    CLS_TRY(launch_reduce<double, double>(RedOp::Sqr, V, nullptr, elems, f->sc + 2, false, control->red_partials,
                                          st));
    double total = 0.0;
    CLS_HIP_TRY(hipMemcpyAsync(&total, f->sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    CLS_HIP_TRY(hipStreamSynchronize(st));
    if (loss) *loss = total / (static_cast<double>(probes) * n);
    if (grad_g_host)
        CLS_HIP_TRY(hipMemcpyAsync(grad_g_host, f->gG, nnzG * sizeof(double), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(chain_kernel, dim3(blocks_for(nnzG)), dim3(256), 0, st, nnzG, f->g_row, f->gG,
                       f->dhat, f->gout);
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(kMlpGrid), dim3(64 * kMlpWaves), 0, st, net->dev(), f->X,
                       f->gout, nnzG, net->partials);
    hipLaunchKernelGGL(grad_reduce_stage1, dim3(blocks_for(kNumParams), kGradChunks), dim3(256), 0, st,
                       net->partials, static_cast<uint32_t>(kMlpGrid), net->chunk);
    hipLaunchKernelGGL(grad_reduce_stage2, dim3(blocks_for(kNumParams)), dim3(256), 0, st, net->chunk,
                       net->grad);
    if (lr > 0.f) {
        net->step += 1;
        const float b1t = std::pow(0.9f, static_cast<float>(net->step));
        const float b2t = std::pow(0.999f, static_cast<float>(net->step));
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kNumParams)), dim3(256), 0, st, net->p, net->m1,
                           net->m2, net->grad, lr, b1t, b2t, net->W1, net->W2, net->W2T);
    }
    CLS_HIP_TRY(hipGetLastError());
    CLS_HIP_TRY(hipStreamSynchronize(st));
    return clsparseSuccess;
}

// Accessors for the solver (cg.hip). fsai_prepare first: the value copies
// the solver reads are brought up to date after training steps.
clsparseStatus fsai_prepare(clsparseAmdFsai f, hipStream_t st)
{
    if (!f->solver_stale) return clsparseSuccess;
    return fsai_assemble(f, nullptr, st, 0, Assemble::Refresh);
}
const clsparseCsrMatrix* fsai_G32(clsparseAmdFsai f) { return &f->G32; }
const clsparseCsrMatrix* fsai_GT32(clsparseAmdFsai f) { return &f->GT32; }
const clsparseCsrMatrix* fsai_G64(clsparseAmdFsai f) { return &f->G64; }
const clsparseCsrMatrix* fsai_G16(clsparseAmdFsai f) { return &f->G16; }
const clsparseCsrMatrix* fsai_GT16(clsparseAmdFsai f) { return &f->GT16; }
const clsparseCsrMatrix* fsai_GT64(clsparseAmdFsai f) { return &f->GT64; }
bool fsai_valid(clsparseAmdFsai f) { return f->valid; }
int fsai_race(clsparseAmdFsai f, double* rate_fsai, double* rate_jacobi)
{
    *rate_fsai = f->race_rate_fsai;
    *rate_jacobi = f->race_rate_jacobi;
    return f->race;
}
void fsai_set_race(clsparseAmdFsai f, int outcome, double rate_fsai, double rate_jacobi)
{
    f->race = outcome;
    f->race_rate_fsai = rate_fsai;
    f->race_rate_jacobi = rate_jacobi;
}
uint32_t fsai_n(clsparseAmdFsai f) { return f->n; }
bool fsai_dia(clsparseAmdFsai f, FsaiDia* out)
{
    if (f->ndiag == 0) return false;
    out->ndiag = f->ndiag;
    for (int d = 0; d < kFsaiMaxDiags; ++d) out->off[d] = f->dia_off[d];
    out->vals = f->dia16;
    return true;
}
