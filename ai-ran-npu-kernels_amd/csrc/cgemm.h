// cgemm.h -- split-precision complex products on bf16 MFMA
// (v_mfma_f32_32x32x16_bf16) for the channel estimator's batched GEMMs.
//
// An fp32 value x is carried as S bf16 parts: x_0 = bf16(x), x_1 =
// bf16(x - x_0), ... (each remainder exact in fp32), 8 S significand bits
// in all.  A product x y is the sum of the part products x_i y_j with
// i + j < S plus, for S = 3, x_1 y_1, accumulated in fp32:
//   S = 2: 3 terms, relative error about 2^-16 (the dropped x_1 y_1);
//   S = 3: 6 terms, relative error about 2^-24, i.e. fp32 products.
// Plain bf16/fp16 operands would not do: the filter's dynamic range and the
// subspace noise estimate (a difference of nearly equal energies) need at
// least fp32 significands in the projection; the interpolation, whose
// output is rounded to fp16, takes S = 2.
//
// Operands live in "fragment planes": a complex matrix X with rows r and K
// columns is stored as bf16
//   [r / 32][k / 16][plane][(k % 16) / 8][r % 32][k % 8]
// with 2 S planes, the real parts x_0 .. x_{S-1} then the imaginary parts.
// The eight K values that lane l of a 32x32x16 MFMA takes for row l % 32
// (k = 8 (l / 32) + j, for both the A and the B operand) are then one
// 16-byte load, and a wave's load of one plane of a 32 x 16 block is 1 KB
// contiguous.  Accumulator element i of lane l holds C row
// 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.
#pragma once

#include "common.h"

namespace airan {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr uint32_t kFragPlane = 32 * 16;  // bf16 per plane of a 32 x 16 block

template <int S>
constexpr uint32_t FragBlock()
{
    return 2u * S * kFragPlane;
}

template <int S>
__device__ inline size_t FragIndex(uint32_t r, uint32_t k, uint32_t plane, uint32_t kblocks)
{
    return (((static_cast<size_t>(r >> 5) * kblocks + (k >> 4)) * (2 * S) + plane) * 2 + ((k >> 3) & 1u)) * 256 +
           (r & 31u) * 8 + (k & 7u);
}

// Round-to-nearest-even bf16 of a finite fp32 value (bits).
__device__ inline uint32_t Bf16Bits(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

template <int S>
__device__ inline void Bf16Parts(float x, uint32_t *part)
{
#pragma unroll
    for (int s = 0; s < S; ++s) {
        part[s] = Bf16Bits(x);
        x -= __uint_as_float(part[s] << 16);
    }
}

// Splits v[0..7] (row r, K values k0 .. k0 + 7, k0 a multiple of 8) into the
// 2 S planes of `base` (fragment order, `kblocks` = K / 16).
template <int S>
__device__ inline void StoreFrag8(uint16_t *__restrict__ base, uint32_t r, uint32_t k0, uint32_t kblocks,
                                  const float2 *v)
{
    uint32_t w[2 * S][4] = {};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        uint32_t pr[S], pi[S];
        Bf16Parts<S>(v[j].x, pr);
        Bf16Parts<S>(v[j].y, pi);
        const uint32_t sh = (j & 1) * 16;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            w[s][j >> 1] |= pr[s] << sh;
            w[S + s][j >> 1] |= pi[s] << sh;
        }
    }
#pragma unroll
    for (uint32_t p = 0; p < 2 * S; ++p)
        *reinterpret_cast<uint4 *>(base + FragIndex<S>(r, k0, p, kblocks)) =
            make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
}

// This lane's 8 K values of every plane of one 32 x 16 block.
template <int S>
struct CFrag {
    uint4 p[2 * S];
};

template <int S>
__device__ inline CFrag<S> LoadFrag(const uint16_t *__restrict__ block, uint32_t lane)
{
    const uint4 *q = reinterpret_cast<const uint4 *>(block) + lane;
    CFrag<S> f;
#pragma unroll
    for (int i = 0; i < 2 * S; ++i) f.p[i] = q[i * (kFragPlane / 8)];
    return f;
}

// Copies one 1 KB plane of a fragment block (64 lanes x 16 B at src) to LDS
// dst in the same order (lane l -> dst + 16 l) by the direct global -> LDS
// path (global_load_lds_dwordx4: no VGPRs, counted by vmcnt); the caller
// waits (WaitVm) and synchronises before reading dst.
__device__ inline void PlaneToLds(const uint16_t *__restrict__ src, uint4 *dst, uint32_t lane)
{
    __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint4 *>(src) + lane,
                                     (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// s_waitcnt vmcnt(N), N < 16, other counters not waited on.
template <int N>
__device__ inline void WaitVm()
{
    static_assert(N >= 0 && N < 16, "vmcnt field");
    __builtin_amdgcn_s_waitcnt(0xF70 | N);
}

__device__ inline uint4 NegBf16x8(uint4 v)
{
    return make_uint4(v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u, v.w ^ 0x80008000u);
}

__device__ inline f32x16_t MfmaBf16(uint4 a, uint4 b, f32x16_t c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

// Part products kept for S parts, in accumulation order: (i, j) pairs.
template <int S>
struct SplitTerms;
template <>
struct SplitTerms<2> {
    static constexpr int n = 3;
    static constexpr int a[3] = {0, 0, 1};
    static constexpr int b[3] = {0, 1, 0};
};
template <>
struct SplitTerms<3> {
    static constexpr int n = 6;
    static constexpr int a[6] = {0, 0, 1, 0, 2, 1};
    static constexpr int b[6] = {0, 1, 0, 2, 0, 1};
};

// C += A B for one 32 x 32 complex block over 16 K values:
// re += ar br + (-ai) bi, im += ar bi + ai br, every product split as above;
// nai holds -ai's parts.  The order of the terms is fixed.
template <int S>
__device__ inline void CMfmaSplit(f32x16_t &cr, f32x16_t &ci, const CFrag<S> &a, const uint4 *nai,
                                  const CFrag<S> &b)
{
    using T = SplitTerms<S>;
#pragma unroll
    for (int t = 0; t < T::n; ++t) cr = MfmaBf16(a.p[T::a[t]], b.p[T::b[t]], cr);
#pragma unroll
    for (int t = 0; t < T::n; ++t) cr = MfmaBf16(nai[T::a[t]], b.p[S + T::b[t]], cr);
#pragma unroll
    for (int t = 0; t < T::n; ++t) ci = MfmaBf16(a.p[T::a[t]], b.p[S + T::b[t]], ci);
#pragma unroll
    for (int t = 0; t < T::n; ++t) ci = MfmaBf16(a.p[S + T::a[t]], b.p[T::b[t]], ci);
}

}  // namespace airan
