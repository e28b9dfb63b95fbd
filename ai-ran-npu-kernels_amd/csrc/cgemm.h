// cgemm.h -- complex GEMM building blocks for the channel estimator's batched
// projections: VALU FMA (CGemmTile) and fp32 MFMA (CGemmTileMfma) tiles with
// LDS-staged fp32 operands, and split-precision bf16 MFMA blocks (below) for
// operands pre-packed in fragment order.  Plain bf16/fp16 operands would not
// do: the filter's dynamic range and the subspace noise estimate need more
// than 11 significand bits.
//
// One workgroup of 256 threads computes a 64 x 64 tile of C = A B with
// K-steps of 16: A (64 x 16) and B (16 x 64) complex tiles are staged in LDS
// through caller loaders (so operands can be gathered or scaled on the fly),
// each thread accumulates a 4 x 4 block in registers, and the caller's store
// functor receives every in-range element.  Every C element is the same
// ordered sum over k in every launch shape, so results do not depend on how
// the N dimension is batched.
#pragma once

#include "common.h"

namespace airan {

constexpr int kGemmThreads = 256;
constexpr int kGemmTile = 64;
constexpr int kGemmK = 16;

__device__ inline void CMacAcc(float2 &c, float2 a, float2 b)
{
#pragma clang fp contract(on)
    c.x += a.x * b.x - a.y * b.y;
    c.y += a.x * b.y + a.y * b.x;
}

// LoadA(m, k) / LoadB(k, n) return the operand (callers handle bounds and
// return zero outside); Store(m, n, value) is called for m < M, n < N.
// B_KFAST: consecutive threads load consecutive k of B (for operands stored
// with k contiguous); otherwise consecutive n.
template <bool B_KFAST, typename LoadA, typename LoadB, typename Store>
__device__ inline void CGemmTile(uint32_t m0, uint32_t n0, uint32_t M, uint32_t N, uint32_t K, LoadA la,
                                 LoadB lb, Store st)
{
    __shared__ float2 as[kGemmK][kGemmTile + 1];
    __shared__ float2 bs[kGemmK][kGemmTile + 1];
    const uint32_t t = threadIdx.x;
    const uint32_t tm = t & 15u, tn = t >> 4;
    float2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = make_float2(0.0f, 0.0f);
    for (uint32_t k0 = 0; k0 < K; k0 += kGemmK) {
        // 1024 elements of each tile, 4 per thread.
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t e = t + r * kGemmThreads;
            const uint32_t am = e & 63u, ak = e >> 6;       // A: m fastest
            as[ak][am] = la(m0 + am, k0 + ak);
            const uint32_t bn = B_KFAST ? e >> 4 : e & 63u;
            const uint32_t bk = B_KFAST ? e & 15u : e >> 6;
            bs[bk][bn] = lb(k0 + bk, n0 + bn);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGemmK; ++kk) {
            float2 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = as[kk][tm + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = bs[kk][tn + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) CMacAcc(acc[i][j], a[i], b[j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t m = m0 + tm + 16 * i, n = n0 + tn + 16 * j;
            if (m < M && n < N) st(m, n, acc[i][j]);
        }
}

typedef float cg_f32x4 __attribute__((ext_vector_type(4)));

// The same tile on fp32 MFMA (v_mfma_f32_16x16x4f32): the 4 waves form a
// 2 x 2 grid of 32 x 32 sub-tiles, each two 16 x 16 blocks in M and N; per
// K-step of 4 a block takes four real MFMAs (re = ar br + ai (-bi),
// im = ar bi + ai br).  Operands are staged in LDS exactly as in CGemmTile.
// MFMA 16x16x4 f32 lane layout: A row = lane % 16, K = lane / 16; B column =
// lane % 16, K = lane / 16; accumulator i holds row 4 (lane / 16) + i of
// column lane % 16.  Every C element is the same ordered sum over K-steps of
// 16 in every launch shape.
template <bool B_KFAST, typename LoadA, typename LoadB, typename Store>
__device__ inline void CGemmTileMfma(uint32_t m0, uint32_t n0, uint32_t M, uint32_t N, uint32_t K, LoadA la,
                                     LoadB lb, Store st)
{
    __shared__ float2 as[kGemmK][kGemmTile + 1];
    __shared__ float2 bs[kGemmK][kGemmTile + 1];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u, wave = t >> 6;
    const uint32_t wm = (wave & 1u) * 32u, wn = (wave >> 1) * 32u;
    const uint32_t l16 = lane & 15u, kq = lane >> 4;
    cg_f32x4 acc_r[2][2], acc_i[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc_r[i][j] = cg_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            acc_i[i][j] = cg_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    for (uint32_t k0 = 0; k0 < K; k0 += kGemmK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t e = t + r * kGemmThreads;
            const uint32_t am = e & 63u, ak = e >> 6;
            as[ak][am] = la(m0 + am, k0 + ak);
            const uint32_t bn = B_KFAST ? e >> 4 : e & 63u;
            const uint32_t bk = B_KFAST ? e & 15u : e >> 6;
            bs[bk][bn] = lb(k0 + bk, n0 + bn);
        }
        __syncthreads();
#pragma unroll
        for (uint32_t ks = 0; ks < kGemmK; ks += 4) {
            float2 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = as[ks + kq][wm + 16 * i + l16];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = bs[ks + kq][wn + 16 * j + l16];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    acc_r[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].x, b[j].x, acc_r[i][j], 0, 0, 0);
                    acc_r[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].y, -b[j].y, acc_r[i][j], 0, 0, 0);
                    acc_i[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].x, b[j].y, acc_i[i][j], 0, 0, 0);
                    acc_i[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].y, b[j].x, acc_i[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t m = m0 + wm + 16 * i + 4 * kq + q, n = n0 + wn + 16 * j + l16;
                if (m < M && n < N) st(m, n, make_float2(acc_r[i][j][q], acc_i[i][j][q]));
            }
}

// ---------------------------------------------------------------------------
// Split-precision complex products on bf16 MFMA (v_mfma_f32_32x32x16_bf16).
// An fp32 value x is carried as x_hi = bf16(x) and x_lo = bf16(x - x_hi)
// (16 significand bits, |x - x_hi - x_lo| <= 2^-17 |x|) and a product as
// x_hi y_hi + x_hi y_lo + x_lo y_hi, accumulated in fp32: relative error
// about 2^-16 per product (the dropped x_lo y_lo is below 2^-16 |x y|), the
// fp32 exponent range throughout.
//
// Operands live in "fragment planes": a complex matrix X with rows r and K
// columns is stored as bf16
//   [r / 32][k / 16][plane][(k % 16) / 8][r % 32][k % 8]
// with planes re_hi, re_lo, im_hi, im_lo.  The eight K values that lane l
// of a 32x32x16 MFMA takes for row l % 32 (k = 8 (l / 32) + j, for both the
// A and the B operand) are then one 16-byte load, and a wave's load of one
// plane of a 32 x 16 block is 1 KB contiguous.  Accumulator element i of lane
// l holds C row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr uint32_t kFragPlane = 32 * 16;       // bf16 per plane of a 32 x 16 block
constexpr uint32_t kFragBlock = 4 * kFragPlane;

__device__ inline size_t FragIndex(uint32_t r, uint32_t k, uint32_t plane, uint32_t kblocks)
{
    return (((static_cast<size_t>(r >> 5) * kblocks + (k >> 4)) * 4 + plane) * 2 + ((k >> 3) & 1u)) * 256 +
           (r & 31u) * 8 + (k & 7u);
}

// Round-to-nearest-even bf16 of a finite fp32 value (bits).
__device__ inline uint32_t Bf16Bits(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// hi | lo << 16 of x.
__device__ inline uint32_t Bf16Split(float x)
{
    const uint32_t hi = Bf16Bits(x);
    return hi | Bf16Bits(x - __uint_as_float(hi << 16)) << 16;
}

// Splits v[0..7] (row r, K values k0 .. k0 + 7, k0 a multiple of 8) into the
// four planes of `base` (fragment order, `kblocks` = K / 16).
__device__ inline void StoreFrag8(uint16_t *__restrict__ base, uint32_t r, uint32_t k0, uint32_t kblocks,
                                  const float2 *v)
{
    uint32_t w[4][4] = {};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t sr = Bf16Split(v[j].x), si = Bf16Split(v[j].y), sh = (j & 1) * 16;
        w[0][j >> 1] |= (sr & 0xffffu) << sh;
        w[1][j >> 1] |= (sr >> 16) << sh;
        w[2][j >> 1] |= (si & 0xffffu) << sh;
        w[3][j >> 1] |= (si >> 16) << sh;
    }
#pragma unroll
    for (uint32_t p = 0; p < 4; ++p)
        *reinterpret_cast<uint4 *>(base + FragIndex(r, k0, p, kblocks)) = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
}

// One complex 32 x 16 operand block, 8 K values per plane for this lane.
struct CFrag {
    uint4 rh, rl, ih, il;
};

__device__ inline CFrag LoadFrag(const uint16_t *__restrict__ block, uint32_t lane)
{
    const uint4 *p = reinterpret_cast<const uint4 *>(block) + lane;
    return CFrag{p[0], p[kFragPlane / 8], p[2 * kFragPlane / 8], p[3 * kFragPlane / 8]};
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

// C += A B for one 32 x 32 complex block over 16 K values (12 MFMAs):
// re += ar br - ai bi, im += ar bi + ai br, each product split as above;
// nai_h / nai_l are -ai's planes.  The order of the terms is fixed.
__device__ inline void CMfmaSplit(f32x16_t &cr, f32x16_t &ci, const CFrag &a, uint4 nai_h, uint4 nai_l,
                                  const CFrag &b)
{
    cr = MfmaBf16(a.rh, b.rh, cr);
    cr = MfmaBf16(a.rh, b.rl, cr);
    cr = MfmaBf16(a.rl, b.rh, cr);
    cr = MfmaBf16(nai_h, b.ih, cr);
    cr = MfmaBf16(nai_h, b.il, cr);
    cr = MfmaBf16(nai_l, b.ih, cr);
    ci = MfmaBf16(a.rh, b.ih, ci);
    ci = MfmaBf16(a.rh, b.il, ci);
    ci = MfmaBf16(a.rl, b.ih, ci);
    ci = MfmaBf16(a.ih, b.rh, ci);
    ci = MfmaBf16(a.ih, b.rl, ci);
    ci = MfmaBf16(a.il, b.rh, ci);
}

}  // namespace airan
