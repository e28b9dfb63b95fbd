// cgemm.h -- complex fp32 GEMM tiles for the channel estimator's batched
// projections: VALU FMA (CGemmTile) and fp32 MFMA (CGemmTileMfma), both with
// LDS-staged operands (fp32 throughout: the filter's dynamic range is too
// wide for bf16/fp16 operands).
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

}  // namespace airan
