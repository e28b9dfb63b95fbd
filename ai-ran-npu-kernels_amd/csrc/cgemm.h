// cgemm.h -- complex fp32 GEMM tile for the channel estimator's batched
// projections (VALU FMA, LDS-staged operands; the matrices are far too
// narrow in one dimension and too sensitive to precision for bf16 MFMA).
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

}  // namespace airan
