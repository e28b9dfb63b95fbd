// gold.h -- length-31 Gold sequence (TS 38.211 5.2.1) on the device with
// GF(2) jump-ahead, so any thread can start at any sequence index.
//
// State convention: a 31-bit word whose bit i is x(n + i).  One step maps
// x(n..n+30) to x(n+1..n+31) with x1(n+31) = x1(n+3) ^ x1(n) and
// x2(n+31) = x2(n+3) ^ x2(n+2) ^ x2(n+1) ^ x2(n).  kGoldJump[s][k] holds the
// 31 columns of the step matrix raised to 2^k for LFSR s (0 = x1, 1 = x2).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace airan {

constexpr int kGoldJumpBits = 40;

struct GoldJumpTable {
    uint32_t col[2][kGoldJumpBits][31];
};

// Filled once per device by InitGoldTables (gold.hip).
extern __constant__ GoldJumpTable kGoldJump;
int InitGoldTables();

__device__ inline uint32_t GfMatVec(const uint32_t *col, uint32_t v)
{
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 31; ++j) r ^= (0u - ((v >> j) & 1u)) & col[j];
    return r;
}

// LFSR states at absolute index n (x1 starts at e_0, x2 at c_init).
__device__ inline void GoldStateAt(uint64_t n, uint32_t cinit, uint32_t *s1, uint32_t *s2)
{
    uint32_t a = 1u, b = cinit & 0x7fffffffu;
    for (int k = 0; k < kGoldJumpBits && n != 0; ++k, n >>= 1) {
        if (n & 1u) {
            a = GfMatVec(kGoldJump.col[0][k], a);
            b = GfMatVec(kGoldJump.col[1][k], b);
        }
    }
    *s1 = a;
    *s2 = b;
}

// Emits c(n..n+31) (bit j = c(n + j)) and advances both states by 32.
__device__ inline uint32_t GoldNext32(uint32_t *s1, uint32_t *s2)
{
    uint64_t w1 = *s1, w2 = *s2;
    w1 |= ((w1 ^ (w1 >> 3)) & 0x0fffffffull) << 31;
    w1 |= (((w1 >> 28) ^ (w1 >> 31)) & 0xfull) << 59;
    w2 |= ((w2 ^ (w2 >> 1) ^ (w2 >> 2) ^ (w2 >> 3)) & 0x0fffffffull) << 31;
    w2 |= (((w2 >> 28) ^ (w2 >> 29) ^ (w2 >> 30) ^ (w2 >> 31)) & 0xfull) << 59;
    *s1 = static_cast<uint32_t>(w1 >> 32) & 0x7fffffffu;
    *s2 = static_cast<uint32_t>(w2 >> 32) & 0x7fffffffu;
    return static_cast<uint32_t>(w1 ^ w2);
}

// Convenience: 32 Gold bits c(start..start+31) of sequence c_init.
__device__ inline uint32_t GoldWordAt(uint32_t cinit, uint64_t start)
{
    uint32_t s1, s2;
    GoldStateAt(start + 1600ull, cinit, &s1, &s2);
    return GoldNext32(&s1, &s2);
}

}  // namespace airan
