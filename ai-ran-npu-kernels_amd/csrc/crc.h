// crc.h -- CRC checks on decoded bits (TS 38.212 5.1), shared by the CRC
// operator and the receive chain.
//
// The remainder of a message is linear in its bits: for an n-bit message
// (data followed by its parity) it is the XOR over set bits j of
// x^(L + n - 1 - j) mod g, with the powers tabulated per polynomial.  A
// transport block's remainder is the XOR over its code blocks of each code
// block's local remainder times x^(bits after it) mod g.
#pragma once

#include "common.h"

namespace airan {

// Everything a kernel needs to check a batch of code blocks and their TBs.
struct CrcJob {
    const uint32_t *pow_b;   // x^(24 + e) mod g_CRC24B
    const uint32_t *pow_tb;  // x^(len + e) mod g_TB (CRC24A or CRC16)
    uint8_t *cb_ok;          // [blocks]; nullptr: no check
    uint32_t *tb_part;       // [blocks] each block's TB remainder part, shifted
    const uint32_t *xpow;    // [C] x^((C - 1 - r) L) mod g_TB per code block r, or nullptr
    uint32_t C, kprime, crc_cb, L, poly_tb, len_tb;
};

// Carry-less product a * b mod g for len-bit residues.
__device__ inline uint32_t CrcMulMod(uint32_t a, uint32_t b, uint32_t poly, uint32_t len)
{
    const uint32_t mask = (1u << len) - 1u;
    uint32_t r = 0;
    for (int i = static_cast<int>(len) - 1; i >= 0; --i) {
        const uint32_t top = (r >> (len - 1)) & 1u;
        r = (r << 1) & mask;
        if (top) r ^= poly;
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

constexpr uint32_t kCrcPolyB = 0x800063u;  // CRC24B without the x^24 term

// x^((C - 1 - r) L) mod g_TB for code block cbg (r = cbg mod C): the shift
// of its TB remainder part (a table load, or square-and-multiply without one).
__device__ inline uint32_t CrcTbShift(const CrcJob &c, uint32_t cbg)
{
    const uint32_t r = cbg % c.C;
    if (c.xpow != nullptr) return c.xpow[r];
    uint32_t xpow = 1u, base = 2u, e = (c.C - 1 - r) * c.L;
    while (e != 0) {
        if (e & 1u) xpow = CrcMulMod(xpow, base, c.poly_tb, c.len_tb);
        base = CrcMulMod(base, base, c.poly_tb, c.len_tb);
        e >>= 1;
    }
    return xpow;
}

// Called by one thread per code block with the block-reduced remainders and
// the block's CrcTbShift.
__device__ inline void CrcFinishShifted(const CrcJob &c, uint32_t cbg, uint32_t rb, uint32_t ra, uint32_t xpow)
{
    c.cb_ok[cbg] = c.crc_cb ? static_cast<uint8_t>(rb == 0) : static_cast<uint8_t>(1);
    c.tb_part[cbg] = CrcMulMod(ra, xpow, c.poly_tb, c.len_tb);
}

__device__ inline void CrcFinishBlock(const CrcJob &c, uint32_t cbg, uint32_t rb, uint32_t ra)
{
    CrcFinishShifted(c, cbg, rb, ra, CrcTbShift(c, cbg));
}

}  // namespace airan
