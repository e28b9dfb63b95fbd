// crc.h -- CRC checks on decoded bits shared by the CRC operator and the
// decoder that checks its code blocks as it writes them (TS 38.212 5.1).
//
// The remainder of a message is linear in its bits: for an n-bit message
// (data followed by its parity) it is the XOR over set bits j of
// x^(L + n - 1 - j) mod g, with the powers tabulated per polynomial.  A
// transport block's remainder is the XOR over its code blocks of each code
// block's local remainder times x^(bits after it) mod g.
#pragma once

#include "common.h"

namespace airan {

// Everything a decoder needs to check its code blocks in its epilogue.
struct CrcFuse {
    const uint32_t *pow_b;   // x^(24 + e) mod g_CRC24B
    const uint32_t *pow_tb;  // x^(len + e) mod g_TB (CRC24A or CRC16)
    uint8_t *cb_ok;          // [blocks]; nullptr: no check
    uint32_t *tb_acc;        // [TBs] XOR accumulators, zero before the decoder
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

// Contribution of bit j (set) of code block r to its CRC24B remainder (rb)
// and to its part of the TB remainder before the shift (ra).
__device__ inline void CrcAddBit(const CrcFuse &c, uint32_t j, uint32_t &rb, uint32_t &ra)
{
    if (j >= c.kprime) return;
    if (c.crc_cb) rb ^= c.pow_b[c.kprime - 1 - j];
    if (j < c.L) ra ^= c.pow_tb[c.L - 1 - j];
}

// Called by one thread per code block with the block-reduced remainders.
__device__ inline void CrcFinishBlock(const CrcFuse &c, uint32_t cbg, uint32_t rb, uint32_t ra)
{
    const uint32_t tb = cbg / c.C, r = cbg - tb * c.C;
    c.cb_ok[cbg] = c.crc_cb ? static_cast<uint8_t>(rb == 0) : static_cast<uint8_t>(1);
    uint32_t xpow = 1u;
    if (c.xpow != nullptr) {
        xpow = c.xpow[r];
    } else {
        // Square-and-multiply from x^1 (serial: a table avoids it).
        uint32_t base = 2u, e = (c.C - 1 - r) * c.L;
        while (e != 0) {
            if (e & 1u) xpow = CrcMulMod(xpow, base, c.poly_tb, c.len_tb);
            base = CrcMulMod(base, base, c.poly_tb, c.len_tb);
            e >>= 1;
        }
    }
    atomicXor(&c.tb_acc[tb], CrcMulMod(ra, xpow, c.poly_tb, c.len_tb));
}

}  // namespace airan
