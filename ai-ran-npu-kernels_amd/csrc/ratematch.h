// ratematch.h -- LDPC de-rate-matching of one decoder input position,
// shared by the standalone rate-dematch operator and the decoder that reads
// the codeword LLRs directly.
//
// fec/rate_dematch_mimo/rate_dematch_mimo_kernel.cpp:33-143 written as a
// gather: decoder position p of a code block holds 0 for the punctured 2Z
// systematic bits and for circular-buffer positions never transmitted,
// +5120 for filler bits, else the clipped sum of every transmitted bit that
// lands on it (kk, kk + NF, kk + 2 NF, ... < E with NF = Ncb - filler), each
// LLR scaled by 8 and clipped to +-5120.  Transmitted bit kk of the code
// block is bit b = kk / (E/Qm) of codeword symbol j = kk % (E/Qm) (the
// 38.212 bit interleaver), read from the bit-plane codeword layout
// [slot][Qm][codeword_stride].
#pragma once

#include "common.h"

namespace airan {

struct RateDev {
    uint32_t qm, cw_symbols, cw_stride, z, kprime, filler, ncb, n_cb;
};

// Per-code-block constants (uniform over the block's positions).
struct RateCb {
    uint32_t e, eq, q0, slot0, rem0;
    float rcp_eq;
};

__device__ inline RateCb RateCbSetup(const RateMatchDescriptor &d, const RateDev &r)
{
    RateCb c;
    c.e = d.e;
    c.eq = d.e / r.qm;
    const uint32_t fstart = r.kprime - 2u * r.z;
    c.q0 = d.k0 <= fstart ? d.k0 : d.k0 - r.filler;
    c.slot0 = d.cw_symbol_offset / r.cw_symbols;
    c.rem0 = d.cw_symbol_offset - c.slot0 * r.cw_symbols;
    c.rcp_eq = 1.0f / static_cast<float>(c.eq > 0 ? c.eq : 1u);
    return c;
}

// kk / eq (eq >= 1): the fp32 quotient is within kk * 2^-22 of the exact
// one (one step for kk < 2^22), corrected both ways.
__device__ inline uint32_t RateDivEq(uint32_t kk, const RateCb &c)
{
    uint32_t b = static_cast<uint32_t>(static_cast<float>(kk) * c.rcp_eq);
    while (b > 0 && b * c.eq > kk) --b;
    while ((b + 1u) * c.eq <= kk) ++b;
    return b;
}

// Every transmitted bit lands on its own circular-buffer position (no
// repetition: E <= Ncb - filler) and the block's codeword symbols lie in one
// slot: position p then takes at most one LLR (RateDematchSingle).
__device__ inline bool RateSingle(const RateCb &c, const RateDev &r)
{
    return c.e <= r.ncb - r.filler && c.rem0 + c.eq <= r.cw_symbols;
}

// RateDematchValue under RateSingle: the source element index of position p
// in the block's slot (planes [Qm][cw_stride] from the block's first
// symbol), or -1 for a zero LLR, -2 for a filler bit (+5120).
__device__ inline int32_t RateSingleIndex(uint32_t p, const RateCb &c, const RateDev &r)
{
    const uint32_t two_z = 2u * r.z;
    if (p < two_z || p >= two_z + r.ncb) return -1;
    const uint32_t k = p - two_z;
    const uint32_t fstart = r.kprime - two_z;
    if (k >= fstart && k < fstart + r.filler) return -2;
    const uint32_t nf = r.ncb - r.filler;
    const uint32_t q = k < fstart ? k : k - r.filler;
    const uint32_t kk = q >= c.q0 ? q - c.q0 : q + nf - c.q0;
    if (kk >= c.e) return -1;
    // kk < E < 2^22 (checked by the launcher): the fp32 quotient is within
    // one of kk / eq, so one correction each way.
    uint32_t b = static_cast<uint32_t>(static_cast<float>(kk) * c.rcp_eq);
    b = b * c.eq > kk ? b - 1u : b;
    b = (b + 1u) * c.eq <= kk ? b + 1u : b;
    return static_cast<int32_t>(b * r.cw_stride + (kk - b * c.eq));
}

__device__ inline int16_t RateDematchValue(uint32_t p, const RateCb &c, const RateDev &r,
                                           const int16_t *__restrict__ in)
{
    const uint32_t two_z = 2u * r.z;
    if (p < two_z || p >= two_z + r.ncb) return 0;
    const uint32_t k = p - two_z;
    const uint32_t fstart = r.kprime - two_z;
    if (k >= fstart && k < fstart + r.filler) return 5120;
    const uint32_t nf = r.ncb - r.filler;
    const uint32_t q = k < fstart ? k : k - r.filler;
    uint32_t kk = q >= c.q0 ? q - c.q0 : q + nf - c.q0;
    int32_t acc = 0;
    for (; kk < c.e; kk += nf) {
        const uint32_t b = RateDivEq(kk, c), j = kk - b * c.eq;
        uint32_t slot = c.slot0, sym = c.rem0 + j;
        while (sym >= r.cw_symbols) {
            sym -= r.cw_symbols;
            ++slot;
        }
        int32_t v = in[(static_cast<size_t>(slot) * r.qm + b) * r.cw_stride + sym] * 8;
        v = v > 5120 ? 5120 : (v < -5120 ? -5120 : v);
        acc += v;
    }
    return static_cast<int16_t>(acc > 5120 ? 5120 : (acc < -5120 ? -5120 : acc));
}

}  // namespace airan
