// detect_demap.h -- the tail of the fused detector shared by the batched
// chain and the ragged multi-grant batch: per layer, the equalised symbol
// and no_eff rounded to fp16, no_eff floored, the QAM max-log LLRs
// (qam256_demod arithmetic), then layer demap and Gold descrambling into the
// codeword LLRs, one NL-element store per bit plane.
#pragma once

#include "detect.h"
#include "qam.h"

namespace airan {

__device__ inline int16_t SatNeg16(int v) { return static_cast<int16_t>(v == -32768 ? 32767 : -v); }

// xh / ne: the RE's LMMSE outputs; gw: the scrambling bits from this RE's
// first codeword bit on (bit l * qm + b for layer l, bit plane b); out: the
// RE's first codeword position in bit plane 0 (planes cw_stride apart).
// QM > 0 fixes the modulation at compile time.
template <int NL, int QM>
__device__ inline void DemapStore(const Cf (&xh)[4], const float (&ne)[4], uint32_t qm_rt, __half floor16,
                                  uint64_t gw, int16_t *out, uint32_t cw_stride)
{
    const uint32_t qm = QM > 0 ? static_cast<uint32_t>(QM) : qm_rt, m = qm / 2;
    const uint32_t g32 = static_cast<uint32_t>(gw);
    // LLRs of all layers first, then one contiguous NL-element store per bit
    // plane (the thread's REs are adjacent within a plane).
    int16_t v[8][NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const __half xr16 = __float2half_rn(xh[l].r), xi16 = __float2half_rn(xh[l].i);
        __half ne16 = __float2half_rn(ne[l]);
        if (__half2float(ne16) < __half2float(floor16)) ne16 = floor16;
        int llr[8];
        QamSymbolLlrs(qm, xr16, xi16, ne16, llr);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const uint32_t stream = (b & 1u) ? m + b / 2 : b / 2;
            const int x = b < qm ? llr[stream] : 0;
            // l qm + b < NL qm <= 32: the RE's bits are the low word's.
            const bool flip = (g32 >> (static_cast<uint32_t>(l) * qm + b)) & 1u;
            v[b][l] = flip ? SatNeg16(x) : static_cast<int16_t>(x);
        }
    }
    for (uint32_t b = 0; b < qm; ++b) {
        int16_t *o = out + static_cast<size_t>(b) * cw_stride;
        if constexpr (NL == 4) {
            *reinterpret_cast<uint2 *>(o) =
                make_uint2(static_cast<uint16_t>(v[b][0]) | static_cast<uint32_t>(static_cast<uint16_t>(v[b][1])) << 16,
                           static_cast<uint16_t>(v[b][2]) | static_cast<uint32_t>(static_cast<uint16_t>(v[b][3])) << 16);
        } else if constexpr (NL == 2) {
            *reinterpret_cast<uint32_t *>(o) =
                static_cast<uint16_t>(v[b][0]) | static_cast<uint32_t>(static_cast<uint16_t>(v[b][1])) << 16;
        } else {
#pragma unroll
            for (int l = 0; l < NL; ++l) o[l] = v[b][l];
        }
    }
}

}  // namespace airan
