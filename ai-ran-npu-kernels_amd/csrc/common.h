// common.h -- internal helpers shared by the gfx950 PUSCH receive operators.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>

#include "airan_pusch_rx.h"

namespace airan {

constexpr uint32_t kMaxRx = 64;
constexpr uint32_t kMaxLayers = 4;
constexpr uint32_t kDetectLayers = 16;
constexpr uint32_t kSymbols = 14;

constexpr uint32_t AlignUp(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

// Per-device FFT twiddle table (op_ofdm.hip); idempotent.
int InitOfdmTables();
// Per-device address of the decoder's no-alloc message array (op_ldpc.hip);
// idempotent.
int32_t InitLdpcScratch();

// Geometry derived from a PuschMimoConfig; one definition for every operator
// so that SameLayout checks agree across the chain.
struct Geom {
    uint32_t nr, nl, qm, nsym, fft, used, pad;
    uint32_t ndmrs, dmrs_sym[4];
    uint32_t ndata, data_sym[14];
    uint32_t n_dmrs_re, ref_pad, pilot_pad, n_occ;
    uint32_t data_re, data_stride, qam_row, qam_plane, cw_symbols, cw_stride;
    uint32_t P, Q;
    uint32_t cp_long, cp_short, samples, sym_start[14];
    uint32_t comb[4], occ[4];  // per layer: comb offset, FD-OCC2 (two ports share the comb)
};

// Checks the geometry the operators of this library support (any
// PRB count, FFT 128..4096 power of two, 1..64 Rx, rank 1..4, Qm 2..8,
// type-1 single-symbol DMRS on ports 1000..1003, 1..4 DMRS symbols, full
// 14-symbol allocation, no transform precoding) and fills `g`.
inline int32_t DeriveGeom(const PuschMimoConfig &c, Geom *g)
{
    if (c.abi_version != AIRAN_PUSCH_MIMO_ABI_VERSION || c.struct_size < sizeof(PuschMimoConfig))
        return AIRAN_INVALID_ARGUMENT;
    for (uint32_t v : c.reserved)
        if (v != 0) return AIRAN_UNSUPPORTED_PROFILE;
    if (c.num_layers < 1 || c.num_layers > kMaxLayers || c.num_tx_ports < c.num_layers ||
        c.num_rx_antennas < 1 || c.num_rx_antennas > kMaxRx ||
        (c.qm != 2 && c.qm != 4 && c.qm != 6 && c.qm != 8) || c.num_symbols != kSymbols ||
        c.fft_size < 128 || c.fft_size > 4096 || (c.fft_size & (c.fft_size - 1)) != 0 ||
        c.num_rb < 1 || c.num_rb > 275 || c.used_subcarriers != 12u * c.num_rb ||
        c.padded_subcarriers < c.used_subcarriers || c.padded_subcarriers > c.fft_size ||
        c.start_symbol != 0 || c.num_allocated_symbols != kSymbols || c.dmrs_type != 1 ||
        c.dmrs_length != 1 || c.transform_precoding != 0 || c.n_scid > 1)
        return AIRAN_UNSUPPORTED_PROFILE;
    std::memset(g, 0, sizeof(*g));
    g->nr = c.num_rx_antennas;
    g->nl = c.num_layers;
    g->qm = c.qm;
    g->nsym = c.num_symbols;
    g->fft = c.fft_size;
    g->used = c.used_subcarriers;
    g->pad = c.padded_subcarriers;
    for (uint32_t s = 0; s < g->nsym; ++s) {
        if ((c.dmrs_symbol_mask >> s) & 1u) {
            if (g->ndmrs >= 4) return AIRAN_UNSUPPORTED_PROFILE;
            g->dmrs_sym[g->ndmrs++] = s;
        } else {
            g->data_sym[g->ndata++] = s;
        }
    }
    if (g->ndmrs < 1 || (c.dmrs_symbol_mask >> kSymbols) != 0) return AIRAN_UNSUPPORTED_PROFILE;
    for (uint32_t l = 0; l < g->nl; ++l) {
        uint16_t p = c.dmrs_ports[l];
        if (p < 1000 || p > 1003) return AIRAN_UNSUPPORTED_PROFILE;
        for (uint32_t q = 0; q < l; ++q)
            if (c.dmrs_ports[q] == p) return AIRAN_UNSUPPORTED_PROFILE;
        g->comb[l] = (p - 1000u) / 2u;
    }
    for (uint32_t l = 0; l < g->nl; ++l) {
        uint32_t share = 0;
        for (uint32_t q = 0; q < g->nl; ++q) share += g->comb[q] == g->comb[l];
        g->occ[l] = share == 2;
    }
    g->n_dmrs_re = g->used / 2;
    g->ref_pad = AlignUp(g->n_dmrs_re, 128);
    g->pilot_pad = AlignUp(g->n_dmrs_re, 64);
    g->n_occ = g->used / 4;
    g->data_re = g->ndata * g->used;
    g->data_stride = AlignUp(g->data_re, 64);
    g->qam_row = AlignUp(g->used, 64);
    g->qam_plane = g->ndata * g->qam_row;
    g->cw_symbols = g->nl * g->data_re;
    g->cw_stride = g->nl * g->data_stride;
    g->Q = 64;
    g->P = g->fft / 64;
    g->cp_short = g->fft * 9 / 128;
    g->cp_long = g->cp_short + g->fft / 64;
    uint32_t pos = 0;
    for (uint32_t s = 0; s < g->nsym; ++s) {
        pos += s == 0 ? g->cp_long : g->cp_short;
        g->sym_start[s] = pos;
        pos += g->fft;
    }
    g->samples = pos;
    return AIRAN_OK;
}

inline void FillLayout(const Geom &g, PuschMimoLayout *layout)
{
    layout->num_dmrs_symbols = static_cast<uint16_t>(g.ndmrs);
    layout->num_data_symbols = static_cast<uint16_t>(g.ndata);
    layout->num_data_re = g.data_re;
    layout->data_stride = g.data_stride;
    layout->codeword_symbols = g.cw_symbols;
    layout->codeword_stride = g.cw_stride;
}

inline bool SameLayout(const PuschMimoLayout &a, const PuschMimoLayout &b)
{
    return a.num_dmrs_symbols == b.num_dmrs_symbols && a.num_data_symbols == b.num_data_symbols &&
           a.num_data_re == b.num_data_re && a.data_stride == b.data_stride &&
           a.codeword_symbols == b.codeword_symbols && a.codeword_stride == b.codeword_stride;
}

// Common OpArgs header check: ABI version, struct size, non-null config.
template <typename T>
inline bool HeaderOk(const T *args, size_t min_size)
{
    return args != nullptr && args->abi_version == AIRAN_PUSCH_MIMO_ABI_VERSION &&
           args->struct_size >= min_size;
}

inline int32_t LaunchStatus()
{
    return hipGetLastError() == hipSuccess ? AIRAN_OK : AIRAN_LAUNCH_FAILED;
}

inline uint32_t CeilDiv(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

// fp16 arithmetic with one rounding per operation, as the reference's fp16
// vector units compute: contraction into v_fma_f16 would skip a product's
// rounding and differ by one ulp.
__device__ inline __half HMul(__half a, __half b)
{
#pragma clang fp contract(off)
    return __float2half_rn(__half2float(a) * __half2float(b));
}
__device__ inline __half HAdd(__half a, __half b)
{
#pragma clang fp contract(off)
    return __float2half_rn(__half2float(a) + __half2float(b));
}
__device__ inline __half HSub(__half a, __half b)
{
#pragma clang fp contract(off)
    return __float2half_rn(__half2float(a) - __half2float(b));
}

// Kernel-argument copy of the geometry (trivially copyable).
struct DevGeom {
    uint32_t nr, nl, qm, nsym, fft, used, pad;
    uint32_t ndmrs, ndata;
    uint32_t dmrs_sym[4];
    uint32_t data_sym[14];
    uint32_t n_dmrs_re, ref_pad, pilot_pad, n_occ;
    uint32_t data_re, data_stride, qam_row, qam_plane, cw_symbols, cw_stride;
    uint32_t P, Q;
    uint32_t comb[4], occ[4];
};

inline DevGeom ToDev(const Geom &g)
{
    DevGeom d{};
    d.nr = g.nr; d.nl = g.nl; d.qm = g.qm; d.nsym = g.nsym; d.fft = g.fft;
    d.used = g.used; d.pad = g.pad; d.ndmrs = g.ndmrs; d.ndata = g.ndata;
    for (int i = 0; i < 4; ++i) { d.dmrs_sym[i] = g.dmrs_sym[i]; d.comb[i] = g.comb[i]; d.occ[i] = g.occ[i]; }
    for (int i = 0; i < 14; ++i) d.data_sym[i] = g.data_sym[i];
    d.n_dmrs_re = g.n_dmrs_re; d.ref_pad = g.ref_pad; d.pilot_pad = g.pilot_pad; d.n_occ = g.n_occ;
    d.data_re = g.data_re; d.data_stride = g.data_stride; d.qam_row = g.qam_row;
    d.qam_plane = g.qam_plane; d.cw_symbols = g.cw_symbols; d.cw_stride = g.cw_stride;
    d.P = g.P; d.Q = g.Q;
    return d;
}

}  // namespace airan
