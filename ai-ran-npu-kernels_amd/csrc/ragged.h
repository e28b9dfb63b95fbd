// ragged.h -- per-grant records and work items of a ragged multi-grant batch
// (airan_pusch_grants_*): grants of different PRB counts, ranks, modulation
// orders and code rates, each its own cell and slot, processed together by
// work-list kernels (one launch per stage, per FFT size for the OFDM stage).
// Every stage computes exactly what the single-grant kernels compute; only
// the indexing differs.
#pragma once

#include "common.h"

namespace airan {

// One grant: its geometry and where its slices of the batch buffers start.
struct RaggedGrant {
    DevGeom g;
    uint32_t dmrs_cinit[4];   // c_init of each DMRS symbol (mimo_dmrs_gen.cpp:38-45)
    uint32_t wf_odd[4];       // w_f = (-1)^m on ports 1001 / 1003
    uint32_t dmrs_off;        // 6 * rb_start (Gold offset of the first pilot)
    uint32_t gold_off;        // words: the grant's data scrambling bits
    uint32_t floor16;         // fp16 bits of the no_eff floor of its Qm
    uint32_t noise_off;       // [NR] fp16 noise estimates
    uint64_t iq_off;          // complex samples: [NR][samples]
    uint64_t rx_off;          // fp16: [NR][14][pad]
    uint64_t ref_off;         // fp16: [L][D][ref_pad]
    uint64_t hls_off;         // fp16: [NR][L][D][pilot_pad]
    uint64_t h_off;           // fp16: [NR][L][14][pad]
    uint64_t cw_off;          // int16: [Qm][codeword_stride]
};

// OFDM: one antenna stream of one grant (all streams of a launch share the
// FFT size).
struct RaggedStream {
    uint64_t iq_off;   // complex samples
    uint64_t out_off;  // fp16 elements of the [14][pad] grid
    uint32_t used, pad;
};

// DMRS generation / LS / noise items.
struct RaggedItem {
    uint32_t grant, a, b;  // DMRS gen: (grant, layer); LS: (grant, rx, layer); noise: (grant, rx)
};

// Detector tile: one data symbol of one grant, `tile` subcarriers from sc0.
struct RaggedTile {
    uint32_t grant, ds, sc0;
};

// Rate dematch + LDPC + CRC: one code block.
struct RaggedCb {
    uint32_t grant;
    uint32_t e, k0, ncb, cw_symbol_offset;  // rate matching (within the grant's codeword)
    uint32_t z, kprime, filler, k, rows;    // decoder
    uint64_t lam_off;                       // int16 [68 z]
    uint64_t bits_off;                      // int8 [k]
    uint64_t ws_off;                        // decoder message state (uint2), rows > 8
    uint32_t crc_cb, len_tb, poly_tb, xpow; // CRC24B length (0 / 24), TB CRC and this block's shift
};

}  // namespace airan
