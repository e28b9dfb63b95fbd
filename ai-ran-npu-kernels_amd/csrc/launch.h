// launch.h -- internal launchers shared by the per-operator entry points
// (one slot) and the batched receive chain (many slots per launch).
#pragma once

#include "common.h"

struct LmmseFilter;

namespace airan {

struct CrcJob;

// DMRS least squares + per-Rx noise for `slots` stacked grids
// (rx [slots,NR,14,pad] -> h_ls [slots,NR,L,D,pilot_pad], noise [slots,NR]).
// ref_re/ref_im: DMRS references [L][D][ref_pad], one per slot
// (ref_per_slot) or one shared by every slot.
int32_t LaunchDmrsLs(const Geom &g, const __half *rx_re, const __half *rx_im, const __half *ref_re,
                     const __half *ref_im, __half *h_re, __half *h_im, __half *noise, uint32_t slots,
                     hipStream_t s, bool ref_per_slot = false);
// DMRS references of `slots` slots [slots][L][D][ref_pad], slot s with the
// c_init of slot number slot_no[s] (device uint16) and the rest of cfg.
int32_t LaunchDmrsGenSlots(const PuschMimoConfig &cfg, const Geom &g, const uint16_t *slot_no, uint32_t slots,
                           __half *ref_re, __half *ref_im, hipStream_t s);

// OFDM demod as two DFT products on the matrix cores (op_ofdm_mfma.hip):
// precision AIRAN_DFT_BF16 or AIRAN_DFT_FP16, fft 1024 / 2048 / 4096;
// out_mode 0 stage-4 order, 1 the used-subcarrier grid.
bool MfmaDftSupported(uint32_t fft);
int32_t InitMfmaDftTables();
int32_t LaunchMfmaDft(const Geom &g, uint32_t precision, uint32_t iq_format, uint32_t batch, const void *iq,
                      int out_mode, __half *out_re, __half *out_im, hipStream_t s);

// SISO LS + linear interpolation (op_siso.hip) for `slots` stacked grids
// rx [slots,1,14,pad] -> H [slots,1,1,14,pad]; ref [1 or slots][D][ref_pad].
// AIRAN_UNSUPPORTED_PROFILE unless the config is 1 Rx, rank 1 on port 1000
// with one or two DMRS symbols.
int32_t LaunchLsInterpSiso(const PuschMimoConfig &cfg, const Geom &g, const __half *rx_re, const __half *rx_im,
                           const __half *ref_re, const __half *ref_im, bool ref_per_slot, __half *h_re,
                           __half *h_im, uint32_t slots, hipStream_t s);

// LMMSE channel estimate for `slots` stacked observations; output
// [slots,NR,out_layers,14,pad] (layers >= L untouched).
size_t CeWorkspaceBytes(const Geom &g, uint32_t slots);
// noise_in [slots,NR] is the fallback noise (DMRS-LS estimate); noise_out
// [slots,NR] (may be null) receives the estimate the filter used.
int32_t LaunchCeApply(const LmmseFilter *f, const __half *h_re, const __half *h_im, const __half *noise_in,
                      __half *noise_out, void *ws, size_t ws_bytes, __half *out_re, __half *out_im,
                      uint32_t out_layers, uint32_t slots, hipStream_t s);

// Several CE applications launched together, one launch per CE step (the
// grant shapes of a ragged batch): each job computes exactly what
// LaunchCeApply computes for it.  Jobs must not share a workspace.
struct CeGroupJob {
    const LmmseFilter *filter;
    const __half *hr, *hi, *noise_in;
    __half *noise_out, *out_re, *out_im;
    void *ws;
    size_t ws_bytes;
    uint32_t out_layers, slots;
};
struct CeGrouped;
int32_t CreateCeGrouped(const CeGroupJob *jobs, uint32_t n, CeGrouped **out);
int32_t LaunchCeGrouped(const CeGrouped *g, hipStream_t s);
void DestroyCeGrouped(CeGrouped *g);
// The filter's geometry must match `g` (same grant shape).
bool CeFilterMatches(const LmmseFilter *f, const Geom &g);

// Rate dematch of num_cb code blocks described by device descriptors
// (cw_symbol_offset counts from the start of the stacked codeword buffer).
int32_t LaunchRateDematch(const Geom &g, const AiranTbPlanV1 &plan, const RateMatchDescriptor *desc,
                          const int16_t *in, int16_t *out, uint32_t num_cb, hipStream_t s);

// LDPC decode (FP16 messages) of num_cb code blocks read straight from the
// bit-plane codeword LLRs through the rate-match descriptors: the rate
// dematch and decoder stages in one launch.  AIRAN_UNSUPPORTED_PROFILE when
// the active rows exceed the fused decoder's limit (8) or Z is not 24 * 2^j.
// queue: null, or two device words, zero before the first launch (each
// launch leaves them zero): the decoder's workgroups then take code blocks
// from it as they free up instead of walking fixed lists (same outputs,
// better balance).
int32_t LaunchLdpcFromCodeword(const Geom &g, const AiranTbPlanV1 &plan, const RateMatchDescriptor *desc,
                               const int16_t *cw, uint32_t num_cb, uint32_t rows, uint32_t iters, bool early_stop,
                               int8_t *bits, uint8_t *iters_out, uint32_t *queue,
                               hipStream_t s);

// CRC checks of decoded bits (crc.h): LaunchCbCrc checks each code block's
// CRC24B (cb_ok) and stores its part of its transport block's CRC remainder
// (tb_part, one word per block), which LaunchCrcFinish turns into tb_ok.
// bits: int8 0/1, [num_cb][K].
int32_t CrcJobSetup(const AiranTbPlanV1 &plan, uint8_t *cb_ok, uint32_t *tb_part, CrcJob *out);
// x^((C - 1 - r) L) mod g_TB: the shift of code block r's remainder into its
// transport block's (a per-plan table for CrcJob::xpow).
uint32_t CrcXpowHost(const AiranTbPlanV1 &plan, uint32_t r);
// Zeroes n device words.
void LaunchCrcZero(uint32_t *words, uint32_t n, hipStream_t s);
// The check's per-plan tables (kCrcTableBytes of 16-byte aligned device
// memory, built by LaunchCrcTables before LaunchCbCrc reads them): the 0/1
// term matrix in MFMA fragment order, then the VALU path's masks.
constexpr size_t kCrcFragBytes = 2 * (8448 / 32) * 64 * 16;
constexpr size_t kCrcTableBytes = kCrcFragBytes + 2 * 24 * 384 * sizeof(uint32_t);
// Code blocks per TB the CRC workspace is sized for (TS 38.214 TB sizes
// need at most 152).
constexpr uint32_t kCrcMaxCb = 256;
// Byte offset of the tables in airan_crc_check_WorkspaceBytes(num_tbs):
// after the num_tbs * kCrcMaxCb remainder parts.
inline size_t CrcTableOffset(uint32_t num_tbs)
{
    return (static_cast<size_t>(num_tbs) * kCrcMaxCb * sizeof(uint32_t) + 255) / 256 * 256;
}
void LaunchCrcTables(const CrcJob &c, uint32_t K, void *tables, hipStream_t s);
void LaunchCbCrc(const int8_t *bits, const CrcJob &c, uint32_t K, uint32_t num_cb, const void *tables,
                 hipStream_t s);
void LaunchCrcFinish(const uint32_t *tb_part, uint8_t *tb_ok, uint32_t num_tbs, uint32_t C, hipStream_t s);

// ---------------------------------------------------- ragged multi-grant batch
// (ragged.h; every launch computes what the single-grant kernels compute.)
struct RaggedGrant;
struct RaggedStream;
struct RaggedItem;
struct RaggedCb;
// OFDM demod + RE demap of num_streams streams of one FFT size.
int32_t LaunchOfdmRagged(uint32_t fft, uint32_t iq_format, const RaggedStream *streams, uint32_t num_streams,
                         const void *iq, __half *out_re, __half *out_im, hipStream_t s);
// DMRS references (items (grant, layer)), DMRS LS (items (grant, rx, layer))
// and noise (items (grant, rx)); the max_* bound the grids.
int32_t LaunchDmrsRagged(const RaggedGrant *grants, const RaggedItem *gen_items, uint32_t n_gen,
                         const RaggedItem *ls_items, uint32_t n_ls, const RaggedItem *noise_items, uint32_t n_noise,
                         uint32_t max_ref_pad, uint32_t max_pilot_pad, uint32_t max_dmrs, const __half *rx_re,
                         const __half *rx_im, __half *ref_re, __half *ref_im, __half *h_re, __half *h_im,
                         __half *noise, hipStream_t s);
// c_init of DMRS symbol `symbol` of cfg (mimo_dmrs_gen.cpp:38-45).
uint32_t HostDmrsCinit(const PuschMimoConfig &cfg, uint32_t symbol);
// Rate dematch of every code block into its decoder input (lam_off).
int32_t LaunchRateDematchRagged(const RaggedGrant *grants, const RaggedCb *cbs, uint32_t num_cb, const int16_t *cw,
                                int16_t *out, hipStream_t s);
// LDPC (FP16 messages) of the listed code blocks: reg_list with <= 8 rows
// (messages in registers), ws_list with more (messages in ws at ws_off).
int32_t LaunchLdpcRagged(const RaggedCb *cbs, const uint32_t *reg_list, uint32_t n_reg, const uint32_t *ws_list,
                         uint32_t n_ws, uint32_t iters, bool early_stop, const int16_t *lam, int8_t *bits,
                         uint8_t *iters_out, uint2 *ws, hipStream_t s);
// CRC24B of every code block and the CRC of every TB (tb_first / tb_count
// index the blocks of each TB).
int32_t LaunchCrcRagged(const RaggedCb *cbs, uint32_t num_cb, const int8_t *bits, uint8_t *cb_ok, uint32_t *tb_part,
                        const uint32_t *tb_first, const uint32_t *tb_count, uint8_t *tb_ok, uint32_t num_tbs,
                        hipStream_t s);
// The TB CRC polynomial (without its top term) of a plan's crc_tb length.
uint32_t CrcTbPoly(uint32_t crc_tb);

}  // namespace airan
