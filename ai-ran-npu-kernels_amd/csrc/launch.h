// launch.h -- internal launchers shared by the per-operator entry points
// (one slot) and the batched receive chain (many slots per launch).
#pragma once

#include "common.h"

struct LmmseFilter;

namespace airan {

struct CrcFuse;

// DMRS least squares + per-Rx noise for `slots` stacked grids
// (rx [slots,NR,14,pad] -> h_ls [slots,NR,L,D,pilot_pad], noise [slots,NR]).
int32_t LaunchDmrsLs(const Geom &g, const __half *rx_re, const __half *rx_im, const __half *ref_re,
                     const __half *ref_im, __half *h_re, __half *h_im, __half *noise, uint32_t slots,
                     hipStream_t s);

// LMMSE channel estimate for `slots` stacked observations; output
// [slots,NR,out_layers,14,pad] (layers >= L untouched).
size_t CeWorkspaceBytes(const Geom &g, uint32_t slots);
// noise_in [slots,NR] is the fallback noise (DMRS-LS estimate); noise_out
// [slots,NR] (may be null) receives the estimate the filter used.
int32_t LaunchCeApply(const LmmseFilter *f, const __half *h_re, const __half *h_im, const __half *noise_in,
                      __half *noise_out, void *ws, size_t ws_bytes, __half *out_re, __half *out_im,
                      uint32_t out_layers, uint32_t slots, hipStream_t s);
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
// With crc non-null the decoder also checks each code block's CRC24B
// (cb_ok) and accumulates the transport blocks' CRC remainders (tb_acc,
// zeroed by LaunchCrcZero before, turned into tb_ok by LaunchCrcFinish after).
// queue: null, or one device word zeroed before the launch (LaunchCrcZero
// with n = 1): the decoder's workgroups then take code blocks from it as
// they free up instead of walking fixed lists (same outputs, better balance).
int32_t LaunchLdpcFromCodeword(const Geom &g, const AiranTbPlanV1 &plan, const RateMatchDescriptor *desc,
                               const int16_t *cw, uint32_t num_cb, uint32_t rows, uint32_t iters, bool early_stop,
                               int8_t *bits, uint8_t *iters_out, const CrcFuse *crc, uint32_t *queue,
                               hipStream_t s);

// CRC check of decoded bits inside the decoder (crc.h).
int32_t CrcFuseSetup(const AiranTbPlanV1 &plan, uint8_t *cb_ok, uint32_t *tb_acc, CrcFuse *out);
// x^((C - 1 - r) L) mod g_TB: the shift of code block r's remainder into its
// transport block's (a per-plan table for CrcFuse::xpow).
uint32_t CrcXpowHost(const AiranTbPlanV1 &plan, uint32_t r);
void LaunchCrcZero(uint32_t *tb_acc, uint32_t num_tbs, hipStream_t s);
void LaunchCrcFinish(const uint32_t *tb_acc, uint8_t *tb_ok, uint32_t num_tbs, hipStream_t s);

}  // namespace airan
