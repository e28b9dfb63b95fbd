/*
 * airan_pusch_chain.h -- the batched PUSCH receive chain on one GPU.
 *
 * Replaces the slot loop of chain/pusch_mimo_rx_chain/rx_chain.cpp:316-724
 * (Run: per slot OFDM demod -> RE demap -> DMRS gen -> DMRS LS -> LMMSE CE ->
 * detector I/O pack -> detection -> QAM demap -> layer demap, then per TB
 * descramble -> rate dematch -> LDPC) for grants of one shape: `num_slots`
 * slots carrying transport blocks of T consecutive slots each (T = 1: one TB
 * per slot; the reference's profile: T = 23, C = 143), run as one pipeline
 * of batched kernels that a HIP graph replays:
 *
 *   OFDM demod + RE demap     airan_ofdm_demod_redemap_Enqueue, all slots
 *   DMRS LS + noise           the mimo_dmrs_ls kernels, all slots
 *   LMMSE CE                  the channel_est_lmmse kernels, all slots
 *                             (receiver LS_ZF: LS + linear interpolation)
 *   detect+demap+descramble   one fused kernel: exact per-RE LMMSE (or
 *                             single-tap zero forcing), the
 *                             QAM demapper, layer demap and Gold descramble,
 *                             writing the descrambled codeword LLRs
 *   rate dematch + LDPC       FP16 messages with at most 8 active rows: one
 *                             decoder launch that de-rate-matches the
 *                             codeword LLRs while loading each code block
 *                             (the rate_dematch_mimo arithmetic); otherwise
 *                             the rate_dematch_mimo kernel then
 *                             airan_ldpc_decode_Enqueue, all code blocks
 *   CRC                       airan_crc_check_Enqueue, all transport blocks
 *                             (skipped for a plan without CRC, such as the
 *                             reference's fixed plan)
 *
 * Each slot's DMRS is that of its own slot number (c_init depends on it,
 * mimo_dmrs_gen.cpp:38-45; the reference re-derives its config per slot,
 * rx_chain.cpp:464): EnqueueSlots takes the slot number of each slot and the
 * pipeline generates each slot's DMRS from it.  The descrambling
 * sequence restarts with each TB and runs on across its slots
 * (descramble_mimo.cpp:176-184); code blocks may span slot boundaries
 * (rate_dematch_mimo.cpp:93-106).
 *
 * Every stage computes exactly what the per-operator entry points compute
 * (the fused stage is bit-identical to detect_io_pack + detect_lmmse +
 * no_eff floor + qam256_demod + layer_demap + descramble), so a slot decoded
 * here equals the same slot run operator by operator.  No counterpart to the
 * fused kernel or the batching in the reference.
 */
#ifndef AIRAN_PUSCH_CHAIN_H
#define AIRAN_PUSCH_CHAIN_H

#include "airan_ce_refiner.h"
#include "airan_pusch_rx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct AiranPuschChainOptionsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t max_slots;         /* slots per Enqueue, >= 1 */
    uint32_t iq_format;         /* AIRAN_IQ_INT16_Q8_8 or AIRAN_IQ_COMPLEX64 */
    uint32_t ldpc_iterations;   /* maximum iterations, 1..64 */
    uint32_t ldpc_mode;         /* AIRAN_LDPC_MODE_* */
    uint32_t early_stop;        /* stop a code block once its syndrome is zero */
    uint32_t ce_max_rank;       /* LMMSE filter eigenpairs, 1..512 */
    uint32_t use_graph;         /* replay the pipeline from a HIP graph */
    double tau_s;               /* CE frequency prior: RMS delay spread */
    double fd_hz;               /* CE time prior: maximum Doppler */
    double scs_hz;              /* subcarrier spacing */
    double eig_floor;           /* CE: drop eigenvalues below eig_floor * max */
    uint32_t receiver;          /* AIRAN_CHAIN_RX_* */
    uint32_t dft_precision;     /* AIRAN_DFT_* of the OFDM stage (FFT 1024..4096 for BF16 / FP16) */
    uint32_t gram_precision;    /* AIRAN_GRAM_* of the detector's Gram and matched filter */
    uint32_t reserved[5];
} AiranPuschChainOptionsV1;
/* Channel estimate and equaliser of the chain.  LMMSE: LMMSE CE (filter
 * built from tau_s, fd_hz) and per-RE LMMSE detection, any geometry.  LS_ZF:
 * the SISO receiver (one Rx, one layer on port 1000): DMRS least squares with
 * linear frequency/time interpolation (airan_channel_est_ls_siso) and
 * single-tap zero forcing (airan_equalize_zf_siso) with n0 = the slot's
 * DMRS-LS noise estimate; no CE filter is built. */
#define AIRAN_CHAIN_RX_LMMSE 0u
#define AIRAN_CHAIN_RX_LS_ZF 1u
/* Gram matrix H^H H and matched filter H^H y of the LMMSE detector.  FP32:
 * per-RE fp32 multiply-adds on the vector units.  FP16 / BF16: on the matrix
 * cores (v_mfma_f32_4x4x4 with 16 REs per instruction, the H and y samples
 * as fp16 -- exactly the fp16 values the chain stores -- or rounded to bf16,
 * fp32 accumulation), then the same fp32 LDL^H solve. */
#define AIRAN_GRAM_FP32 0u
#define AIRAN_GRAM_BF16 1u
#define AIRAN_GRAM_FP16 2u

typedef struct AiranPuschChain AiranPuschChain;

/* Device buffers of the last Enqueue (valid until the next one): bits int8
 * [TBs][C][K], tb_ok / cb_ok / iterations uint8 (TB-major; tb_ok and cb_ok
 * are NULL when the plan has no CRC), and the slot-major intermediates in
 * the operator layouts for inspection. */
typedef struct AiranPuschChainOutputsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t num_slots;
    const void *bits;           /* int8 [TBs][C][K] */
    const void *tb_ok;          /* uint8 [TBs], NULL without CRC */
    const void *cb_ok;          /* uint8 [TBs][C], NULL without CRC */
    const void *iterations;     /* uint8 [TBs][C] */
    const void *rx_grid_re;     /* fp16 [slots][NR][14][padded] */
    const void *rx_grid_im;
    const void *h_ls_re;        /* fp16 [slots][NR][L][D][pilot_pad] */
    const void *h_ls_im;
    const void *noise_ls;       /* fp16 [slots][NR] (DMRS LS) */
    const void *noise_ce;       /* fp16 [slots][NR] (CE subspace estimate); NULL for receiver LS_ZF */
    const void *h_re;           /* fp16 [slots][NR][L][14][padded] */
    const void *h_im;
    const void *cw_llr;         /* int16 [slots][Qm][codeword_stride], descrambled */
    const void *ldpc_llr;       /* int16 [TBs][C][68 Z]; NULL when the decoder reads cw_llr directly
                                   (FP16 messages and at most 8 active rows) */
} AiranPuschChainOutputsV1;

/* Builds the CE filter, the Gold bits and every buffer for max_slots slots,
 * then runs the pipeline once on zero slots (loading the per-device tables)
 * and synchronises `stream`.  plan->G must be T slots' coded bits
 * (T * codeword_symbols * Qm, T >= 1) and max_slots a multiple of T. */
int32_t airan_pusch_chain_Create(const PuschMimoConfig *config, const AiranTbPlanV1 *plan,
                                 const AiranPuschChainOptionsV1 *options, void *stream,
                                 AiranPuschChain **chain);
void airan_pusch_chain_Destroy(AiranPuschChain *chain);
/* iq: device [num_slots][NR][samples][2] (int16 Q8.8 or float per iq_format);
 * num_slots a multiple of T.  slot_numbers: device uint16 [num_slots], the
 * slot number of each slot (the DMRS c_init term of mimo_dmrs_gen.cpp:38-45),
 * read by the DMRS stage when the pipeline runs; the caller owns the buffer
 * and may rewrite it between calls.  Asynchronous on `stream`;
 * with use_graph, the first call for a given (iq, slot_numbers, num_slots,
 * stream) captures the graph. */
int32_t airan_pusch_chain_EnqueueSlots(AiranPuschChain *chain, const void *iq, uint32_t num_slots,
                                       const uint16_t *slot_numbers, void *stream);
/* EnqueueSlots with config->slot_number for every slot. */
int32_t airan_pusch_chain_Enqueue(AiranPuschChain *chain, const void *iq, uint32_t num_slots,
                                  void *stream);
/* T, the slots per transport block (0 on a bad argument). */
uint32_t airan_pusch_chain_SlotsPerTb(const AiranPuschChain *chain);
int32_t airan_pusch_chain_Outputs(const AiranPuschChain *chain, AiranPuschChainOutputsV1 *out);
/* Device bytes the chain holds for max_slots (0 on a bad argument). */
size_t airan_pusch_chain_DeviceBytes(const AiranPuschChain *chain);
/* Learned estimator (include/airan_ce_refiner.h; no reference counterpart):
 * after the LMMSE CE, every slot whose SNR estimate 10 log10(1 / mean noise
 * of the CE) lies in [snr_min_db, snr_max_db] gets the refiner's channel
 * estimate instead; the others keep LMMSE.  The refiner must hold at least
 * as many slots as each Enqueue; NULL detaches it.  Drops the cached graph.
 * The refiner's activations live in the refiner handle: a refiner attached
 * to two chains must not be used by both at once (one refiner per chain for
 * chains on concurrent streams). */
int32_t airan_pusch_chain_SetRefiner(AiranPuschChain *chain, AiranCeRefiner *refiner, float snr_min_db,
                                     float snr_max_db);

/* Stage timing.  With timing on, Enqueue launches the stages directly (no
 * graph) and records a HIP event on the stream before the first stage and
 * after each one; StageTimes waits for the last event and returns the
 * milliseconds of each stage of the last Enqueue, in pipeline order:
 * 0 OFDM, 1 DMRS LS, 2 CE, 3 detect+demap, 4 rate dematch, 5 LDPC, 6 CRC. */
#define AIRAN_CHAIN_STAGES 7
int32_t airan_pusch_chain_SetTiming(AiranPuschChain *chain, uint32_t enabled);
int32_t airan_pusch_chain_StageTimes(const AiranPuschChain *chain, float *ms, uint32_t count);

#ifdef __cplusplus
}
#endif

#endif
