/*
 * airan_pusch_tx.h -- synthetic PUSCH transmitter for loopback tests and
 * benchmarks: the exact inverse of the receive chain's bit-level stages.
 *
 * No counterpart in the receive-chain operator set; it stands in for the
 * reference's test-vector generators (chain/pusch_mimo_rx_chain/scripts and
 * the per-operator gen_data.py scripts), which produce the TX side of every
 * golden input.
 *
 * Per transport block: CRC attach (CRC24A, or CRC16 for small TBs; none for
 * the reference's fixed plan), code-block segmentation with CRC24B, filler
 * bits, LDPC base graph 1 encoding, rate matching from k0 skipping filler,
 * bit interleaving, Gold scrambling, QAM mapping with the reference mapper's
 * polarity (mapping/qam_mod_256_siso/scripts/qam256_mod_ref.py:38-45: TS
 * 38.211 5.1 Gray mapping with every bit inverted; unit average power),
 * layer mapping (TS 38.211 7.3.1.3) and the resource grid with type-1 DMRS
 * (|r| = 1, FD-OCC on ports 1001/1003).  Host code, and (airan_pusch_txgpu_*)
 * the same on the device.
 */
#ifndef AIRAN_PUSCH_TX_H
#define AIRAN_PUSCH_TX_H

#include "airan_pusch_rx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Encodes num_tbs transport blocks, each spanning slots_per_tb consecutive
 * slots of `config` (plan->G == slots_per_tb * codeword_symbols * Qm).
 * payload: [num_tbs][plan->tbs] bits, one bit per byte (0/1).
 * grid: complex64 [num_tbs * slots_per_tb][L][14][used_subcarriers]
 * (interleaved re/im); every slot carries DMRS of config->slot_number.
 * Scrambling restarts at Gold index 0 for every TB and runs on across the
 * TB's slots.  coded_bits (optional): [num_tbs][G] scrambled bits in
 * codeword order, one per byte. */
int32_t airan_pusch_tx_EncodeTbs(const PuschMimoConfig *config, const AiranTbPlanV1 *plan,
                                 uint32_t num_tbs, uint32_t slots_per_tb, const uint8_t *payload,
                                 float *grid, uint8_t *coded_bits);
/* As EncodeTbs, slot s of the batch (TB s / slots_per_tb) carrying the DMRS
 * of slot number slot_numbers[s] (host uint16 [num_tbs * slots_per_tb];
 * NULL: config->slot_number for every slot). */
int32_t airan_pusch_tx_EncodeTbsSlots(const PuschMimoConfig *config, const AiranTbPlanV1 *plan,
                                      uint32_t num_tbs, uint32_t slots_per_tb,
                                      const uint16_t *slot_numbers, const uint8_t *payload,
                                      float *grid, uint8_t *coded_bits);

/* The transmitter on the device, and a fading channel to IQ (no
 * counterpart in the reference: its test vectors come from host scripts).
 * EncodeTbs computes exactly what airan_pusch_tx_EncodeTbsSlots computes,
 * bit for bit, for transport blocks whose payload is already in device
 * memory; Channel passes the grid through a tapped-delay-line channel and
 * OFDM-modulates it to the chain's IQ input.  One handle per grant shape
 * (config + plan) and batch size; calls on one handle must not overlap. */
typedef struct AiranPuschTxGpu AiranPuschTxGpu;
/* plan->G == slots_per_tb * codeword_symbols * Qm; max_tbs TBs per call. */
int32_t airan_pusch_txgpu_Create(const PuschMimoConfig *config, const AiranTbPlanV1 *plan, uint32_t max_tbs,
                                 uint32_t slots_per_tb, AiranPuschTxGpu **tx);
void airan_pusch_txgpu_Destroy(AiranPuschTxGpu *tx);
/* payload: device uint8 [num_tbs][plan->tbs] (0/1); slot_numbers: device
 * uint16 [num_tbs * slots_per_tb] or NULL (config->slot_number); grid:
 * device complex64 [num_tbs * slots_per_tb][L][14][used], every RE written;
 * coded_bits: device uint8 [num_tbs][G] scrambled bits in codeword order, or
 * NULL.  Asynchronous on `stream`. */
int32_t airan_pusch_txgpu_EncodeTbs(AiranPuschTxGpu *tx, const uint8_t *payload, uint32_t num_tbs,
                                    const uint16_t *slot_numbers, float *grid, uint8_t *coded_bits, void *stream);
/* grid (device complex64 [slots][L][14][used]) through the channel
 * H[rx][l][sym][u] = sum_t gains[slot][rx][l][t][sym] e^(-j 2 pi f_u delays[slot][t]),
 * f_u = (u - used/2) scs_hz (gains: device complex64 [slots][NR][L][taps][14],
 * taps <= 32; delays: device float [slots][taps] in seconds), then OFDM
 * modulation (subcarrier u on bin (u - used/2) mod N, unitary IFFT, the NR
 * cyclic prefixes) and complex AWGN of variance noise_var per sample from a
 * counter-based generator keyed by `seed`.  iq: device [slots][NR][samples][2],
 * int16 Q8.8 (rounded, saturated) or float per iq_format. */
int32_t airan_pusch_txgpu_Channel(AiranPuschTxGpu *tx, const float *grid, uint32_t slots, const float *gains,
                                  const float *delays, uint32_t taps, float scs_hz, float noise_var, uint64_t seed,
                                  uint32_t iq_format, void *iq, void *stream);

/* LDPC base graph 1 systematic encoder: info [22Z] bits -> codeword [68Z]
 * bits (columns 0..67; filler positions must be 0 in info). */
int32_t airan_ldpc_bg1_Encode(uint32_t z, const uint8_t *info, uint8_t *codeword);

/* CRC remainder of `nbits` bits (MSB-first, one per byte).  kind: 0 =
 * CRC24A, 1 = CRC24B, 2 = CRC16 (TS 38.212 5.1). */
uint32_t airan_crc_Compute(uint32_t kind, const uint8_t *bits, uint64_t nbits);

#ifdef __cplusplus
}
#endif

#endif
