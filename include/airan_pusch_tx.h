/*
 * airan_pusch_tx.h -- synthetic PUSCH transmitter for loopback tests and
 * benchmarks: the exact inverse of the receive chain's bit-level stages.
 *
 * No counterpart in the receive-chain operator set; it stands in for the
 * reference's test-vector generators (chain/pusch_mimo_rx_chain/scripts and
 * the per-operator gen_data.py scripts), which produce the TX side of every
 * golden input.  Host code, no GPU needed.
 *
 * Per transport block: CRC attach (CRC24A, or CRC16 for small TBs; none for
 * the reference's fixed plan), code-block segmentation with CRC24B, filler
 * bits, LDPC base graph 1 encoding, rate matching from k0 skipping filler,
 * bit interleaving, Gold scrambling, QAM mapping with the reference mapper's
 * polarity (mapping/qam_mod_256_siso/scripts/qam256_mod_ref.py:38-45: TS
 * 38.211 5.1 Gray mapping with every bit inverted; unit average power),
 * layer mapping (TS 38.211 7.3.1.3) and the resource grid with type-1 DMRS
 * (|r| = 1, FD-OCC on ports 1001/1003).
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
