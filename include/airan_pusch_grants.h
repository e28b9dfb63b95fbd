/*
 * airan_pusch_grants.h -- ragged multi-grant batches on one GPU.
 *
 * No counterpart in the reference, whose chain serves one fixed profile
 * (chain/pusch_mimo_rx_chain/rx_chain.cpp:316-724).  A batch holds N grants
 * of different shapes -- PRB counts, ranks 1..4, modulation orders, code
 * rates, FFT sizes -- each its own cell and its own slot with one transport
 * block (the 192-byte PuschMimoRuntimeConfig of every grant compiled and
 * validated by the caller, see puschrx/grants.py), and runs the receive chain
 * over all of them with one launch per stage on work lists:
 *
 *   OFDM demod + RE demap   one launch per FFT size present
 *   DMRS gen, LS, noise     one launch each over (grant, layer / Rx) items
 *   LMMSE CE                one filter per distinct grant shape, built once
 *                           and cached for the process (keyed by the shape
 *                           and the CE prior); one launch per CE step over
 *                           every shape, each shape's grants together
 *   detect + demap + descramble   one launch over (grant, symbol, tile) items
 *   rate dematch            one launch over all code blocks
 *   LDPC (FP16 messages)    one launch over the code blocks with <= 8 active
 *                           rows, one over the rest
 *   CRC                     one launch over all code blocks + one per TB
 *
 * replayed from one HIP graph per batch.  Every stage computes what the
 * single-grant chain (airan_pusch_chain_*) computes for that grant, bit for
 * bit; only the indexing differs.
 */
#ifndef AIRAN_PUSCH_GRANTS_H
#define AIRAN_PUSCH_GRANTS_H

#include "airan_pusch_rx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct AiranGrantBatchOptionsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t iq_format;         /* AIRAN_IQ_INT16_Q8_8 or AIRAN_IQ_COMPLEX64 */
    uint32_t ldpc_iterations;   /* 1..64 */
    uint32_t early_stop;
    uint32_t ce_max_rank;       /* 1..512 */
    uint32_t use_graph;
    double tau_s;               /* CE prior of every grant: RMS delay spread */
    double fd_hz;               /* CE prior: maximum Doppler */
    double scs_hz;
    uint32_t detect_tile;       /* subcarriers per detector workgroup: 64, 128 or 256 (tuned) */
    uint32_t reserved[7];
} AiranGrantBatchOptionsV1;

/* Where grant g's data lives in the batch buffers. */
typedef struct AiranGrantLayoutV1 {
    uint64_t iq_offset;         /* complex samples into the IQ buffer: [NR][samples] of this grant */
    uint64_t iq_samples;        /* NR * samples */
    uint64_t bits_offset;       /* int8 elements into `bits`: [C][K] */
    uint64_t cw_offset;         /* int16 elements into `cw_llr`: [Qm][codeword_stride] */
    uint32_t first_cb;          /* index of its first code block in cb_ok / iterations */
    uint32_t num_cb;            /* C */
    uint32_t K;
    uint32_t fft;
} AiranGrantLayoutV1;

typedef struct AiranGrantBatchOutputsV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    uint32_t num_grants;
    const void *bits;           /* int8, grant g's [C][K] at its bits_offset */
    const void *tb_ok;          /* uint8 [grants] */
    const void *cb_ok;          /* uint8 [code blocks] */
    const void *iterations;     /* uint8 [code blocks] */
    const void *cw_llr;         /* int16: grant g's descrambled [Qm][codeword_stride] at its cw_offset */
    uint64_t num_cb;
} AiranGrantBatchOutputsV1;

typedef struct AiranGrantBatch AiranGrantBatch;

/* configs / plans: one per grant (plan->G = one slot's coded bits).  Builds
 * (or takes from the cache) the CE filter of every distinct grant shape and
 * every buffer; synchronises `stream`. */
int32_t airan_pusch_grants_Create(const PuschMimoConfig *configs, const AiranTbPlanV1 *plans, uint32_t num_grants,
                                  const AiranGrantBatchOptionsV1 *options, void *stream, AiranGrantBatch **batch);
void airan_pusch_grants_Destroy(AiranGrantBatch *batch);
/* layouts [num_grants]; *iq_samples_total: the IQ buffer's complex samples. */
int32_t airan_pusch_grants_Layout(const AiranGrantBatch *batch, AiranGrantLayoutV1 *layouts, uint32_t num_grants,
                                  uint64_t *iq_samples_total);
/* iq: device buffer of every grant's IQ at its layout offset.  Asynchronous
 * on `stream`; with use_graph the first call for (iq, stream) captures the
 * graph. */
int32_t airan_pusch_grants_Enqueue(AiranGrantBatch *batch, const void *iq, void *stream);
int32_t airan_pusch_grants_Outputs(const AiranGrantBatch *batch, AiranGrantBatchOutputsV1 *out);
/* Detector tile (64, 128 or 256 subcarriers per workgroup); drops the graph. */
int32_t airan_pusch_grants_SetDetectTile(AiranGrantBatch *batch, uint32_t tile);
/* Distinct grant shapes (CE filters) of the batch, and filters in the cache. */
uint32_t airan_pusch_grants_NumShapes(const AiranGrantBatch *batch);
uint32_t airan_pusch_grants_FilterCacheSize(void);
/* Frees every cached filter that no live batch uses. */
void airan_pusch_grants_ClearFilterCache(void);

#ifdef __cplusplus
}
#endif

#endif
