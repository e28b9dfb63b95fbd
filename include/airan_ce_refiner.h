/*
 * airan_ce_refiner.h -- learned channel-estimation refiner (no counterpart
 * in the reference): a convolutional network that maps the DMRS
 * least-squares estimate, linearly interpolated onto the 14 x used grid, to
 * the full-grid channel, with BF16 training and inference kernels.
 *
 * Network (4 Rx x 4 layers): input X [slots][14][used][32] bf16, channel
 * c = 2 (rx * 4 + layer) + {0: Re, 1: Im}; six 3x3 convolutions (stride 1,
 * zero padding, optionally dilated along the subcarriers:
 * airan_ce_refiner_SetDilation) 32 -> 64 -> 64 -> 64 -> 64 -> 64 -> 32 with ReLU after the
 * first five, and a residual: H = X + s conv6(...) (s = 1, or the slot's
 * noise standard deviation: airan_ce_refiner_SetResidualScale), X added
 * in fp32.  Weights and activations
 * bf16, accumulation fp32 (MFMA), master weights fp32 trained with Adam on
 * the device.  Parameters, in order per layer: weights [9][cout][cin]
 * (tap = 3 kh + kw over symbol offset kh - 1 and subcarrier offset kw - 1),
 * then bias [cout]; AIRAN_REFINER_PARAMS floats in all.
 *
 * Training data is synthesised on the device: true channels from per-slot
 * tapped-delay-line gains, the DMRS symbols of the received grid with
 * AWGN, the library's DMRS least-squares kernels, then the interpolation
 * that forms X; the loss is the mean squared error against the true channel.
 * In input mode 1 (airan_ce_refiner_SetInput) X is instead the LMMSE
 * estimate (the channel_est_lmmse output), so the network learns a residual
 * correction of it.
 */
#ifndef AIRAN_CE_REFINER_H
#define AIRAN_CE_REFINER_H

#include <stddef.h>
#include <stdint.h>

#include "airan_pusch_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AIRAN_REFINER_PARAMS 184672u
#define AIRAN_REFINER_CHANNELS 32u

typedef struct AiranCeRefiner AiranCeRefiner;

/* Refiner for grants of `config`'s shape (4 Rx, 4 layers, DMRS type 1 on
 * ports 1000-1003) and batches of up to max_slots slots.  params: host fp32
 * [AIRAN_REFINER_PARAMS], or NULL for He-normal weights from `seed`.
 * Synchronises `stream`. */
int32_t airan_ce_refiner_Create(const PuschMimoConfig *config, uint32_t max_slots, const float *params,
                                uint64_t seed, void *stream, AiranCeRefiner **refiner);
void airan_ce_refiner_Destroy(AiranCeRefiner *refiner);
/* Master weights to / from the host (synchronous). */
int32_t airan_ce_refiner_GetParams(AiranCeRefiner *refiner, float *params);
int32_t airan_ce_refiner_SetParams(AiranCeRefiner *refiner, const float *params);

/* Inference on DMRS least-squares estimates in the chain layout
 * (h_ls fp16 [slots][NR][L][D][pilot_pad], from mimo_dmrs_ls): writes the
 * channel grid fp16 [slots][NR][L][14][padded] (zero beyond used) for every
 * slot whose flag keep[slot] is non-zero (keep may be NULL: all slots). */
int32_t airan_ce_refiner_Enqueue(AiranCeRefiner *refiner, const void *h_ls_re, const void *h_ls_im,
                                 uint32_t slots, const uint8_t *keep, void *h_re, void *h_im, void *stream);

/* Enqueue with each slot's noise variance (fp16 [slots][NR], the CE's
 * estimate): needed when the residual is scaled by the noise
 * (airan_ce_refiner_SetResidualScale mode 1); Enqueue is this call with
 * noise NULL. */
int32_t airan_ce_refiner_EnqueueWithNoise(AiranCeRefiner *refiner, const void *h_ls_re, const void *h_ls_im,
                                          const void *noise, uint32_t slots, const uint8_t *keep, void *h_re,
                                          void *h_im, void *stream);

/* Residual scale: the network's output is X + s * conv6(...), s per slot.
 * mode 0 (the default): s = 1.  mode 1: s = the slot's noise standard
 * deviation (training: 10^(-snr/20) of the synthesised slot; inference: the
 * square root of the mean noise variance given to EnqueueWithNoise), so the
 * last layer works in units of the noise and an optimiser step moves the
 * output by an amount proportional to the error it corrects.  Part of the
 * architecture: weights are trained for one mode.  SetSlotScale sets the
 * scales explicitly (for airan_ce_refiner_Gradients: host float [slots]). */
int32_t airan_ce_refiner_SetResidualScale(AiranCeRefiner *refiner, uint32_t mode);
int32_t airan_ce_refiner_SetSlotScale(AiranCeRefiner *refiner, const float *scale, uint32_t slots);

/* Network input.  mode 0 (the default): the LS interpolation of h_ls.
 * mode 1: a residual correction of the LMMSE estimate -- Enqueue reads the
 * estimate from h_re / h_im (the channel_est_lmmse output grid, which it then
 * overwrites for the kept slots; h_ls is not read), and SynthBatch computes
 * it from the batch's DMRS LS with `filter` (an LMMSE filter of the same
 * grant shape; required for SynthBatch in mode 1, not owned).  Weights are
 * trained for one mode. */
int32_t airan_ce_refiner_SetInput(AiranCeRefiner *refiner, uint32_t mode, const struct LmmseFilter *filter);

/* Subcarrier dilation of each of the six layers' taps (count = 6, each
 * 1..16; default 1): layer i reads subcarrier offsets -d_i, 0, +d_i (symbol
 * offsets stay -1, 0, +1).  The dilations are part of the architecture: the
 * weights are trained for one setting.  Dilations 1, 2, 4, 8, 16, 1 give the
 * network a 65-subcarrier view of the band instead of 13. */
int32_t airan_ce_refiner_SetDilation(AiranCeRefiner *refiner, const uint32_t *dilation, uint32_t count);
int32_t airan_ce_refiner_GetDilation(const AiranCeRefiner *refiner, uint32_t *dilation, uint32_t count);

/* Synthesises a training or evaluation batch on the device into the
 * refiner's batch buffers: gains host complex64 [slots][NR][L][taps][14]
 * (tapped-delay-line gains per OFDM symbol), tap_delay_s host [slots][taps]
 * (at most 32 taps), snr_db host [slots] (AWGN per resource element
 * relative to unit channel power), scs_hz the subcarrier spacing.  Also
 * leaves the DMRS least-squares estimate and its noise estimate (the LMMSE
 * estimator's inputs) in the batch buffers. */
int32_t airan_ce_refiner_SynthBatch(AiranCeRefiner *refiner, const float *gains, const float *tap_delay_s,
                                    uint32_t taps, const float *snr_db, uint32_t slots, float scs_hz,
                                    uint64_t seed, void *stream);
/* Device pointers of the batch buffers (valid until Destroy). */
typedef struct AiranCeRefinerBatchV1 {
    uint16_t abi_version;
    uint16_t struct_size;
    const void *x;              /* bf16 [slots][14][used][32] network input */
    const void *target;         /* fp32 [slots][14][used][32] true channel */
    const void *out;            /* fp32 [slots][14][used][32] output of the last Forward,
                                   TrainStep or Gradients (Enqueue in input mode 1 writes
                                   its grid directly and leaves this unchanged) */
    const void *h_ls_re;        /* fp16 [slots][NR][L][D][pilot_pad] */
    const void *h_ls_im;
    const void *noise_ls;       /* fp16 [slots][NR] */
} AiranCeRefinerBatchV1;
int32_t airan_ce_refiner_Batch(const AiranCeRefiner *refiner, AiranCeRefinerBatchV1 *batch);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the synthesised batch.
 * The loss weights each slot's squared error by the slot's linear SNR
 * (weights of mean 1 over the batch), so every SNR counts alike; loss_out
 * (device fp32, may be NULL) receives it before the update. */
int32_t airan_ce_refiner_TrainStep(AiranCeRefiner *refiner, uint32_t slots, float lr, float *loss_out,
                                   void *stream);
/* Forward pass of the current weights on the synthesised batch (result in
 * the batch `out` buffer). */
int32_t airan_ce_refiner_Forward(AiranCeRefiner *refiner, uint32_t slots, void *stream);

/* Loss and gradients of every parameter for an arbitrary input:
 * x bf16 [slots][14][width][32], target fp32 [slots][14][width][32]
 * (width <= used), grad device fp32 [AIRAN_REFINER_PARAMS], loss device
 * fp32 [1].  The loss is the mean over all slots * 14 * width * 32 values. */
int32_t airan_ce_refiner_Gradients(AiranCeRefiner *refiner, const void *x, const void *target, uint32_t slots,
                                   uint32_t width, float *grad, float *loss, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* AIRAN_CE_REFINER_H */
