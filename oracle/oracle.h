/*
 * oracle.h -- CPU restatement of the reference PUSCH MIMO receive chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; nothing in the product library links it or calls it.
 *
 * Every function names the reference file:line it restates.  Where the
 * reference computes in fp16 and states a bit-exact contract (DMRS, DMRS-LS
 * h_ls, QAM demap inputs), fp16 rounding is emulated operation by operation.
 *
 * Parity status per function (how each one is pinned is listed in the test
 * that pins it, tests/test_oracle_pins.py):
 *   or_gold, or_dmrs_gen, or_re_demap, or_dmrs_ls, or_qam_demod,
 *   or_layer_demap, or_descramble, or_rate_dematch, or_crc, or_tbs,
 *   or_runtime_config_*        pinned (closed forms, brute force, reference
 *                              known answers)
 *   or_ofdm_demod              pinned (textbook DFT, Parseval, single tone)
 *   or_ce_lmmse                pinned (noiseless/flat closed forms, full-rank
 *                              identity against a brute-force solve)
 *   or_detect_lmmse            pinned (closed form vs brute-force inverse)
 *   or_ls_siso, or_zf_siso     pinned (pilots equal the DMRS-LS h_ls of port
 *                              1000; interpolation exact on channels linear
 *                              in frequency and time; ZF x = y / h closed
 *                              form, the fp32 arithmetic against numpy)
 *   or_refiner_grad            pinned (exact mode against central finite
 *                              differences of its own loss)
 *   or_ldpc_decode             pinned at outcome level only (encode -> decode
 *                              round trip, syndrome); the reference's fp16/int8
 *                              message trajectory has no host oracle anywhere
 *                              ("parity unpinned" below the outcome level:
 *                              the int8 message rounding of MulCast and the
 *                              rounding count of MulAddDst are choices).  The
 *                              REF hard decision follows the reference's
 *                              int16 CAST_RINT round trip.
 */
#ifndef AIRAN_ORACLE_H
#define AIRAN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/airan_pusch_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- geometry */
/* Derived per-grant geometry.  The reference fixes these numbers per profile
 * (e.g. re_demap_batch.cpp:62-67, mimo_dmrs_ls.h:27-32, qam256_demod_batch.h:
 * 18-24); here they follow the same alignment rules for any geometry:
 * data_stride = align64(data_re) (mimo_profile_compiler.py:562), QAM row
 * stride = align64(used) (pusch_grants/full_band.json qam_row_alignment). */
typedef struct OrGeom {
    int nr, nl, qm, nsym, fft, used, pad;
    int ndmrs, dmrs_sym[4];
    int ndata, data_sym[14];
    int n_dmrs_re;   /* pilots per comb = used/2 */
    int ref_pad;     /* align128(n_dmrs_re) */
    int pilot_pad;   /* align64(n_dmrs_re) */
    int n_occ;       /* used/4 */
    int data_re, data_stride, qam_row, qam_plane;
    int cw_symbols, cw_stride;
    int P, Q;        /* stage-4 DFT factorisation, fft = P*Q, Q = 64 */
    int cp_long, cp_short, samples;
    int sym_start[14];
} OrGeom;

int or_geom(const PuschMimoConfig *cfg, OrGeom *g);

/* ------------------------------------------------------------------- fp16 */
uint16_t or_f2h(float f);
float or_h2f(uint16_t h);

/* ------------------------------------------------------- runtime config */
int or_validate_runtime_config(const PuschMimoRuntimeConfig *c, char *why, int why_len);
int or_derive_config(const PuschMimoRuntimeConfig *rt, uint16_t rx_capacity,
                     uint16_t slot_number, PuschMimoConfig *out, char *why, int why_len);

/* ------------------------------------------------------------------- Gold */
/* c(n) for n in [start, start+count), TS 38.211 5.2.1 */
void or_gold(uint32_t cinit, uint64_t start, uint64_t count, uint8_t *c);
uint32_t or_dmrs_cinit(const PuschMimoConfig *cfg, int symbol);

/* --------------------------------------------------------------- RE demap */
void or_re_demap_gather_index(const PuschMimoConfig *cfg, uint32_t *index_bytes);
void or_re_demap(const PuschMimoConfig *cfg, const uint16_t *fft_re, const uint16_t *fft_im,
                 uint16_t *rx_re, uint16_t *rx_im);

/* ------------------------------------------------------------------- DMRS */
void or_dmrs_gen(const PuschMimoConfig *cfg, uint16_t *re, uint16_t *im);

/* ---------------------------------------------------------------- DMRS LS */
void or_dmrs_ls(const PuschMimoConfig *cfg, const uint16_t *rx_re, const uint16_t *rx_im,
                const uint16_t *ref_re, const uint16_t *ref_im,
                uint16_t *h_re, uint16_t *h_im, uint16_t *pilot_sc, uint16_t *pilot_count,
                double *noise);

/* --------------------------------------------------------------- CE LMMSE */
typedef struct OrCeParams {
    double tau_s;       /* RMS delay spread of the exponential prior */
    double fd_hz;       /* maximum Doppler of the Jakes time prior */
    double scs_hz;      /* subcarrier spacing */
} OrCeParams;
/* h_ls natural [NR,L,D,pilot_pad] (double complex, interleaved) -> H
 * [NR,L,14,used] complex double.  Observation models in layer order: a
 * layer alone on its comb is one model (step 2); an FD-OCC pair is two
 * models g+ = h_ls0 + h_ls1 (offset comb) and g- = h_ls0 - h_ls1 (offset
 * comb + 2), step 4.  noise_rel[q] is model q's relative noise
 * sigma^2 / P_q (or_ce_noise_rel computes it from h_ls and sigma^2). */
int or_ce_lmmse(const PuschMimoConfig *cfg, const OrCeParams *prm, const double *noise_rel,
                const double *h_ls, double *h_out);
int or_ce_noise_rel(const PuschMimoConfig *cfg, const double *h_ls, double sigma2, double *noise_rel);

/* ----------------------------------------------------------- SISO LS / ZF */
/* channel_est_ls_siso (LS + linear frequency/time interpolation, fp16 op
 * order) and equalize_zf_siso (single-tap ZF): or_core.c. */
void or_ls_siso(const PuschMimoConfig *cfg, const uint16_t *rx_re, const uint16_t *rx_im, const uint16_t *ref_re,
                const uint16_t *ref_im, uint16_t *h_re, uint16_t *h_im);
void or_zf_siso(size_t n, const uint16_t *y_re, const uint16_t *y_im, const uint16_t *h_re, const uint16_t *h_im,
                const uint16_t *n0, uint16_t *x_re, uint16_t *x_im, uint16_t *no_eff);

/* ------------------------------------------------------------- detection */
/* H [NR,L,14,used], y [NR,14,used] -> xhat [L,14,used] complex, no_eff
 * [L,14,used]: unbiased per-RE LMMSE (mimo_detect_bri_ref.py:72-86). */
void or_detect_lmmse(const PuschMimoConfig *cfg, const double *h, const double *y,
                     const double *noise_re, double *xhat, double *no_eff);

/* ------------------------------------------------------------ QAM demap */
void or_qam_axis_llr(int qm, float y, float noise, int16_t *llr_out /* qm/2 */);
void or_qam_axis_llr_bruteforce(int qm, float y, float noise, int32_t *llr_out);
void or_qam_demod(const PuschMimoConfig *cfg, const uint16_t *x_re, const uint16_t *x_im,
                  const uint16_t *no_eff, int16_t *layer_llr);

/* ------------------------------------------------------------ layer demap */
void or_layer_demap(const PuschMimoConfig *cfg, const int16_t *layer_llr, int16_t *cw_llr);

/* ------------------------------------------------------------- descramble */
void or_descramble(const PuschMimoConfig *cfg, uint32_t num_slots, uint64_t gold_start,
                   const int16_t *llr_qam, int16_t *llr_nr);

/* --------------------------------------------------------- TB / CB math */
typedef struct OrTbPlan {
    uint32_t tbs, C, K, Kprime, Z, filler, G, E_low, E_high, n_low, N, Ncb, k0, bg;
    uint32_t crc_tb, crc_cb;
} OrTbPlan;
uint32_t or_tbs(uint32_t n_prb, uint32_t n_re_per_prb, uint32_t qm, uint32_t layers,
                double rate);
int or_tb_plan(uint32_t tbs, uint32_t G, uint32_t qm, uint32_t layers, uint32_t rv,
               OrTbPlan *p);

/* ------------------------------------------------------------ rate dematch */
void or_rate_dematch(const OrTbPlan *p, uint32_t layers, uint32_t qm, uint32_t cw_symbols,
                     uint32_t cw_stride, uint32_t num_slots, const int16_t *llr_nr,
                     int16_t *ldpc_llr /* [C, 68*Z] */);

/* -------------------------------------------------------------------- LDPC */
#define OR_LDPC_MODE_REF 0   /* fp16 posterior in Q8.8 units, int8 messages */
#define OR_LDPC_MODE_FP16 1  /* fp16 posterior in LLR units, fp16 messages */
int or_ldpc_active_rows(uint32_t Z, uint32_t Kprime, uint32_t E_max, uint32_t k0, uint32_t Ncb);
void or_ldpc_decode(uint32_t Z, uint32_t n_cb, int16_t *lam /* in/out [n_cb,68Z] */,
                    int mode, int iterations, int active_rows, uint32_t K, int early_stop,
                    int8_t *bits /* [n_cb, K] */, uint8_t *iters_out /* [n_cb] or NULL */);
/* As or_ldpc_decode, also writing the final fp16 posterior (before the
 * clamp and the int16 conversion) as float [n_cb, 68Z]. */
void or_ldpc_decode_posterior(uint32_t Z, uint32_t n_cb, int16_t *lam, int mode, int iterations,
                              int active_rows, uint32_t K, int early_stop, int8_t *bits,
                              uint8_t *iters_out, float *posterior);
void or_ldpc_encode(uint32_t Z, const uint8_t *info /* 22Z */, uint8_t *cw /* 68Z */);
int or_ldpc_syndrome_ok(uint32_t Z, const uint8_t *cw);

/* -------------------------------------------------------------------- CRC */
#define OR_CRC24A 0
#define OR_CRC24B 1
#define OR_CRC16 2
uint32_t or_crc(int kind, const uint8_t *bits, uint64_t nbits);

/* ------------------------------------------------------------------ OFDM */
/* iq: complex64 interleaved [NR, samples] -> grid natural-bin complex double
 * [NR, 14, fft] (unitary DFT, no fftshift). */
void or_ofdm_demod(const PuschMimoConfig *cfg, const float *iq, double *grid);

/* ----------------------------------------------------------------- chain */
/* One slot, one TB: iq complex64 [NR, samples] -> bits int8 [C, K] and the
 * TB CRC outcome (or_chain.c).  0 on success. */
int or_chain_slot(const PuschMimoConfig *cfg, const OrTbPlan *p, const OrCeParams *prm, int ldpc_iters,
                  const float *iq, int8_t *bits, uint8_t *tb_ok);
int or_num_threads(void);

/* --------------------------------------------------------- CE refiner */
/* The refiner network of include/airan_ce_refiner.h (or_refiner.c):
 * params fp32 [or_refiner_param_offset(OR_RF_LAYERS)], x and target fp32
 * [S][14][W][32] -> loss, gradients (double, same layout as params) and the
 * output (fp32 [S][14][W][32], may be NULL).  device_rounding: round to bf16
 * where the device stores bf16. */
#define OR_RF_LAYERS 6
size_t or_refiner_param_offset(int layer);
int or_refiner_grad(const float *params, const float *x, const float *target, int S, int W, int device_rounding,
                    double *grad, double *loss, float *out);
/* The same with a subcarrier dilation per layer (dil [6], each >= 1). */
int or_refiner_grad_dil(const float *params, const float *x, const float *target, int S, int W, int device_rounding,
                        const int *dil, double *grad, double *loss, float *out);
/* ... and a residual scale per slot (slot_scale [S] or NULL for 1):
 * out = x + slot_scale[s] z_5. */
int or_refiner_grad_full(const float *params, const float *x, const float *target, int S, int W, int device_rounding,
                         const int *dil, const float *slot_scale, double *grad, double *loss, float *out);
double or_bf16_round(double v);

#ifdef __cplusplus
}
#endif

#endif
