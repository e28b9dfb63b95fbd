/*
 * or_chain.c -- the CPU receive chain (test infrastructure and the CPU
 * baseline): the oracle stages composed in the order of
 * chain/pusch_mimo_rx_chain/rx_chain.cpp:440-640 for one slot carrying one
 * transport block.  FP64 OFDM, LMMSE channel estimation and detection, the
 * reference's integer demapping, descrambling and rate dematching, the
 * layered min-sum LDPC decoder and the TB CRC check.
 *
 * Differences from the GPU chain, by design of the baseline: the filter's
 * noise comes from the DMRS-LS estimate (the GPU chain re-estimates it from
 * the filter's residual subspace), the filter is full rank, and detection
 * runs in FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* The chain's demapper noise floor: 0.04 (2^(Qm/2) - 1) / norm(Qm). */
static float noise_floor(int qm)
{
    const float norm = qm == 2 ? 2.0f : (qm == 4 ? 10.0f : (qm == 6 ? 42.0f : 170.0f));
    return 0.04f * (float)((1 << (qm / 2)) - 1) / norm;
}

int or_chain_slot(const PuschMimoConfig *cfg, const OrTbPlan *p, const OrCeParams *prm, int ldpc_iters,
                  const float *iq, int8_t *bits, uint8_t *tb_ok)
{
    OrGeom g;
    if (or_geom(cfg, &g) != 0) return -1;
    const int NR = g.nr, L = g.nl, U = g.used, PAD = g.pad, NS = g.nsym;
    int rc = -1;
    double *fg = malloc(sizeof(double) * 2 * (size_t)NR * NS * g.fft);
    uint16_t *rx_re = calloc((size_t)NR * NS * PAD, 2), *rx_im = calloc((size_t)NR * NS * PAD, 2);
    uint16_t *ref_re = calloc((size_t)L * g.ndmrs * g.ref_pad, 2), *ref_im = calloc((size_t)L * g.ndmrs * g.ref_pad, 2);
    const size_t nls = (size_t)NR * L * g.ndmrs * g.pilot_pad;
    uint16_t *h_re = calloc(nls, 2), *h_im = calloc(nls, 2);
    uint16_t *psc = calloc((size_t)L * g.ndmrs * g.pilot_pad, 2), *pcnt = calloc(16, 2);
    double *noise = calloc(NR, sizeof(double));
    double *hls = malloc(sizeof(double) * 2 * nls);
    double *H = malloc(sizeof(double) * 2 * (size_t)NR * L * NS * U);
    double *y = malloc(sizeof(double) * 2 * (size_t)NR * NS * U);
    double *xh = malloc(sizeof(double) * 2 * (size_t)L * NS * U), *ne = malloc(sizeof(double) * (size_t)L * NS * U);
    uint16_t *x_re = calloc((size_t)L * NS * PAD, 2), *x_im = calloc((size_t)L * NS * PAD, 2);
    uint16_t *n16 = calloc((size_t)L * NS * PAD, 2);
    int16_t *llr = malloc(2 * (size_t)L * g.qm * g.qam_plane);
    int16_t *cw = malloc(2 * (size_t)g.qm * g.cw_stride), *nr = malloc(2 * (size_t)g.qm * g.cw_stride);
    int16_t *lam = malloc(2 * (size_t)p->C * 68 * p->Z);
    uint8_t *tb = malloc((size_t)p->C * p->K);
    if (!fg || !rx_re || !rx_im || !ref_re || !ref_im || !h_re || !h_im || !psc || !pcnt || !noise || !hls || !H ||
        !y || !xh || !ne || !x_re || !x_im || !n16 || !llr || !cw || !nr || !lam || !tb)
        goto done;
    /* OFDM + RE demap (fftshift order) to the fp16 grid. */
    or_ofdm_demod(cfg, iq, fg);
    for (int r = 0; r < NR; ++r)
        for (int s = 0; s < NS; ++s)
            for (int u = 0; u < U; ++u) {
                const int bin = u < U / 2 ? g.fft - U / 2 + u : u - U / 2;
                const double *v = fg + 2 * (((size_t)r * NS + s) * g.fft + bin);
                const size_t o = ((size_t)r * NS + s) * PAD + u;
                rx_re[o] = or_f2h((float)v[0]);
                rx_im[o] = or_f2h((float)v[1]);
            }
    /* DMRS and least squares. */
    or_dmrs_gen(cfg, ref_re, ref_im);
    or_dmrs_ls(cfg, rx_re, rx_im, ref_re, ref_im, h_re, h_im, psc, pcnt, noise);
    /* LMMSE CE with the DMRS-LS noise. */
    double s2 = 0.0;
    for (int r = 0; r < NR; ++r) s2 += noise[r];
    s2 /= NR;
    for (size_t i = 0; i < nls; ++i) {
        hls[2 * i] = or_h2f(h_re[i]);
        hls[2 * i + 1] = or_h2f(h_im[i]);
    }
    double noise_rel[4];
    or_ce_noise_rel(cfg, hls, s2, noise_rel);
    if (or_ce_lmmse(cfg, prm, noise_rel, hls, H) != 0) goto done;
    /* Detection (FP64) and the fp16 demapper inputs. */
    for (int r = 0; r < NR; ++r)
        for (int s = 0; s < NS; ++s)
            for (int u = 0; u < U; ++u) {
                const size_t i = ((size_t)r * NS + s) * PAD + u;
                double *o = y + 2 * (((size_t)r * NS + s) * U + u);
                o[0] = or_h2f(rx_re[i]);
                o[1] = or_h2f(rx_im[i]);
            }
    or_detect_lmmse(cfg, H, y, &s2, xh, ne);
    const uint16_t fl = or_f2h(noise_floor(g.qm));
    for (int l = 0; l < L; ++l)
        for (int s = 0; s < NS; ++s)
            for (int u = 0; u < U; ++u) {
                const size_t i = ((size_t)l * NS + s) * U + u, o = ((size_t)l * NS + s) * PAD + u;
                x_re[o] = or_f2h((float)xh[2 * i]);
                x_im[o] = or_f2h((float)xh[2 * i + 1]);
                uint16_t nv = or_f2h((float)ne[i]);
                if (or_h2f(nv) < or_h2f(fl)) nv = fl;
                n16[o] = nv;
            }
    /* Integer stages, LDPC and CRC. */
    or_qam_demod(cfg, x_re, x_im, n16, llr);
    or_layer_demap(cfg, llr, cw);
    or_descramble(cfg, 1, 0, cw, nr);
    or_rate_dematch(p, L, g.qm, g.cw_symbols, g.cw_stride, 1, nr, lam);
    const int rows = or_ldpc_active_rows(p->Z, p->Kprime, p->E_high > p->E_low ? p->E_high : p->E_low, p->k0, p->Ncb);
    or_ldpc_decode(p->Z, p->C, lam, OR_LDPC_MODE_FP16, ldpc_iters, rows, p->K, 1, bits, NULL);
    const uint32_t per_cb = p->Kprime - p->crc_cb;
    for (uint32_t r = 0; r < p->C; ++r)
        for (uint32_t k = 0; k < per_cb; ++k) tb[(size_t)r * per_cb + k] = (uint8_t)bits[(size_t)r * p->K + k];
    *tb_ok = p->crc_tb == 0 ? 1
                            : (uint8_t)(or_crc(p->crc_tb == 24 ? OR_CRC24A : OR_CRC16, tb, (uint64_t)p->tbs + p->crc_tb) == 0);
    rc = 0;
done:
    free(fg); free(rx_re); free(rx_im); free(ref_re); free(ref_im); free(h_re); free(h_im); free(psc);
    free(pcnt); free(noise); free(hls); free(H); free(y); free(xh); free(ne); free(x_re); free(x_im);
    free(n16); free(llr); free(cw); free(nr); free(lam); free(tb);
    return rc;
}
