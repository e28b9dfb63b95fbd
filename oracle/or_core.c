/*
 * or_core.c -- oracle: geometry, fp16, runtime config, Gold, RE demap, DMRS,
 * DMRS-LS, QAM demap, layer demap, descramble, TBS/segmentation, rate
 * dematch, CRC.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int align_up(int v, int a) { return (v + a - 1) / a * a; }

/* ------------------------------------------------------------ geometry */
int or_geom(const PuschMimoConfig *cfg, OrGeom *g)
{
    memset(g, 0, sizeof(*g));
    g->nr = cfg->num_rx_antennas;
    g->nl = cfg->num_layers;
    g->qm = cfg->qm;
    g->nsym = cfg->num_symbols;
    g->fft = cfg->fft_size;
    g->used = cfg->used_subcarriers;
    g->pad = cfg->padded_subcarriers;
    for (int s = 0; s < g->nsym; ++s) {
        if ((cfg->dmrs_symbol_mask >> s) & 1u) {
            if (g->ndmrs >= 4) return -1;
            g->dmrs_sym[g->ndmrs++] = s;
        } else {
            g->data_sym[g->ndata++] = s;
        }
    }
    g->n_dmrs_re = g->used / 2;
    g->ref_pad = align_up(g->n_dmrs_re, 128);
    g->pilot_pad = align_up(g->n_dmrs_re, 64);
    g->n_occ = g->used / 4;
    g->data_re = g->ndata * g->used;
    g->data_stride = align_up(g->data_re, 64);
    g->qam_row = align_up(g->used, 64);
    g->qam_plane = g->ndata * g->qam_row;
    g->cw_symbols = g->nl * g->data_re;
    g->cw_stride = g->nl * g->data_stride;
    g->Q = 64;
    g->P = g->fft / 64;
    /* Normal CP at 30 kHz: 144/2048 and 176/2048 of the FFT size
     * (ofdm/ofdm_demod_batch/ofdm_demod.h:18-22 for N=2048). */
    g->cp_short = g->fft * 9 / 128;
    g->cp_long = g->cp_short + g->fft / 64;
    int pos = 0;
    for (int s = 0; s < g->nsym; ++s) {
        pos += (s == 0) ? g->cp_long : g->cp_short;
        g->sym_start[s] = pos;
        pos += g->fft;
    }
    g->samples = pos;
    return 0;
}

/* ---------------------------------------------------------------- fp16 */
uint16_t or_f2h(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) {
        float a;
        memcpy(&a, &ax, 4);
        double v = (double)a * 16777216.0; /* units of 2^-24 */
        double r = nearbyint(v);           /* default rounding: nearest-even */
        return (uint16_t)(sign | (uint32_t)r);
    }
    uint32_t mant = ax & 0x7fffffu;
    uint32_t exp = (ax >> 23) - 127u + 15u;
    uint32_t h = (exp << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float or_h2f(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t mant = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        float v = (float)mant * (1.0f / 16777216.0f);
        memcpy(&x, &v, 4);
        x |= sign;
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (mant << 13);
    } else {
        x = sign | ((exp - 15u + 127u) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static inline uint16_t hmul(uint16_t a, uint16_t b) { return or_f2h(or_h2f(a) * or_h2f(b)); }
static inline uint16_t hadd(uint16_t a, uint16_t b) { return or_f2h(or_h2f(a) + or_h2f(b)); }
static inline uint16_t hsub(uint16_t a, uint16_t b) { return or_f2h(or_h2f(a) - or_h2f(b)); }

/* ----------------------------------------------------- runtime config */
static int fail(int status, const char *msg, char *why, int len)
{
    if (why && len > 0) {
        strncpy(why, msg, (size_t)len - 1);
        why[len - 1] = 0;
    }
    return status;
}

/* Restates common/pusch_mimo_runtime_config.cpp:23-170. */
int or_validate_runtime_config(const PuschMimoRuntimeConfig *c, char *why, int len)
{
    if (c->abi_version != AIRAN_PUSCH_MIMO_RUNTIME_ABI_VERSION || c->struct_size != 192)
        return fail(AIRAN_RTCFG_ABI_MISMATCH, "runtime config ABI version/size mismatch", why, len);
    if (c->flags & ~AIRAN_RUNTIME_KNOWN_FLAGS)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "runtime config contains unknown flags", why, len);
    if (c->flags & AIRAN_RUNTIME_FLAG_PROFILE_HASH) {
        int zero = 1;
        for (int i = 0; i < 32; ++i) zero &= c->profile_sha256[i] == 0;
        if (zero)
            return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "profile hash flag set with an all-zero digest", why, len);
    }
    if (c->architecture < 1 || c->architecture > 3)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "unsupported MIMO architecture", why, len);
#define VTX(v) ((v) >= 1 && (v) <= 16)
#define VRX(v) ((v) >= 1 && (v) <= 64)
    if (!VTX(c->num_tx_antennas) || !VTX(c->num_tx_rf_chains) || !VRX(c->num_rx_antennas) ||
        !VRX(c->num_rx_rf_chains) || c->num_tx_rf_chains > c->num_tx_antennas ||
        c->num_rx_rf_chains > c->num_rx_antennas)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid antenna/RF-chain topology", why, len);
    if (c->architecture == 1 && (c->num_tx_rf_chains != c->num_tx_antennas ||
                                 c->num_rx_rf_chains != c->num_rx_antennas))
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "full-digital topology requires one RF chain per antenna", why, len);
    if (c->num_layers < 1 || c->num_layers > 4 || c->num_layers > c->num_tx_rf_chains ||
        c->num_layers > c->num_rx_rf_chains || c->num_tx_ports < c->num_layers ||
        c->num_tx_ports > 4 || c->num_tx_ports > c->num_tx_rf_chains)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid layer/port dimensions", why, len);
    if (c->max_rx_antennas < c->num_rx_antennas || c->max_rx_antennas > 64 ||
        c->max_layers < c->num_layers || c->max_layers > 4 ||
        c->rx_bucket < c->num_rx_antennas || c->rx_bucket > c->max_rx_antennas ||
        c->layer_bucket < c->num_layers || c->layer_bucket > 16)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "active dimensions exceed declared bucket/capacity", why, len);
    if ((c->num_rx_antennas != c->rx_bucket || c->num_layers != c->layer_bucket) &&
        !(c->flags & AIRAN_RUNTIME_FLAG_ADAPTER_REQUIRED))
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "capacity padding requires the adapter flag", why, len);
    if (c->qm != 2 && c->qm != 4 && c->qm != 6 && c->qm != 8)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "unsupported modulation order", why, len);
    if (c->num_symbols < 1 || c->num_symbols > 14 || c->fft_size < 128 || c->fft_size > 4096 ||
        c->num_rb < 1 || c->num_rb > 275 || c->rb_start + c->num_rb > 275 ||
        c->used_subcarriers != c->num_rb * 12 || c->used_subcarriers > c->padded_subcarriers ||
        c->padded_subcarriers > c->fft_size)
        return fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "invalid waveform geometry", why, len);
    if (c->num_dmrs_symbols < 1 || c->num_dmrs_symbols >= c->num_symbols ||
        c->dmrs_port_count != c->num_layers || c->dmrs_port_count > 16 ||
        (c->dmrs_type != 1 && c->dmrs_type != 2) || (c->dmrs_length != 1 && c->dmrs_length != 2) ||
        c->num_cdm_groups_without_data < 1 || c->num_cdm_groups_without_data > 3 ||
        c->n_scid > 1 || __builtin_popcount(c->dmrs_symbol_mask) != c->num_dmrs_symbols)
        return fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "invalid DMRS dimensions", why, len);
    for (int i = 0; i < c->dmrs_port_count; ++i)
        for (int j = i + 1; j < c->dmrs_port_count; ++j)
            if (c->dmrs_ports[i] == c->dmrs_ports[j])
                return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "duplicate active DMRS port", why, len);
    for (int i = c->dmrs_port_count; i < 16; ++i)
        if (c->dmrs_ports[i] != 0)
            return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "inactive DMRS ports must be zero", why, len);
    if (c->precoding_mode > 2 || (c->precoding_mode == 0 && c->num_tx_ports != c->num_layers))
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid precoding configuration", why, len);
    if (c->slot_number > 1023 || c->data_scrambling_id > 1023 || c->rnti == 0 ||
        c->start_symbol > 13 || c->num_allocated_symbols == 0 ||
        c->start_symbol + c->num_allocated_symbols > 14 || c->dmrs_additional_position > 3 ||
        c->mapping_type > 1 || c->codeword_index > 1 || c->reserved16 != 0)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid dynamic PUSCH scheduler grant", why, len);
    if (!isfinite(c->channel_gain) || c->channel_gain <= 0.0f || c->channel_gain > 1.0f ||
        !isfinite(c->awgn_std_int16) || c->awgn_std_int16 < 0.0f || c->awgn_std_int16 > 200.0f)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid channel gain/noise", why, len);
    uint32_t data_re = (uint32_t)(c->num_symbols - c->num_dmrs_symbols) * c->used_subcarriers;
    if (c->data_re_per_layer != data_re || c->data_stride_per_layer < data_re ||
        c->qam_row_stride < c->used_subcarriers ||
        c->grid_re_per_port != (uint32_t)c->num_symbols * c->padded_subcarriers)
        return fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "derived tensor geometry mismatch", why, len);
    for (int i = 0; i < 5; ++i)
        if (c->reserved[i] != 0)
            return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "reserved runtime config fields must be zero", why, len);
    if (why && len > 0) why[0] = 0;
    return AIRAN_RTCFG_SUCCESS;
#undef VTX
#undef VRX
}

/* Restates common/pusch_mimo_runtime_config.cpp:216-272. */
int or_derive_config(const PuschMimoRuntimeConfig *rt, uint16_t rx_capacity, uint16_t slot_number,
                     PuschMimoConfig *out, char *why, int len)
{
    int st = or_validate_runtime_config(rt, why, len);
    if (st) return st;
    if (rt->dmrs_port_count > 4 || rx_capacity < rt->num_rx_antennas || rx_capacity > 64 ||
        rt->num_symbols > 255)
        return fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "runtime config cannot be represented by operator ABI v1", why, len);
    PuschMimoConfig r;
    memset(&r, 0, sizeof(r));
    r.abi_version = 1;
    r.struct_size = sizeof(r);
    r.num_layers = rt->num_layers;
    r.num_tx_ports = rt->num_tx_ports;
    r.num_rx_antennas = rx_capacity;
    r.qm = rt->qm;
    r.num_symbols = rt->num_symbols;
    r.fft_size = rt->fft_size;
    r.num_rb = rt->num_rb;
    r.rb_start = rt->rb_start;
    r.slot_number = slot_number == 0 ? rt->slot_number : slot_number;
    r.start_symbol = (uint8_t)rt->start_symbol;
    r.num_allocated_symbols = (uint8_t)rt->num_allocated_symbols;
    r.used_subcarriers = rt->used_subcarriers;
    r.padded_subcarriers = rt->padded_subcarriers;
    r.dmrs_symbol_mask = rt->dmrs_symbol_mask;
    for (int i = 0; i < rt->dmrs_port_count; ++i) r.dmrs_ports[i] = rt->dmrs_ports[i];
    r.dmrs_type = (uint8_t)rt->dmrs_type;
    r.dmrs_length = (uint8_t)rt->dmrs_length;
    r.dmrs_scrambling_id = rt->dmrs_scrambling_id;
    r.data_scrambling_id = rt->data_scrambling_id;
    r.rnti = rt->rnti;
    r.dmrs_additional_position = (uint8_t)rt->dmrs_additional_position;
    r.mapping_type = (uint8_t)rt->mapping_type;
    r.num_cdm_groups_without_data = (uint8_t)rt->num_cdm_groups_without_data;
    r.n_scid = (uint8_t)rt->n_scid;
    r.codeword_index = (uint8_t)rt->codeword_index;
    r.codebook_enabled = rt->precoding_mode == 1;
    r.tpmi = rt->tpmi;
    r.prg_size_rb = rt->prg_size_rb;
    *out = r;
    if (why && len > 0) why[0] = 0;
    return AIRAN_RTCFG_SUCCESS;
}

/* ---------------------------------------------------------------- Gold */
/* TS 38.211 5.2.1; same recursion as fec/descramble_mimo/descramble_mimo.cpp:
 * 53-91 and mimo/mimo_dmrs_gen/mimo_dmrs_gen.cpp:149-176. */
void or_gold(uint32_t cinit, uint64_t start, uint64_t count, uint8_t *c)
{
    uint8_t x1[31], x2[31];
    memset(x1, 0, sizeof(x1));
    x1[0] = 1;
    for (int i = 0; i < 31; ++i) x2[i] = (uint8_t)((cinit >> i) & 1u);
    unsigned pos = 0;
    uint64_t skip = 1600ull + start;
    for (uint64_t n = 0; n < skip + count; ++n) {
        if (n >= skip) c[n - skip] = x1[pos] ^ x2[pos];
        unsigned p1 = (pos + 1) % 31, p2 = (pos + 2) % 31, p3 = (pos + 3) % 31;
        uint8_t n1 = x1[p3] ^ x1[pos];
        uint8_t n2 = x2[p3] ^ x2[p2] ^ x2[p1] ^ x2[pos];
        x1[pos] = n1;
        x2[pos] = n2;
        pos = (pos + 1) % 31;
    }
}

/* mimo/mimo_dmrs_gen/mimo_dmrs_gen.cpp:38-45 */
uint32_t or_dmrs_cinit(const PuschMimoConfig *cfg, int symbol)
{
    uint64_t term = (1ull << 17) * (14ull * cfg->slot_number + (uint64_t)symbol + 1ull) *
                    (2ull * cfg->dmrs_scrambling_id + 1ull);
    uint64_t tail = 2ull * cfg->dmrs_scrambling_id + cfg->n_scid;
    return (uint32_t)((term + tail) & 0x7fffffffu);
}

/* ------------------------------------------------------------ RE demap */
/* mapping/re_demap_batch/re_demap_batch.cpp:105-123, for fft = P*Q, Q = 64. */
void or_re_demap_gather_index(const PuschMimoConfig *cfg, uint32_t *index)
{
    OrGeom g;
    or_geom(cfg, &g);
    for (int u = 0; u < g.pad; ++u) index[u] = 0;
    for (int u = 0; u < g.used; ++u) {
        int bin = (u < g.used / 2) ? g.fft - g.used / 2 + u : u - g.used / 2;
        int element = (bin % g.P) * g.Q + bin / g.P;
        index[u] = (uint32_t)element * 2u;
    }
}

/* mapping/re_demap_batch/re_demap_batch.cpp:125-163 */
void or_re_demap(const PuschMimoConfig *cfg, const uint16_t *fre, const uint16_t *fim,
                 uint16_t *rre, uint16_t *rim)
{
    OrGeom g;
    or_geom(cfg, &g);
    uint32_t *index = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)g.pad);
    or_re_demap_gather_index(cfg, index);
    for (int rx = 0; rx < g.nr; ++rx)
        for (int s = 0; s < g.nsym; ++s) {
            size_t ib = ((size_t)rx * g.nsym + s) * g.fft;
            size_t ob = ((size_t)rx * g.nsym + s) * g.pad;
            for (int u = 0; u < g.pad; ++u) {
                if (u < g.used) {
                    rre[ob + u] = fre[ib + index[u] / 2];
                    rim[ob + u] = fim[ib + index[u] / 2];
                } else {
                    rre[ob + u] = 0;
                    rim[ob + u] = 0;
                }
            }
        }
    free(index);
}

/* ---------------------------------------------------------------- DMRS */
/* Type-1 single-symbol DMRS, ports 1000..1003 (mimo_dmrs_gen.cpp:49-64 port
 * semantics; value r(m) per mimo_dmrs_gen_kernel.cpp:163 and
 * scripts/dmrs_ref.py:32-58).  Output [L,D,ref_pad] fp16 planes. */
void or_dmrs_gen(const PuschMimoConfig *cfg, uint16_t *re, uint16_t *im)
{
    OrGeom g;
    or_geom(cfg, &g);
    const uint16_t pos = or_f2h(0.70710678f), neg = (uint16_t)(pos ^ 0x8000u);
    int off = 6 * cfg->rb_start;
    uint8_t *c = (uint8_t *)malloc((size_t)2 * (g.n_dmrs_re + off) + 2);
    memset(re, 0, sizeof(uint16_t) * (size_t)g.nl * g.ndmrs * g.ref_pad);
    memset(im, 0, sizeof(uint16_t) * (size_t)g.nl * g.ndmrs * g.ref_pad);
    for (int d = 0; d < g.ndmrs; ++d) {
        or_gold(or_dmrs_cinit(cfg, g.dmrs_sym[d]), 0, (uint64_t)2 * (g.n_dmrs_re + off), c);
        for (int l = 0; l < g.nl; ++l) {
            int wf_odd = (cfg->dmrs_ports[l] - 1000) & 1;
            size_t base = ((size_t)l * g.ndmrs + d) * g.ref_pad;
            for (int m = 0; m < g.n_dmrs_re; ++m) {
                int ms = m + off;
                int flip = wf_odd && (m & 1);
                uint16_t vr = c[2 * ms] ? neg : pos;
                uint16_t vi = c[2 * ms + 1] ? neg : pos;
                if (flip) {
                    vr ^= 0x8000u;
                    vi ^= 0x8000u;
                }
                re[base + m] = vr;
                im[base + m] = vi;
            }
        }
    }
    free(c);
}

/* ------------------------------------------------------------- DMRS LS */
/* channel_est/mimo_dmrs_ls/mimo_dmrs_ls_kernel.cpp:146-196,301-337 with the
 * observation model of mimo_dmrs_ls.cpp:77-99; fp16 op order per
 * scripts/gen_data.py:94-107.  The noise reduction is done in double here
 * (the reference reduces in fp16 with an unspecified tree order). */
void or_dmrs_ls(const PuschMimoConfig *cfg, const uint16_t *rx_re, const uint16_t *rx_im,
                const uint16_t *ref_re, const uint16_t *ref_im, uint16_t *h_re, uint16_t *h_im,
                uint16_t *pilot_sc, uint16_t *pilot_count, double *noise)
{
    OrGeom g;
    or_geom(cfg, &g);
    int comb[4], occ[4];
    for (int l = 0; l < g.nl; ++l) comb[l] = (cfg->dmrs_ports[l] - 1000) / 2;
    for (int l = 0; l < g.nl; ++l) {
        int share = 0;
        for (int p = 0; p < g.nl; ++p) share += comb[p] == comb[l];
        occ[l] = share == 2;
    }
    memset(pilot_count, 0, 16 * sizeof(uint16_t));
    memset(pilot_sc, 0, sizeof(uint16_t) * (size_t)g.nl * g.ndmrs * g.pilot_pad);
    for (int l = 0; l < g.nl; ++l)
        for (int d = 0; d < g.ndmrs; ++d) {
            int cnt = occ[l] ? g.n_occ : g.n_dmrs_re;
            pilot_count[l * g.ndmrs + d] = (uint16_t)cnt;
            for (int p = 0; p < cnt; ++p)
                pilot_sc[((size_t)l * g.ndmrs + d) * g.pilot_pad + p] =
                    (uint16_t)(comb[l] + (occ[l] ? 1 + 4 * p : 2 * p));
        }
    uint16_t *hr = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)g.n_dmrs_re);
    uint16_t *hi = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)g.n_dmrs_re);
    const uint16_t half_one = or_f2h(0.5f);
    for (int rx = 0; rx < g.nr; ++rx) {
        double acc = 0.0;
        for (int d = 0; d < g.ndmrs; ++d) {
            const uint16_t *yr = rx_re + ((size_t)rx * g.nsym + g.dmrs_sym[d]) * g.pad;
            const uint16_t *yi = rx_im + ((size_t)rx * g.nsym + g.dmrs_sym[d]) * g.pad;
            for (int l = 0; l < g.nl; ++l) {
                const uint16_t *xr = ref_re + ((size_t)l * g.ndmrs + d) * g.ref_pad;
                const uint16_t *xi = ref_im + ((size_t)l * g.ndmrs + d) * g.ref_pad;
                for (int m = 0; m < g.n_dmrs_re; ++m) {
                    int sc = comb[l] + 2 * m;
                    uint16_t a = hmul(yr[sc], xr[m]), b = hmul(yi[sc], xi[m]);
                    hr[m] = hadd(a, b);
                    uint16_t e = hmul(yi[sc], xr[m]), f = hmul(yr[sc], xi[m]);
                    hi[m] = hsub(e, f);
                }
                size_t ob = (((size_t)rx * g.nl + l) * g.ndmrs + d) * g.pilot_pad;
                int cnt;
                for (int p = 0; p < g.pilot_pad; ++p) {
                    h_re[ob + p] = 0;
                    h_im[ob + p] = 0;
                }
                if (occ[l]) {
                    cnt = g.n_occ;
                    for (int p = 0; p < cnt; ++p) {
                        h_re[ob + p] = hmul(hadd(hr[2 * p], hr[2 * p + 1]), half_one);
                        h_im[ob + p] = hmul(hadd(hi[2 * p], hi[2 * p + 1]), half_one);
                    }
                } else {
                    cnt = g.n_dmrs_re;
                    for (int p = 0; p < cnt; ++p) {
                        h_re[ob + p] = hr[p];
                        h_im[ob + p] = hi[p];
                    }
                }
                if (l == 0) {
                    for (int p = 0; p + 1 < cnt; ++p) {
                        double dr = (double)or_h2f(h_re[ob + p + 1]) - or_h2f(h_re[ob + p]);
                        double di = (double)or_h2f(h_im[ob + p + 1]) - or_h2f(h_im[ob + p]);
                        acc += dr * dr + di * di;
                    }
                }
            }
        }
        int cnt0 = occ[0] ? g.n_occ : g.n_dmrs_re;
        double v = acc / ((double)g.ndmrs * (cnt0 - 1));
        noise[rx] = occ[0] ? v : 0.5 * v;
    }
    free(hr);
    free(hi);
}

/* ------------------------------------------------------------ QAM demap */
/* Constellation normalisation for Qm = 2,4,6,8 (TS 38.211 5.1). */
static double qam_norm(int qm)
{
    switch (qm) {
    case 2: return 1.0 / sqrt(2.0);
    case 4: return 1.0 / sqrt(10.0);
    case 6: return 1.0 / sqrt(42.0);
    default: return 1.0 / sqrt(170.0);
    }
}

/* fp16 bit patterns of D*32 and D^2*32 per Qm (qam256_demod.h:50-56 gives
 * the 256QAM pair; the others follow the same definition). */
static void qam_consts(int qm, uint16_t *dq, uint16_t *d2q)
{
    double d = qam_norm(qm);
    *dq = or_f2h((float)(d * 32.0));
    *d2q = or_f2h((float)(d * d * 32.0));
}

/* Quantised axis value u and threshold t, restating
 * mapping/qam_demod_256_mimo_batch/scripts/gen_data.py:20-27: every fp16
 * operation is one IEEE fp16 rounding.  A noise value that is not a positive
 * normal fp16 number is raised to the smallest normal (2^-14); a non-finite
 * product is saturated to +-65504 before the +-1000 clip. */
static void qam_ut(int qm, float y, float noise, int *u_out, int *t_out)
{
    uint16_t dq, d2q;
    qam_consts(qm, &dq, &d2q);
    uint16_t yh = or_f2h(y);
    uint16_t nh = or_f2h(noise);
    float nf = or_h2f(nh);
    if (!(nf >= 6.103515625e-05f) || !isfinite(nf)) {
        nh = 0x0400u;
        nf = or_h2f(nh);
    }
    uint16_t scale = or_f2h(or_h2f(dq) / nf);
    uint16_t thr = or_f2h(or_h2f(d2q) / nf);
    float prod = or_h2f(or_f2h(or_h2f(yh) * or_h2f(scale)));
    if (!isfinite(prod)) prod = prod > 0 ? 65504.0f : -65504.0f;
    if (isnan(prod)) prod = 0.0f;
    int u = (int)nearbyint(prod);
    float tf = or_h2f(thr);
    if (!isfinite(tf)) tf = 65504.0f;
    int t = (int)nearbyint(tf);
    if (u > 1000) u = 1000;
    if (u < -1000) u = -1000;
    *u_out = u;
    *t_out = t;
}

/* 256QAM integer closed form, gen_data.py:28-40 (output already negated and
 * clipped, i.e. positive => bit 0). */
static void axis_llr_256(int u, int t, int16_t *out)
{
    int a = u < 0 ? -u : u;
    int cl[15];
    int ss = 0;
    for (int k = 2; k <= 14; k += 2) {
        cl[k] = a < k * t ? a : k * t;
        int lo = -k * t, hi = k * t;
        ss += u < lo ? lo : (u > hi ? hi : u);
    }
    int v[4];
    v[0] = 32 * u - 4 * ss;
    v[1] = 80 * t - 16 * a - 4 * (cl[2] + cl[4] + cl[6]) + 4 * (cl[10] + cl[12] + cl[14]);
    v[2] = -24 * t - 8 * a + 4 * cl[2] - 4 * cl[6] + 16 * cl[8] - 4 * cl[10] + 4 * cl[14];
    v[3] = -8 * t - 4 * a + 8 * cl[4] - 8 * cl[8] + 8 * cl[12];
    for (int i = 0; i < 4; ++i) {
        int c = v[i] < -2560 ? -2560 : (v[i] > 2560 ? 2560 : v[i]);
        out[i] = (int16_t)(-c);
    }
}


/* PAM level of the m axis bits in `code` (bit i of code = c_i, c_0 = sign bit)
 * under the reference mapper polarity, d = 2c - 1:
 *   level = d0 * e1,  e_m = 1,  e_i = 2^(m-i) - d_i * e_{i+1}
 * (for m = 4: d0(8 - d1(4 - d2(2 - d3))), mapping/qam_mod_256_siso/scripts/
 * qam256_mod_ref.py:41-45). */
static int pam_level(int m, int code)
{
    int e = 1;
    for (int i = m - 1; i >= 1; --i) {
        int d = 2 * ((code >> i) & 1) - 1;
        e = (1 << (m - i)) - d * e;
    }
    int d0 = 2 * (code & 1) - 1;
    return d0 * e;
}

/* Max-log axis LLRs by exhaustion over the 2^m PAM points, in the same
 * integer units: cost(s) = t*s^2 - 2*u*s, LLR_i = min_{c_i=1} - min_{c_i=0}.
 * Positive => bit 0.  Not clipped. */
void or_qam_axis_llr_bruteforce(int qm, float y, float noise, int32_t *out)
{
    int u, t, m = qm / 2;
    qam_ut(qm, y, noise, &u, &t);
    for (int i = 0; i < m; ++i) {
        long best0 = 1L << 60, best1 = 1L << 60;
        for (int code = 0; code < (1 << m); ++code) {
            long s = pam_level(m, code);
            long cost = (long)t * s * s - 2L * u * s;
            if ((code >> i) & 1) {
                if (cost < best1) best1 = cost;
            } else if (cost < best0) {
                best0 = cost;
            }
        }
        out[i] = (int32_t)(best1 - best0);
    }
}

/* One axis: 256QAM uses the reference closed form (gen_data.py:20-40); the
 * other orders use the max-log exhaustion above, which the reference closed
 * form equals for every integer (u, t) (tests/test_oracle_pins.py). */
void or_qam_axis_llr(int qm, float y, float noise, int16_t *out)
{
    int m = qm / 2;
    if (qm == 8) {
        int u, t;
        qam_ut(qm, y, noise, &u, &t);
        axis_llr_256(u, t, out);
        return;
    }
    int32_t v[4];
    or_qam_axis_llr_bruteforce(qm, y, noise, v);
    for (int i = 0; i < m; ++i) {
        int32_t c = v[i] < -2560 ? -2560 : (v[i] > 2560 ? 2560 : v[i]);
        out[i] = (int16_t)c;
    }
}

/* mapping/qam_demod_256_mimo_batch/qam256_demod_batch_kernel.cpp:147-180:
 * x_re/x_im/no_eff are detector planes [Lcap,14,pad] fp16 (layer stride
 * 14*pad); output [L,qm,ndata,qam_row] int16, streams 0..m-1 from the I axis
 * and m..2m-1 from the Q axis; row tail [used,qam_row) is zero. */
void or_qam_demod(const PuschMimoConfig *cfg, const uint16_t *x_re, const uint16_t *x_im,
                  const uint16_t *no_eff, int16_t *llr)
{
    OrGeom g;
    or_geom(cfg, &g);
    int m = g.qm / 2;
    size_t lstride = (size_t)g.nsym * g.pad;
    memset(llr, 0, sizeof(int16_t) * (size_t)g.nl * g.qm * g.qam_plane);
    for (int l = 0; l < g.nl; ++l)
        for (int ds = 0; ds < g.ndata; ++ds) {
            int s = g.data_sym[ds];
            for (int sc = 0; sc < g.used; ++sc) {
                size_t in = (size_t)l * lstride + (size_t)s * g.pad + sc;
                float nv = or_h2f(no_eff[in]);
                int16_t ai[4], aq[4];
                or_qam_axis_llr(g.qm, or_h2f(x_re[in]), nv, ai);
                or_qam_axis_llr(g.qm, or_h2f(x_im[in]), nv, aq);
                for (int i = 0; i < m; ++i) {
                    llr[((size_t)l * g.qm + i) * g.qam_plane + (size_t)ds * g.qam_row + sc] = ai[i];
                    llr[((size_t)l * g.qm + m + i) * g.qam_plane + (size_t)ds * g.qam_row + sc] = aq[i];
                }
            }
        }
}

/* ------------------------------------------------------------ layer demap */
/* mapping/layer_demap_mimo/layer_demap.cpp:126-157: cw[q][re*L+l] =
 * llr[l][q][ds][sc] with re = ds*used + sc; codeword tail zero. */
void or_layer_demap(const PuschMimoConfig *cfg, const int16_t *llr, int16_t *cw)
{
    OrGeom g;
    or_geom(cfg, &g);
    memset(cw, 0, sizeof(int16_t) * (size_t)g.qm * g.cw_stride);
    for (int q = 0; q < g.qm; ++q)
        for (int ds = 0; ds < g.ndata; ++ds)
            for (int sc = 0; sc < g.used; ++sc) {
                size_t re = (size_t)ds * g.used + sc;
                for (int l = 0; l < g.nl; ++l)
                    cw[(size_t)q * g.cw_stride + re * g.nl + l] =
                        llr[((size_t)l * g.qm + q) * g.qam_plane + (size_t)ds * g.qam_row + sc];
            }
}

/* ------------------------------------------------------------- descramble */
/* fec/descramble_mimo/descramble_mimo.cpp:155-222 with the stream map of
 * descramble_mimo.h:29-30 generalised to Qm: NR bit b comes from QAM stream
 * b/2 (b even, I axis) or m + b/2 (b odd, Q axis).  Layout [slot][Qm]
 * [cw_stride]; the Gold index of (slot, symbol, bit) is gold_start +
 * slot*cw_symbols*Qm + symbol*Qm + bit; c_init = rnti*2^15 + q*2^14 + n_ID. */
void or_descramble(const PuschMimoConfig *cfg, uint32_t num_slots, uint64_t gold_start,
                   const int16_t *in, int16_t *out)
{
    OrGeom g;
    or_geom(cfg, &g);
    int m = g.qm / 2;
    uint32_t cinit = ((uint32_t)cfg->rnti << 15) | ((uint32_t)(cfg->codeword_index & 1u) << 14) |
                     cfg->data_scrambling_id;
    uint64_t nbits = (uint64_t)num_slots * g.cw_symbols * g.qm;
    uint8_t *c = (uint8_t *)malloc(nbits);
    or_gold(cinit, gold_start, nbits, c);
    memset(out, 0, sizeof(int16_t) * (size_t)num_slots * g.qm * g.cw_stride);
    for (uint32_t s = 0; s < num_slots; ++s)
        for (int b = 0; b < g.qm; ++b) {
            int stream = (b & 1) ? m + b / 2 : b / 2;
            size_t ib = ((size_t)s * g.qm + stream) * g.cw_stride;
            size_t ob = ((size_t)s * g.qm + b) * g.cw_stride;
            for (int sym = 0; sym < g.cw_symbols; ++sym) {
                int v = in[ib + sym];
                if (c[(uint64_t)s * g.cw_symbols * g.qm + (uint64_t)sym * g.qm + b]) v = -v;
                if (v > 32767) v = 32767;
                if (v < -32768) v = -32768;
                out[ob + sym] = (int16_t)v;
            }
        }
    free(c);
}

/* ------------------------------------------------------------- TBS / plan */
/* TS 38.214 5.1.3.2 transport block size (N_oh = 0, one codeword). */
static const uint32_t TBS_TABLE[] = {
    24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144, 152, 160, 168, 176,
    184, 192, 208, 224, 240, 256, 272, 288, 304, 320, 336, 352, 368, 384, 408, 432, 456, 480,
    504, 528, 552, 576, 608, 640, 672, 704, 736, 768, 808, 848, 888, 928, 984, 1032, 1064,
    1128, 1160, 1192, 1224, 1256, 1288, 1320, 1352, 1416, 1480, 1544, 1608, 1672, 1736, 1800,
    1864, 1928, 2024, 2088, 2152, 2216, 2280, 2408, 2472, 2536, 2600, 2664, 2728, 2792, 2856,
    2976, 3104, 3240, 3368, 3496, 3624, 3752, 3824};

uint32_t or_tbs(uint32_t n_prb, uint32_t n_re_per_prb, uint32_t qm, uint32_t layers, double rate)
{
    uint32_t nre = (n_re_per_prb < 156 ? n_re_per_prb : 156) * n_prb;
    double ninfo = (double)nre * rate * qm * layers;
    if (ninfo <= 3824.0) {
        int n = (int)floor(log2(ninfo)) - 6;
        if (n < 3) n = 3;
        double p = ldexp(1.0, n);
        double np = p * floor(ninfo / p);
        if (np < 24.0) np = 24.0;
        for (size_t i = 0; i < sizeof(TBS_TABLE) / sizeof(TBS_TABLE[0]); ++i)
            if (TBS_TABLE[i] >= np) return TBS_TABLE[i];
        return 3824;
    }
    int n = (int)floor(log2(ninfo - 24.0)) - 5;
    double p = ldexp(1.0, n);
    double np = p * nearbyint((ninfo - 24.0) / p);
    if (np < 3840.0) np = 3840.0;
    uint32_t C;
    if (rate <= 0.25) {
        C = (uint32_t)ceil((np + 24.0) / 3816.0);
    } else if (np > 8424.0) {
        C = (uint32_t)ceil((np + 24.0) / 8424.0);
    } else {
        C = 1;
    }
    return (uint32_t)(8.0 * C * ceil((np + 24.0) / (8.0 * C)) - 24.0);
}

/* TS 38.212 5.1 (CRC length), 5.2.2 (segmentation, BG1), 5.4.2.1 (E, k0).
 * Only base graph 1 with a set-1 lifting size (Z = 3*2^j) is supported. */
int or_tb_plan(uint32_t tbs, uint32_t G, uint32_t qm, uint32_t layers, uint32_t rv, OrTbPlan *p)
{
    memset(p, 0, sizeof(*p));
    p->tbs = tbs;
    p->bg = 1;
    uint32_t Ltb = tbs > 3824 ? 24 : 16;
    uint32_t B = tbs + Ltb;
    const uint32_t Kcb = 8448;
    uint32_t C, Bp;
    if (B <= Kcb) {
        C = 1;
        Bp = B;
        p->crc_cb = 0;
    } else {
        C = (B + (Kcb - 24) - 1) / (Kcb - 24);
        Bp = B + 24 * C;
        p->crc_cb = 24;
    }
    p->crc_tb = Ltb;
    p->C = C;
    p->Kprime = Bp / C;
    uint32_t Z = 0;
    for (uint32_t z = 3; z <= 384; z *= 2)
        if (22 * z >= p->Kprime) {
            Z = z;
            break;
        }
    if (Z == 0) return -1;
    p->Z = Z;
    p->K = 22 * Z;
    p->filler = p->K - p->Kprime;
    p->N = 66 * Z;
    p->Ncb = p->N;
    p->G = G;
    uint32_t q = layers * qm;
    if (G % q) return -2;
    uint32_t Gp = G / q;
    uint32_t rem = Gp % C;
    p->n_low = C - rem;
    p->E_low = q * (Gp / C);
    p->E_high = q * ((Gp + C - 1) / C);
    static const uint32_t kc[4] = {0, 17, 33, 56};
    p->k0 = (uint32_t)((uint64_t)kc[rv & 3] * p->Ncb / (66ull * Z)) * Z;
    return 0;
}

/* ------------------------------------------------------------ rate dematch */
/* Position in the circular buffer (0..Ncb-1) of the q-th non-filler bit. */
static inline uint32_t nf_pos(const OrTbPlan *p, uint32_t q)
{
    uint32_t fstart = p->Kprime - 2 * p->Z;
    return q < fstart ? q : q + p->filler;
}

/* fec/rate_dematch_mimo/rate_dematch_mimo.cpp:153-201 (inverse bit
 * interleave, circular buffer, x8 scale, +-5120 clip, repetition combining)
 * generalised with filler bits (+5120, a known 0) and puncturing (E < Ncb).
 * llr_nr layout [slot][Qm][cw_stride]; output [C][68Z] with the first 2Z
 * entries zero. */
void or_rate_dematch(const OrTbPlan *p, uint32_t layers, uint32_t qm, uint32_t cw_symbols,
                     uint32_t cw_stride, uint32_t num_slots, const int16_t *in, int16_t *out)
{
    (void)layers;
    const uint32_t Z = p->Z, NF = p->Ncb - p->filler;
    uint32_t fstart = p->Kprime - 2 * Z;
    uint32_t q0 = p->k0 <= fstart ? p->k0 : p->k0 - p->filler;
    int32_t *acc = (int32_t *)malloc(sizeof(int32_t) * p->Ncb);
    uint64_t sym_off = 0;
    for (uint32_t r = 0; r < p->C; ++r) {
        uint32_t E = r < p->n_low ? p->E_low : p->E_high;
        uint32_t eq = E / qm;
        memset(acc, 0, sizeof(int32_t) * p->Ncb);
        for (uint32_t b = 0; b < qm; ++b)
            for (uint32_t j = 0; j < eq; ++j) {
                uint64_t logical = sym_off + j;
                uint64_t slot = logical / cw_symbols, sym = logical % cw_symbols;
                (void)num_slots;
                int32_t v = in[(slot * qm + b) * cw_stride + sym] * 8;
                if (v > 5120) v = 5120;
                if (v < -5120) v = -5120;
                uint32_t k = b * eq + j;
                uint32_t pos = nf_pos(p, (uint32_t)((q0 + (uint64_t)k) % NF));
                acc[pos] += v;
            }
        int16_t *row = out + (size_t)r * 68 * Z;
        for (uint32_t k = 0; k < 2 * Z; ++k) row[k] = 0;
        for (uint32_t k = 0; k < p->Ncb; ++k) {
            int32_t v = acc[k];
            if (k >= fstart && k < fstart + p->filler) v = 5120;
            row[2 * Z + k] = (int16_t)(v > 5120 ? 5120 : (v < -5120 ? -5120 : v));
        }
        for (uint32_t k = 2 * Z + p->Ncb; k < 68 * Z; ++k) row[k] = 0;
        sym_off += eq;
    }
    free(acc);
}

/* -------------------------------------------------------------------- CRC */
/* TS 38.212 5.1: MSB-first polynomial division, zero initial state. */
uint32_t or_crc(int kind, const uint8_t *bits, uint64_t nbits)
{
    uint32_t poly, len;
    switch (kind) {
    case OR_CRC24A: poly = 0x864CFBu; len = 24; break;
    case OR_CRC24B: poly = 0x800063u; len = 24; break;
    default: poly = 0x1021u; len = 16; break;
    }
    uint32_t mask = (len == 32) ? 0xffffffffu : ((1u << len) - 1u);
    uint32_t reg = 0;
    for (uint64_t i = 0; i < nbits; ++i) {
        uint32_t top = ((reg >> (len - 1)) & 1u) ^ (bits[i] & 1u);
        reg = (reg << 1) & mask;
        if (top) reg ^= poly;
    }
    return reg;
}

/* ----------------------------------------------------- SISO LS + ZF */
/* channel_est/channel_est_ls_siso/channel_est_ls_kernel.cpp Phase1_Ls (h_ls
 * = y x* at the even subcarriers, fp16 op order of ComplexMulConj),
 * Phase2_FreqInterp (even: h_ls + 0, odd: (h_ls[p] + h_ls[p+1]) * 0.5, the
 * pilot past the last one zero: "zero_extended_padded_grid",
 * operator.yaml:82-93) and Phase3_TimeInterpAndNaturalWrite (H_0 * fp16((b -
 * sym) / (b - a)) + H_1 * fp16((sym - a) / (b - a)), the reference's k / 9
 * weights for DMRS symbols {2, 11}); one DMRS symbol: H_0 on every symbol.
 * rx [14][pad], ref [D][ref_pad] (airan_mimo_dmrs_gen layout) -> H
 * [14][pad] fp16 bits; subcarriers >= used are zero. */
void or_ls_siso(const PuschMimoConfig *cfg, const uint16_t *rx_re, const uint16_t *rx_im, const uint16_t *ref_re,
                const uint16_t *ref_im, uint16_t *h_re, uint16_t *h_im)
{
    OrGeom g;
    or_geom(cfg, &g);
    const int n = g.n_dmrs_re;
    const uint16_t zero = 0, half_one = or_f2h(0.5f);
    uint16_t *er = (uint16_t *)calloc((size_t)2 * g.pad, sizeof(uint16_t));
    uint16_t *ei = (uint16_t *)calloc((size_t)2 * g.pad, sizeof(uint16_t));
    uint16_t *lr = (uint16_t *)calloc((size_t)n + 1, sizeof(uint16_t));
    uint16_t *li = (uint16_t *)calloc((size_t)n + 1, sizeof(uint16_t));
    for (int d = 0; d < g.ndmrs && d < 2; ++d) {
        const uint16_t *yr = rx_re + (size_t)g.dmrs_sym[d] * g.pad, *yi = rx_im + (size_t)g.dmrs_sym[d] * g.pad;
        const uint16_t *xr = ref_re + (size_t)d * g.ref_pad, *xi = ref_im + (size_t)d * g.ref_pad;
        for (int m = 0; m < n; ++m) {
            lr[m] = hadd(hmul(yr[2 * m], xr[m]), hmul(yi[2 * m], xi[m]));
            li[m] = hsub(hmul(yi[2 * m], xr[m]), hmul(yr[2 * m], xi[m]));
        }
        lr[n] = li[n] = 0;
        for (int m = 0; m < n; ++m) {
            er[(size_t)d * g.pad + 2 * m] = hadd(lr[m], zero);
            ei[(size_t)d * g.pad + 2 * m] = hadd(li[m], zero);
            er[(size_t)d * g.pad + 2 * m + 1] = hmul(hadd(lr[m], lr[m + 1]), half_one);
            ei[(size_t)d * g.pad + 2 * m + 1] = hmul(hadd(li[m], li[m + 1]), half_one);
        }
    }
    for (int sym = 0; sym < 14; ++sym) {
        uint16_t al = 0, om = 0;
        if (g.ndmrs == 2) {
            const float span = (float)g.dmrs_sym[1] - (float)g.dmrs_sym[0];
            al = or_f2h(((float)sym - (float)g.dmrs_sym[0]) / span);
            om = or_f2h(((float)g.dmrs_sym[1] - (float)sym) / span);
        }
        for (int sc = 0; sc < g.pad; ++sc) {
            uint16_t vr = 0, vi = 0;
            if (sc < g.used) {
                if (g.ndmrs == 1) {
                    vr = er[sc];
                    vi = ei[sc];
                } else {
                    vr = hadd(hmul(er[sc], om), hmul(er[(size_t)g.pad + sc], al));
                    vi = hadd(hmul(ei[sc], om), hmul(ei[(size_t)g.pad + sc], al));
                }
            }
            h_re[(size_t)sym * g.pad + sc] = vr;
            h_im[(size_t)sym * g.pad + sc] = vi;
        }
    }
    free(er);
    free(ei);
    free(lr);
    free(li);
}

/* equalization/equalize_zf_siso/equalize_kernel.cpp Process (operator.yaml:
 * 72-84): fp32 from fp16 inputs, g = max(hr^2 + hi^2, 1e-4), x = conj(h) y /
 * g, no_eff = n0 / g, each rounded to fp16.  n elements. */
void or_zf_siso(size_t n, const uint16_t *y_re, const uint16_t *y_im, const uint16_t *h_re, const uint16_t *h_im,
                const uint16_t *n0, uint16_t *x_re, uint16_t *x_im, uint16_t *no_eff)
{
    for (size_t i = 0; i < n; ++i) {
        const float hr = or_h2f(h_re[i]), hi = or_h2f(h_im[i]), yr = or_h2f(y_re[i]), yi = or_h2f(y_im[i]);
        const float g = fmaxf(hr * hr + hi * hi, 1.0e-4f);
        const float a = hr * yr + hi * yi, b = hr * yi - hi * yr;
        x_re[i] = or_f2h(a / g);
        x_im[i] = or_f2h(b / g);
        no_eff[i] = or_f2h(or_h2f(n0[i]) / g);
    }
}
