/*
 * or_ldpc.c -- oracle: BG1 LDPC encode / syndrome / layered normalized
 * min-sum decode.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Decoder schedule (fec/ldpc_decode/ldpc_decode_kernel.cpp:533-713):
 * rows in order, for each row and each of its Z checks i:
 *   gathered_k = lam[col_k][(i + sh_k) mod Z]
 *   ext_k      = gathered_k - msg_old_k
 *   min1/min2 over |ext_k|, sign product over sgn(ext_k) (sgn(0) = +1)
 *   msg_k      = alpha * (prod / sgn_k) * (|ext_k| == min1 ? min2 : min1)
 *   lam[col_k][(i + sh_k) mod Z] = ext_k + msg_k
 * then the posterior is clamped to +-5120, converted to int16 with CAST_RINT
 * (:735, the exported posterior), converted back and bit = (value < 0)
 * (:753-767), so a posterior in [-0.5, 0) decodes to 0.  alpha = 0.75
 * (ldpc_decode.h:50).
 *
 * MODE_REF restates the reference arithmetic: fp16 posterior without
 * per-row saturation (:659-661), messages stored as int8 (:666).  Two
 * roundings of the reference are not fixed by its source and are choices
 * here ("parity unpinned"): MulCast<int8_t, half> (:666) is taken as
 * round-half-even with saturation to +-127 (PREV_CLIP_FX, ldpc_decode.h:59),
 * and MulAddDst (:657) as two fp16 roundings (the product, then the sum).
 * MODE_FP16 is the build's decoder: the same fp16 arithmetic in LLR units
 * (input / 256) with fp16 messages.  An integer posterior saturated at int16
 * range drifts on high-degree columns (increases are clipped, decreases are
 * not) until hard decisions flip; fp16 in LLR units never reaches its range.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "oracle.h"
#include "../include/airan_ldpc_bg1.h"

static inline uint32_t edge_shift(int e, uint32_t Z) { return AIRAN_BG1_SHIFT[e] % Z; }

/* ---------------------------------------------------------------- encode */
static void shift_xor(uint8_t *acc, const uint8_t *x, uint32_t sh, uint32_t Z)
{
    for (uint32_t i = 0; i < Z; ++i) acc[i] ^= x[(i + sh) % Z];
}

static int find_edge(int row, int col)
{
    for (int e = AIRAN_BG1_ROW_START[row]; e < AIRAN_BG1_ROW_START[row + 1]; ++e)
        if (AIRAN_BG1_COL[e] == col) return e;
    return -1;
}

/* Systematic BG1 encoding using the double-diagonal core (TS 38.212 5.3.2). */
void or_ldpc_encode(uint32_t Z, const uint8_t *info, uint8_t *cw)
{
    memset(cw, 0, 68u * Z);
    memcpy(cw, info, 22u * Z);
    uint8_t *lam = (uint8_t *)calloc(4u * Z, 1);
    for (int r = 0; r < 4; ++r)
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            if (AIRAN_BG1_COL[e] < 22) shift_xor(lam + r * Z, info + AIRAN_BG1_COL[e] * Z, edge_shift(e, Z), Z);
    uint8_t *p0 = cw + 22 * Z, *p1 = cw + 23 * Z, *p2 = cw + 24 * Z, *p3 = cw + 25 * Z;
    uint32_t a = edge_shift(find_edge(0, 22), Z);
    uint32_t b = edge_shift(find_edge(1, 22), Z);
    /* sum of rows 0..3: sum(lam) + S_b(p0) = 0  =>  p0[(i+b)%Z] = sum[i] */
    for (uint32_t i = 0; i < Z; ++i) {
        uint8_t s = lam[i] ^ lam[Z + i] ^ lam[2 * Z + i] ^ lam[3 * Z + i];
        p0[(i + b) % Z] = s;
    }
    for (uint32_t i = 0; i < Z; ++i) {
        p1[i] = lam[i] ^ p0[(i + a) % Z];          /* row 0 */
        p3[i] = lam[3 * Z + i] ^ p0[(i + a) % Z];  /* row 3 */
        p2[i] = lam[2 * Z + i] ^ p3[i];            /* row 2 */
    }
    for (int r = 4; r < AIRAN_BG1_ROWS; ++r) {
        uint8_t *pc = cw + (22 + r) * Z;
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            if (AIRAN_BG1_COL[e] != 22 + r) shift_xor(pc, cw + AIRAN_BG1_COL[e] * Z, edge_shift(e, Z), Z);
    }
    free(lam);
}

int or_ldpc_syndrome_ok(uint32_t Z, const uint8_t *cw)
{
    uint8_t *acc = (uint8_t *)malloc(Z);
    int ok = 1;
    for (int r = 0; r < AIRAN_BG1_ROWS && ok; ++r) {
        memset(acc, 0, Z);
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            shift_xor(acc, cw + AIRAN_BG1_COL[e] * Z, edge_shift(e, Z), Z);
        for (uint32_t i = 0; i < Z; ++i) ok &= acc[i] == 0;
    }
    free(acc);
    return ok;
}

/* Rows whose extension parity column is never transmitted contribute zero
 * messages in MODE_FP16 (their degree-1 parity variable has ext == 0 at
 * every visit: lam = 0 + m and ext = lam - m = 0 exactly in fp16), so the decoder may stop at the last row whose parity column
 * holds a transmitted position.  E_max is the largest E of the TB. */
int or_ldpc_active_rows(uint32_t Z, uint32_t Kprime, uint32_t E_max, uint32_t k0, uint32_t Ncb)
{
    uint32_t filler = 22 * Z - Kprime;
    uint32_t NF = Ncb - filler;
    if (k0 != 0 || E_max >= NF) return AIRAN_BG1_ROWS;
    uint32_t fstart = Kprime - 2 * Z;
    uint32_t last = E_max - 1;                       /* last non-filler index */
    uint32_t pos = last < fstart ? last : last + filler;
    uint32_t col = (pos + 2 * Z) / Z;                /* variable column */
    int rows = (int)col - 22 + 1;
    if (rows < 4) rows = 4;
    if (rows > AIRAN_BG1_ROWS) rows = AIRAN_BG1_ROWS;
    return rows;
}

/* ---------------------------------------------------------------- decode */
static inline float h16(float v) { return or_h2f(or_f2h(v)); }

/* Hard decision.  MODE_REF: rint(clamp(v, +-5120)) < 0, the reference's
 * int16 round trip (ldpc_decode_kernel.cpp:735-762); a NaN gives 0.
 * MODE_FP16 (no reference counterpart): v < 0. */
static inline int hard_bit(float v, int mode)
{
    if (mode == OR_LDPC_MODE_REF) {
        float c = v > 5120.0f ? 5120.0f : (v < -5120.0f ? -5120.0f : v);
        return nearbyintf(c) < 0.0f;
    }
    return v < 0.0f;
}

/* Both modes keep an fp16 posterior and fp16 arithmetic, one rounding per
 * operation.  MODE_REF: posterior in the input Q8.8 units, messages stored as
 * int8 (round-half-even, saturated to +-127).  MODE_FP16: posterior in LLR
 * units (input / 256, exact scaling), messages stored as fp16. */
static void decode_one(uint32_t Z, int16_t *lam_io, int mode, int iters, int rows, uint32_t K,
                       int early_stop, int8_t *bits, uint8_t *iters_out, float *post_out)
{
    const uint32_t NV = 68 * Z;
    const int E = AIRAN_BG1_ROW_START[rows];
    const float in_scale = mode == OR_LDPC_MODE_REF ? 1.0f : 1.0f / 256.0f;
    float *lam = (float *)malloc(sizeof(float) * NV);
    float *msg = (float *)calloc((size_t)E * Z, sizeof(float));
    for (uint32_t k = 0; k < NV; ++k) lam[k] = h16((float)lam_io[k] * in_scale);
    float ext[AIRAN_BG1_MAX_DEG];
    int it = 0;
    while (it < iters) {
        for (int r = 0; r < rows; ++r) {
            int e0 = AIRAN_BG1_ROW_START[r], deg = AIRAN_BG1_ROW_START[r + 1] - e0;
            for (uint32_t i = 0; i < Z; ++i) {
                float sp = 1.0f, m1 = INFINITY, m2 = INFINITY;
                for (int k = 0; k < deg; ++k) {
                    uint32_t v = AIRAN_BG1_COL[e0 + k] * Z + (i + edge_shift(e0 + k, Z)) % Z;
                    ext[k] = h16(lam[v] - msg[(size_t)(e0 + k) * Z + i]);
                    float a = fabsf(ext[k]);
                    sp *= ext[k] >= 0.0f ? 1.0f : -1.0f;
                    if (a < m1) {
                        m2 = m1;
                        m1 = a;
                    } else if (a < m2) {
                        m2 = a;
                    }
                }
                sp *= 0.75f;
                for (int k = 0; k < deg; ++k) {
                    float s = sp * (ext[k] >= 0.0f ? 1.0f : -1.0f);
                    float mag = fabsf(ext[k]) == m1 ? m2 : m1;
                    float nm = h16(s * mag);
                    uint32_t v = AIRAN_BG1_COL[e0 + k] * Z + (i + edge_shift(e0 + k, Z)) % Z;
                    lam[v] = h16(ext[k] + nm);
                    if (mode == OR_LDPC_MODE_REF) {
                        float q = nearbyintf(nm);
                        if (q > 127.0f) q = 127.0f;
                        if (q < -127.0f) q = -127.0f;
                        nm = q;
                    }
                    msg[(size_t)(e0 + k) * Z + i] = nm;
                }
            }
        }
        ++it;
        if (early_stop) {
            int ok = 1;
            for (int r = 0; r < rows && ok; ++r)
                for (uint32_t i = 0; i < Z && ok; ++i) {
                    int par = 0;
                    for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
                        par ^= hard_bit(lam[AIRAN_BG1_COL[e] * Z + (i + edge_shift(e, Z)) % Z], mode);
                    ok &= par == 0;
                }
            if (ok) break;
        }
    }
    if (iters_out) *iters_out = (uint8_t)it;
    const float clip = mode == OR_LDPC_MODE_REF ? 5120.0f : 20.0f;
    for (uint32_t k = 0; k < NV; ++k) {
        float v = lam[k];
        if (post_out) post_out[k] = v;
        if (k < K) bits[k] = (int8_t)hard_bit(v, mode);
        if (v > clip) v = clip;
        if (v < -clip) v = -clip;
        lam_io[k] = (int16_t)nearbyintf(v / in_scale);
    }
    free(lam);
    free(msg);
}

void or_ldpc_decode_posterior(uint32_t Z, uint32_t n_cb, int16_t *lam, int mode, int iterations,
                              int active_rows, uint32_t K, int early_stop, int8_t *bits, uint8_t *iters_out,
                              float *posterior)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int cb = 0; cb < (int)n_cb; ++cb)
        decode_one(Z, lam + (size_t)cb * 68 * Z, mode, iterations, active_rows, K, early_stop,
                   bits + (size_t)cb * K, iters_out ? iters_out + cb : NULL,
                   posterior ? posterior + (size_t)cb * 68 * Z : NULL);
}

void or_ldpc_decode(uint32_t Z, uint32_t n_cb, int16_t *lam, int mode, int iterations,
                    int active_rows, uint32_t K, int early_stop, int8_t *bits, uint8_t *iters_out)
{
    or_ldpc_decode_posterior(Z, n_cb, lam, mode, iterations, active_rows, K, early_stop, bits, iters_out, NULL);
}
