/*
 * or_refiner.c -- oracle: the channel-estimation refiner network of
 * include/airan_ce_refiner.h (no reference counterpart), forward pass, loss
 * and every parameter gradient, in double precision.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the device kernels).
 *
 * With `device_rounding` set, values are rounded to bf16 where the device
 * stores them in bf16: the weights (bf16 copies of the fp32 master weights),
 * the hidden activations a_i = bf16(relu(z_i)), and every back-propagated
 * pre-activation gradient dz; sums stay in double.  With it clear the same
 * code is the exact network, which the tests pin against central finite
 * differences of the loss.
 *
 * Notation: layer i (0..5) maps a_i (a_0 = X, 32 channels) to
 * z_i = conv3x3(a_i, W_i) + b_i; a_{i+1} = relu(z_i) for i < 5; the taps
 * of layer i sit at symbol offsets -1, 0, +1 and subcarrier offsets
 * -d_i, 0, +d_i (d_i its dilation, 1 by default): off(tap) = (kh - 1,
 * (kw - 1) d_i);
 * out = X + s_slot z_5 (s_slot the slot's residual scale, 1 by default);
 * loss = mean((out - T)^2) over all N values.
 * Backward: dz_5 = 2 s_slot (out - T) / N; gW_i[tap][co][ci] = sum_p dz_i[p][co]
 * a_i[p + off(tap)][ci]; gb_i = sum_p dz_i; dz_{i-1} = convT(dz_i, W_i)
 * masked by a_i > 0.  Tensors are NHWC [S][14][W][C].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const int kRfCin[OR_RF_LAYERS] = {32, 64, 64, 64, 64, 64};
static const int kRfCout[OR_RF_LAYERS] = {64, 64, 64, 64, 64, 32};

size_t or_refiner_param_offset(int layer)
{
    size_t off = 0;
    for (int i = 0; i < layer; ++i) off += (size_t)9 * kRfCout[i] * kRfCin[i] + (size_t)kRfCout[i];
    return off;
}

/* Round to the nearest bf16 (ties to even), returned as a double. */
static double rbf16(double v)
{
    float f = (float)v;
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7f800000u) == 0x7f800000u) return f;
    x += 0x7fffu + ((x >> 16) & 1u);
    x &= 0xffff0000u;
    memcpy(&f, &x, 4);
    return f;
}

double or_bf16_round(double v) { return rbf16(v); }

/* out[p][co] = b[co] + sum_{tap,ci} in[p + off(tap)][ci] w[tap][co][ci] (zero padding). */
static void conv_fwd(const double *in, int cin, const double *w, const double *b, int cout, int S, int W, int d,
                     double *out)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < S; ++s)
        for (int h = 0; h < 14; ++h)
            for (int x = 0; x < W; ++x)
                for (int co = 0; co < cout; ++co) {
                    double acc = b ? b[co] : 0.0;
                    for (int kh = 0; kh < 3; ++kh) {
                        const int hh = h + kh - 1;
                        if (hh < 0 || hh >= 14) continue;
                        for (int kw = 0; kw < 3; ++kw) {
                            const int ww = x + (kw - 1) * d;
                            if (ww < 0 || ww >= W) continue;
                            const double *ip = in + (((size_t)s * 14 + hh) * W + ww) * cin;
                            const double *wp = w + ((size_t)(kh * 3 + kw) * cout + co) * cin;
                            for (int ci = 0; ci < cin; ++ci) acc += ip[ci] * wp[ci];
                        }
                    }
                    out[(((size_t)s * 14 + h) * W + x) * cout + co] = acc;
                }
}

/* da[q][ci] = sum_{tap,co} dz[q - off(tap)][co] w[tap][co][ci]. */
static void conv_bwd_data(const double *dz, int cout, const double *w, int cin, int S, int W, int d, double *da)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < S; ++s)
        for (int h = 0; h < 14; ++h)
            for (int x = 0; x < W; ++x)
                for (int ci = 0; ci < cin; ++ci) {
                    double acc = 0.0;
                    for (int kh = 0; kh < 3; ++kh) {
                        const int hh = h - (kh - 1);
                        if (hh < 0 || hh >= 14) continue;
                        for (int kw = 0; kw < 3; ++kw) {
                            const int ww = x - (kw - 1) * d;
                            if (ww < 0 || ww >= W) continue;
                            const double *dp = dz + (((size_t)s * 14 + hh) * W + ww) * cout;
                            for (int co = 0; co < cout; ++co)
                                acc += dp[co] * w[((size_t)(kh * 3 + kw) * cout + co) * cin + ci];
                        }
                    }
                    da[(((size_t)s * 14 + h) * W + x) * cin + ci] = acc;
                }
}

/* gw[tap][co][ci] = sum_p dz[p][co] a[p + off(tap)][ci]; gb[co] = sum_p dz[p][co]. */
static void conv_bwd_weight(const double *dz, int cout, const double *a, int cin, int S, int W, int d, double *gw,
                            double *gb)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int tap = 0; tap < 9; ++tap)
        for (int co = 0; co < cout; ++co) {
            const int kh = tap / 3, kw = tap % 3;
            for (int ci = 0; ci < cin; ++ci) {
                double acc = 0.0;
                for (int s = 0; s < S; ++s)
                    for (int h = 0; h < 14; ++h) {
                        const int hh = h + kh - 1;
                        if (hh < 0 || hh >= 14) continue;
                        for (int x = 0; x < W; ++x) {
                            const int ww = x + (kw - 1) * d;
                            if (ww < 0 || ww >= W) continue;
                            acc += dz[(((size_t)s * 14 + h) * W + x) * cout + co] *
                                   a[(((size_t)s * 14 + hh) * W + ww) * cin + ci];
                        }
                    }
                gw[((size_t)tap * cout + co) * cin + ci] = acc;
            }
        }
    for (int co = 0; co < cout; ++co) {
        double acc = 0.0;
        for (size_t p = 0; p < (size_t)S * 14 * W; ++p) acc += dz[p * cout + co];
        gb[co] = acc;
    }
}

int or_refiner_grad(const float *params, const float *x, const float *target, int S, int W, int device_rounding,
                    double *grad, double *loss, float *out)
{
    static const int ones[OR_RF_LAYERS] = {1, 1, 1, 1, 1, 1};
    return or_refiner_grad_dil(params, x, target, S, W, device_rounding, ones, grad, loss, out);
}

int or_refiner_grad_dil(const float *params, const float *x, const float *target, int S, int W, int device_rounding,
                        const int *dil, double *grad, double *loss, float *out)
{
    return or_refiner_grad_full(params, x, target, S, W, device_rounding, dil, NULL, grad, loss, out);
}

int or_refiner_grad_full(const float *params, const float *x, const float *target, int S, int W, int device_rounding,
                         const int *dil, const float *slot_scale, double *grad, double *loss, float *out)
{
    const size_t P = (size_t)S * 14 * W;
    const size_t np = or_refiner_param_offset(OR_RF_LAYERS);
    double *wq = malloc(np * sizeof(double));
    double *act[OR_RF_LAYERS + 1];
    double *z5 = malloc(P * 32 * sizeof(double));
    double *dz = malloc(P * 64 * sizeof(double));
    double *da = malloc(P * 64 * sizeof(double));
    if (!wq || !z5 || !dz || !da) return -1;
    for (size_t i = 0; i < np; ++i) wq[i] = device_rounding ? rbf16(params[i]) : params[i];
    /* Biases stay fp32 on the device. */
    for (int i = 0; i < OR_RF_LAYERS; ++i) {
        const size_t b0 = or_refiner_param_offset(i) + (size_t)9 * kRfCout[i] * kRfCin[i];
        for (int co = 0; co < kRfCout[i]; ++co) wq[b0 + co] = params[b0 + co];
    }
    act[0] = malloc(P * 32 * sizeof(double));
    for (size_t i = 0; i < P * 32; ++i) act[0][i] = x[i];
    for (int i = 0; i < OR_RF_LAYERS; ++i) {
        const double *w = wq + or_refiner_param_offset(i);
        const double *b = w + (size_t)9 * kRfCout[i] * kRfCin[i];
        if (i < OR_RF_LAYERS - 1) {
            act[i + 1] = malloc(P * kRfCout[i] * sizeof(double));
            conv_fwd(act[i], kRfCin[i], w, b, kRfCout[i], S, W, dil[i], act[i + 1]);
            for (size_t k = 0; k < P * kRfCout[i]; ++k) {
                double v = act[i + 1][k] > 0.0 ? act[i + 1][k] : 0.0;
                act[i + 1][k] = device_rounding ? rbf16(v) : v;
            }
        } else {
            conv_fwd(act[i], kRfCin[i], w, b, kRfCout[i], S, W, dil[i], z5);
        }
    }
    const double n = (double)P * 32;
    double l = 0.0;
    const size_t per_slot = (size_t)14 * W * 32;
    for (size_t k = 0; k < P * 32; ++k) {
        const double sc = slot_scale ? (double)slot_scale[k / per_slot] : 1.0;
        const double o = act[0][k] + sc * z5[k];
        if (out) out[k] = (float)o;
        const double e = o - target[k];
        l += e * e;
        const double g = 2.0 * sc * e / n;
        dz[k] = device_rounding ? rbf16(g) : g;
    }
    *loss = l / n;
    for (int i = OR_RF_LAYERS - 1; i >= 0; --i) {
        const double *w = wq + or_refiner_param_offset(i);
        double *gw = grad + or_refiner_param_offset(i);
        conv_bwd_weight(dz, kRfCout[i], act[i], kRfCin[i], S, W, dil[i], gw, gw + (size_t)9 * kRfCout[i] * kRfCin[i]);
        if (i == 0) break;
        conv_bwd_data(dz, kRfCout[i], w, kRfCin[i], S, W, dil[i], da);
        for (size_t k = 0; k < P * kRfCin[i]; ++k) {
            const double v = act[i][k] > 0.0 ? da[k] : 0.0;
            dz[k] = device_rounding ? rbf16(v) : v;
        }
    }
    for (int i = 0; i < OR_RF_LAYERS; ++i) free(act[i]);
    free(wq);
    free(z5);
    free(dz);
    free(da);
    return 0;
}
