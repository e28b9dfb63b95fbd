// qam.h -- QAM max-log soft demapping in the reference's integer units,
// shared by the per-operator demapper and the fused detector of the chain.
#pragma once

#include "common.h"

namespace airan {

// fp16 bit patterns of D*32 and D^2*32, D = constellation normalisation
// (qam256_demod.h:50-56 for 256QAM; the same definition for Qm = 2, 4, 6).
__host__ __device__ inline void QamConsts(uint32_t qm, uint16_t *dq, uint16_t *d2q)
{
    switch (qm) {
    case 2: *dq = 0x4da8; *d2q = 0x4c00; break;   // 22.627, 16.0
    case 4: *dq = 0x490f; *d2q = 0x4266; break;   // 10.117, 3.199
    case 6: *dq = 0x44f0; *d2q = 0x3a18; break;   // 4.9375, 0.7617
    default: *dq = 0x40e9; *d2q = 0x3206; break;  // 2.455, 0.1882
    }
}

// Quantisation of gen_data.py:20-27: every fp16 operation rounds once.
__device__ inline void QamUT(uint32_t qm, __half y, __half n, int *u, int *t)
{
    uint16_t dq, d2q;
    QamConsts(qm, &dq, &d2q);
    float nf = __half2float(n);
    if (!(nf >= 6.103515625e-05f) || !isfinite(nf)) nf = 6.103515625e-05f;
    const float dqf = __half2float(__ushort_as_half(dq));
    const float d2qf = __half2float(__ushort_as_half(d2q));
    const __half scale = __float2half_rn(dqf / nf);
    const __half thr = __float2half_rn(d2qf / nf);
    float prod = __half2float(__float2half_rn(__half2float(y) * __half2float(scale)));
    if (!isfinite(prod)) prod = prod > 0.0f ? 65504.0f : -65504.0f;
    int uu = static_cast<int>(rintf(prod));
    float tf = __half2float(thr);
    if (!isfinite(tf)) tf = 65504.0f;
    *t = static_cast<int>(rintf(tf));
    *u = uu > 1000 ? 1000 : (uu < -1000 ? -1000 : uu);
}

__device__ inline int Clip2560(int v) { return v > 2560 ? 2560 : (v < -2560 ? -2560 : v); }

// 256QAM closed form (gen_data.py:28-40), returned already negated (positive
// => bit 0).
__device__ inline void Axis256(int u, int t, int *out)
{
    const int a = u < 0 ? -u : u;
    int cl[8];
    int sa = 0;
#pragma unroll
    for (int i = 1; i <= 7; ++i) {
        const int kt = 2 * i * t;
        cl[i] = a < kt ? a : kt;
        sa += cl[i];
    }
    // cl[i] = min(a, 2i t) for k = 2i, and clamp(u, -2i t, 2i t) =
    // sign(u) min(|u|, 2i t) (t >= 0), so the sum of the clamps is +-sa.
    const int ss = u < 0 ? -sa : sa;
    const int v0 = 32 * u - 4 * ss;
    const int v1 = 80 * t - 16 * a - 4 * (cl[1] + cl[2] + cl[3]) + 4 * (cl[5] + cl[6] + cl[7]);
    const int v2 = -24 * t - 8 * a + 4 * cl[1] - 4 * cl[3] + 16 * cl[4] - 4 * cl[5] + 4 * cl[7];
    const int v3 = -8 * t - 4 * a + 8 * cl[2] - 8 * cl[4] + 8 * cl[6];
    out[0] = -Clip2560(v0);
    out[1] = -Clip2560(v1);
    out[2] = -Clip2560(v2);
    out[3] = -Clip2560(v3);
}

// PAM level under the reference mapper polarity (bit i of code = c_i, c_0 the
// sign bit): level = d0 e1, e_m = 1, e_i = 2^(m-i) - d_i e_{i+1}, d = 2c - 1.
__device__ inline int PamLevel(int m, int code)
{
    int e = 1;
    for (int i = m - 1; i >= 1; --i) e = (1 << (m - i)) - (2 * ((code >> i) & 1) - 1) * e;
    return (2 * (code & 1) - 1) * e;
}

// Max-log by exhaustion over the 2^m levels: LLR_i = min_{c_i=1} cost -
// min_{c_i=0} cost with cost = t s^2 - 2 u s; equal to the 256QAM closed form
// for every integer (u, t).
__device__ inline void AxisExhaustive(int m, int u, int t, int *out)
{
    int b0[4], b1[4];
    for (int i = 0; i < m; ++i) b0[i] = b1[i] = 0x7fffffff;
    for (int code = 0; code < (1 << m); ++code) {
        const int s = PamLevel(m, code);
        const int cost = t * s * s - 2 * u * s;
        for (int i = 0; i < m; ++i) {
            if ((code >> i) & 1)
                b1[i] = cost < b1[i] ? cost : b1[i];
            else
                b0[i] = cost < b0[i] ? cost : b0[i];
        }
    }
    for (int i = 0; i < m; ++i) out[i] = Clip2560(b1[i] - b0[i]);
}

// All 2m axis LLRs of one symbol: streams 0..m-1 from the I axis (NR bits
// b0, b2, ...), m..2m-1 from the Q axis (b1, b3, ...).
__device__ inline void QamSymbolLlrs(uint32_t qm, __half xr, __half xi, __half n, int *out)
{
    int ui, ti, uq, tq;
    QamUT(qm, xr, n, &ui, &ti);
    QamUT(qm, xi, n, &uq, &tq);
    const int m = static_cast<int>(qm / 2);
    if (qm == 8) {
        Axis256(ui, ti, out);
        Axis256(uq, tq, out + 4);
    } else {
        AxisExhaustive(m, ui, ti, out);
        AxisExhaustive(m, uq, tq, out + m);
    }
}

}  // namespace airan
