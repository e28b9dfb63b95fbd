// op_siso.hip -- the SISO receiver's channel estimate and equaliser: DMRS
// least squares with linear frequency/time interpolation
// (channel_est/channel_est_ls_siso) and single-tap zero forcing
// (equalization/equalize_zf_siso).  Both are elementwise, HBM-bound stages.
#include "common.h"
#include "detect.h"
#include "launch.h"

using airan::Geom;
using airan::HAdd;
using airan::HMul;
using airan::HSub;

namespace {

struct LsSisoArgs {
    uint32_t used, pad, nsym, ndmrs, n_pilots, ref_pad, ref_per_slot;
    uint32_t dmrs_sym[2];
    uint16_t alpha[14], oma[14];  // fp16 bits of the time weights of H_1 and H_0
};

// channel_est_ls_kernel.cpp Phase1_Ls / Phase2_FreqInterp /
// Phase3_TimeInterpAndNaturalWrite in one pass: one thread per subcarrier
// pair (2p, 2p + 1) of one slot writes that pair on all 14 symbols.  The
// pilot of p + 1 (the odd subcarrier's right neighbour) is recomputed by the
// thread that needs it; pilots past used / 2 are zero (the reference's
// zero-extended padded grid).  Grid (pad / 512, slots).
__global__ void __launch_bounds__(256) LsInterpSisoKernel(const __half *__restrict__ yr,
                                                          const __half *__restrict__ yi,
                                                          const __half *__restrict__ xr,
                                                          const __half *__restrict__ xi, __half *__restrict__ hr,
                                                          __half *__restrict__ hi, LsSisoArgs a)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * p >= a.pad) return;
    const uint32_t slot = blockIdx.y;
    const size_t grid = static_cast<size_t>(slot) * a.nsym * a.pad;
    const size_t ref = static_cast<size_t>(a.ref_per_slot ? slot : 0u) * a.ndmrs * a.ref_pad;
    const __half zero = __ushort_as_half(0), half_one = __float2half_rn(0.5f);
    __half er[2], ei[2], orr[2], oi[2];
#pragma unroll
    for (uint32_t d = 0; d < 2; ++d) {
        if (d >= a.ndmrs) break;
        const __half *y_r = yr + grid + static_cast<size_t>(a.dmrs_sym[d]) * a.pad;
        const __half *y_i = yi + grid + static_cast<size_t>(a.dmrs_sym[d]) * a.pad;
        const __half *x_r = xr + ref + static_cast<size_t>(d) * a.ref_pad;
        const __half *x_i = xi + ref + static_cast<size_t>(d) * a.ref_pad;
        auto ls = [&](uint32_t m, __half *r, __half *i) {
            if (m >= a.n_pilots) {
                *r = zero;
                *i = zero;
                return;
            }
            const __half yre = y_r[2 * m], yim = y_i[2 * m], c = x_r[m], e = x_i[m];
            *r = HAdd(HMul(yre, c), HMul(yim, e));
            *i = HSub(HMul(yim, c), HMul(yre, e));
        };
        __half r0, i0, r1, i1;
        ls(p, &r0, &i0);
        ls(p + 1, &r1, &i1);
        er[d] = HAdd(r0, zero);
        ei[d] = HAdd(i0, zero);
        orr[d] = HMul(HAdd(r0, r1), half_one);
        oi[d] = HMul(HAdd(i0, i1), half_one);
    }
    const bool live = 2 * p < a.used;  // used is even: both subcarriers or neither
    for (uint32_t sym = 0; sym < a.nsym; ++sym) {
        __half2 vr, vi;
        if (!live) {
            vr = __halves2half2(zero, zero);
            vi = vr;
        } else if (a.ndmrs == 1) {
            vr = __halves2half2(er[0], orr[0]);
            vi = __halves2half2(ei[0], oi[0]);
        } else {
            const __half al = __ushort_as_half(a.alpha[sym]), om = __ushort_as_half(a.oma[sym]);
            vr = __halves2half2(HAdd(HMul(er[0], om), HMul(er[1], al)), HAdd(HMul(orr[0], om), HMul(orr[1], al)));
            vi = __halves2half2(HAdd(HMul(ei[0], om), HMul(ei[1], al)), HAdd(HMul(oi[0], om), HMul(oi[1], al)));
        }
        const size_t o = grid + static_cast<size_t>(sym) * a.pad + 2 * p;
        *reinterpret_cast<__half2 *>(hr + o) = vr;
        *reinterpret_cast<__half2 *>(hi + o) = vi;
    }
}

// equalize_kernel.cpp Process, one thread per resource element.
__global__ void __launch_bounds__(256) ZfSisoKernel(const __half *__restrict__ yr, const __half *__restrict__ yi,
                                                    const __half *__restrict__ hr, const __half *__restrict__ hi,
                                                    const __half *__restrict__ n0, __half *__restrict__ xr,
                                                    __half *__restrict__ xi, __half *__restrict__ ne, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x_r, x_i, n_e;
    airan::ZfSolve(__half2float(hr[i]), __half2float(hi[i]), __half2float(yr[i]), __half2float(yi[i]),
                   __half2float(n0[i]), &x_r, &x_i, &n_e);
    xr[i] = __float2half_rn(x_r);
    xi[i] = __float2half_rn(x_i);
    ne[i] = __float2half_rn(n_e);
}

// Host fp32 -> fp16 bits, round to nearest even (|f| <= 2, normal or zero).
uint16_t HostHalfBits(float f)
{
    return __half_as_ushort(__float2half_rn(f));
}

bool SisoGeomOk(const PuschMimoConfig &cfg, Geom *g)
{
    if (airan::DeriveGeom(cfg, g) != AIRAN_OK) return false;
    return g->nr == 1 && g->nl == 1 && cfg.dmrs_ports[0] == 1000 && g->ndmrs >= 1 && g->ndmrs <= 2 &&
           (g->pad & 1u) == 0;
}

}  // namespace

namespace airan {

int32_t LaunchLsInterpSiso(const PuschMimoConfig &cfg, const Geom &g, const __half *rx_re, const __half *rx_im,
                           const __half *ref_re, const __half *ref_im, bool ref_per_slot, __half *h_re,
                           __half *h_im, uint32_t slots, hipStream_t s)
{
    Geom chk;
    if (!SisoGeomOk(cfg, &chk) || slots == 0) return AIRAN_UNSUPPORTED_PROFILE;
    LsSisoArgs a{};
    a.used = g.used;
    a.pad = g.pad;
    a.nsym = g.nsym;
    a.ndmrs = g.ndmrs;
    a.n_pilots = g.n_dmrs_re;
    a.ref_pad = g.ref_pad;
    a.ref_per_slot = ref_per_slot ? 1u : 0u;
    a.dmrs_sym[0] = g.dmrs_sym[0];
    a.dmrs_sym[1] = g.ndmrs > 1 ? g.dmrs_sym[1] : g.dmrs_sym[0];
    if (g.ndmrs == 2) {
        // channel_est_ls_kernel.cpp Phase3: the weights (k / 9) are fp32
        // quotients rounded to fp16; here (sym - a) / (b - a) and (b - sym) / (b - a).
        const float span = static_cast<float>(g.dmrs_sym[1]) - static_cast<float>(g.dmrs_sym[0]);
        for (uint32_t sym = 0; sym < 14; ++sym) {
            a.alpha[sym] = HostHalfBits((static_cast<float>(sym) - static_cast<float>(g.dmrs_sym[0])) / span);
            a.oma[sym] = HostHalfBits((static_cast<float>(g.dmrs_sym[1]) - static_cast<float>(sym)) / span);
        }
    }
    const dim3 grid(CeilDiv(g.pad / 2, 256), slots);
    hipLaunchKernelGGL(LsInterpSisoKernel, grid, dim3(256), 0, s, rx_re, rx_im, ref_re, ref_im, h_re, h_im, a);
    return LaunchStatus();
}

}  // namespace airan

extern "C" {

int32_t airan_channel_est_ls_siso_ValidateOpArgs(const ChannelEstLsSisoOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(ChannelEstLsSisoOpArgsV1)) || a->config == nullptr || a->rx_grid_re == nullptr ||
        a->rx_grid_im == nullptr || a->dmrs_ref_re == nullptr || a->dmrs_ref_im == nullptr || a->h_re == nullptr ||
        a->h_im == nullptr || a->stream == nullptr || a->batch == 0 || a->ref_per_slot > 1)
        return AIRAN_INVALID_ARGUMENT;
    const void *outs[2] = {a->h_re, a->h_im};
    const void *ins[4] = {a->rx_grid_re, a->rx_grid_im, a->dmrs_ref_re, a->dmrs_ref_im};
    if (a->h_re == a->h_im) return AIRAN_INVALID_ARGUMENT;
    for (const void *o : outs)
        for (const void *i : ins)
            if (o == i) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    if (airan::DeriveGeom(*a->config, &g) != AIRAN_OK || !SisoGeomOk(*a->config, &g))
        return AIRAN_UNSUPPORTED_PROFILE;
    return AIRAN_OK;
}

int32_t airan_channel_est_ls_siso_Enqueue(const ChannelEstLsSisoOpArgsV1 *a)
{
    const int32_t st = airan_channel_est_ls_siso_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    return airan::LaunchLsInterpSiso(*a->config, g, static_cast<const __half *>(a->rx_grid_re),
                                     static_cast<const __half *>(a->rx_grid_im),
                                     static_cast<const __half *>(a->dmrs_ref_re),
                                     static_cast<const __half *>(a->dmrs_ref_im), a->ref_per_slot != 0,
                                     static_cast<__half *>(a->h_re), static_cast<__half *>(a->h_im), a->batch,
                                     static_cast<hipStream_t>(a->stream));
}

int32_t airan_equalize_zf_siso_ValidateOpArgs(const EqualizeZfSisoOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(EqualizeZfSisoOpArgsV1)) || a->y_re == nullptr || a->y_im == nullptr ||
        a->h_re == nullptr || a->h_im == nullptr || a->n0 == nullptr || a->x_re == nullptr || a->x_im == nullptr ||
        a->no_eff == nullptr || a->stream == nullptr || a->n == 0)
        return AIRAN_INVALID_ARGUMENT;
    const void *outs[3] = {a->x_re, a->x_im, a->no_eff};
    const void *ins[5] = {a->y_re, a->y_im, a->h_re, a->h_im, a->n0};
    if (a->x_re == a->x_im || a->x_re == a->no_eff || a->x_im == a->no_eff) return AIRAN_INVALID_ARGUMENT;
    for (const void *o : outs)
        for (const void *i : ins)
            if (o == i) return AIRAN_INVALID_ARGUMENT;
    return AIRAN_OK;
}

int32_t airan_equalize_zf_siso_Enqueue(const EqualizeZfSisoOpArgsV1 *a)
{
    const int32_t st = airan_equalize_zf_siso_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    hipLaunchKernelGGL(ZfSisoKernel, dim3(airan::CeilDiv(a->n, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(a->stream), static_cast<const __half *>(a->y_re),
                       static_cast<const __half *>(a->y_im), static_cast<const __half *>(a->h_re),
                       static_cast<const __half *>(a->h_im), static_cast<const __half *>(a->n0),
                       static_cast<__half *>(a->x_re), static_cast<__half *>(a->x_im),
                       static_cast<__half *>(a->no_eff), a->n);
    return airan::LaunchStatus();
}

}  // extern "C"
