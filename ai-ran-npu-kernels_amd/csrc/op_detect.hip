// op_detect.hip -- detector I/O pack (reference layout transform) and the
// unbiased per-RE LMMSE MIMO detector.
//
// Detection per resource element (mimo_detect_bri_ref.py:72-86):
//   A = H^H H + no I (L x L, L <= 4),  m = H^H y,
//   x = A^-1 m,  d_l = max(1 - no Re (A^-1)_ll, 1e-3),
//   xhat_l = x_l / d_l,  no_eff_l = (1 - d_l) / d_l.
// A is Hermitian positive definite; it is factored A = L D L^H in FP32
// registers (one thread per RE), which gives both A^-1 m and diag(A^-1).
#include "common.h"
#include "detect.h"

using airan::DevGeom;
using airan::Geom;

namespace {

using airan::Cf;
using airan::Detect;

constexpr uint32_t kIoMagic = 0x4d495031u;  // "MIP1"

// mimo_detect_io_pack_kernel.cpp:175-194,326-357,404: hrm[re][rx][16] =
// H[rx][l][re] (l < L, else 0), grouped y[g][rx][j] = y[rx][8g + j] (j < 8,
// else 0), no[re] = fp16(sum_rx noise_rx * (1/NR)).  One thread per (re, rx).
__global__ void __launch_bounds__(256) IoPackKernel(const __half *__restrict__ yr, const __half *__restrict__ yi,
                                                    const __half *__restrict__ hr, const __half *__restrict__ hi,
                                                    const __half *__restrict__ noise, __half *__restrict__ hrm_r,
                                                    __half *__restrict__ hrm_i, __half *__restrict__ yv_r,
                                                    __half *__restrict__ yv_i, __half *__restrict__ no,
                                                    uint32_t nr, uint32_t nl, uint32_t n_re)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_re * nr) return;
    const uint32_t re = t / nr, rx = t - re * nr;
    const size_t h_base = static_cast<size_t>(rx) * 16 * n_re + re;
    __half vr[16], vi[16];
#pragma unroll
    for (uint32_t l = 0; l < 16; ++l) {
        vr[l] = l < nl ? hr[h_base + static_cast<size_t>(l) * n_re] : __ushort_as_half(0);
        vi[l] = l < nl ? hi[h_base + static_cast<size_t>(l) * n_re] : __ushort_as_half(0);
    }
    __half *dr = hrm_r + (static_cast<size_t>(re) * nr + rx) * 16;
    __half *di = hrm_i + (static_cast<size_t>(re) * nr + rx) * 16;
#pragma unroll
    for (uint32_t l = 0; l < 16; ++l) {
        dr[l] = vr[l];
        di[l] = vi[l];
    }
    const uint32_t g = re / 8, j = re - g * 8;
    const size_t yo = (static_cast<size_t>(g) * nr + rx) * 16;
    yv_r[yo + j] = yr[static_cast<size_t>(rx) * n_re + re];
    yv_i[yo + j] = yi[static_cast<size_t>(rx) * n_re + re];
    yv_r[yo + 8 + j] = __ushort_as_half(0);
    yv_i[yo + 8 + j] = __ushort_as_half(0);
    if (rx == 0) {
        float s = 0.0f;
        for (uint32_t r = 0; r < nr; ++r) s += __half2float(noise[r]);
        no[re] = __float2half_rn(s * (1.0f / static_cast<float>(nr)));
    }
}

template <int NL>
__global__ void __launch_bounds__(256) DetectPackedKernel(const __half *__restrict__ hrm_r,
                                                          const __half *__restrict__ hrm_i,
                                                          const __half *__restrict__ yv_r,
                                                          const __half *__restrict__ yv_i,
                                                          const __half *__restrict__ no_in,
                                                          __half *__restrict__ xr, __half *__restrict__ xi,
                                                          __half *__restrict__ ne_out, uint32_t nr, uint32_t n_re)
{
    const uint32_t re = blockIdx.x * blockDim.x + threadIdx.x;
    if (re >= n_re) return;
    const float no = __half2float(no_in[re]);
    const size_t hb = static_cast<size_t>(re) * nr * 16;
    const size_t yb = static_cast<size_t>(re / 8) * nr * 16 + (re % 8);
    auto loadh = [&](int r, int l) {
        return Cf{__half2float(hrm_r[hb + r * 16 + l]), __half2float(hrm_i[hb + r * 16 + l])};
    };
    auto loady = [&](int r) { return Cf{__half2float(yv_r[yb + r * 16]), __half2float(yv_i[yb + r * 16])}; };
    Cf xh[4];
    float ne[4];
    Detect<NL>(static_cast<int>(nr), no, loadh, loady, xh, ne);
    for (uint32_t l = 0; l < 16; ++l) {
        const size_t o = static_cast<size_t>(l) * n_re + re;
        if (l < NL) {
            xr[o] = __float2half_rn(xh[l].r);
            xi[o] = __float2half_rn(xh[l].i);
            ne_out[o] = __float2half_rn(ne[l]);
        } else {
            xr[o] = __ushort_as_half(0);
            xi[o] = __ushort_as_half(0);
            ne_out[o] = __ushort_as_half(0);
        }
    }
}

bool PlanOk(const MimoDetectLayerPlan *p, const Geom &g)
{
    return p != nullptr && p->abi_version == AIRAN_PUSCH_MIMO_ABI_VERSION &&
           p->struct_size >= sizeof(MimoDetectLayerPlan) && p->num_rx_antennas == g.nr &&
           p->layer_capacity == airan::kDetectLayers && p->total_layers == g.nl && p->num_allocations == 1 &&
           p->allocations[0].pusch_index == 0 && p->allocations[0].layer_offset == 0 &&
           p->allocations[0].num_layers == g.nl;
}

}  // namespace

extern "C" {

int32_t airan_mimo_detect_io_pack_BuildCurrentProfile(const PuschMimoConfig *configs, size_t num_configs,
                                                      const MimoDetectLayerPlan *plan,
                                                      MimoDetectIoPackKernelMetadata *m)
{
    if (configs == nullptr || m == nullptr || num_configs != 1) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(configs[0], &g);
    if (st != AIRAN_OK) return st;
    if (!PlanOk(plan, g)) return AIRAN_LAYOUT_MISMATCH;
    std::memset(m, 0, sizeof(*m));
    m->magic = kIoMagic;
    m->active_layers = g.nl;
    m->num_rx_antennas = g.nr;
    m->layer_capacity = airan::kDetectLayers;
    m->n_re = g.nsym * g.pad;
    m->abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    return AIRAN_OK;
}

int32_t airan_mimo_detect_io_pack_ValidateOpArgs(const MimoDetectIoPackOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(MimoDetectIoPackOpArgsV1)) || a->rx_grid_re == nullptr ||
        a->rx_grid_im == nullptr || a->h_grid_re == nullptr || a->h_grid_im == nullptr ||
        a->noise_var_rx == nullptr || a->hrm_re == nullptr || a->hrm_im == nullptr || a->yvpad_re == nullptr ||
        a->yvpad_im == nullptr || a->no == nullptr || a->configs == nullptr || a->layer_plan == nullptr ||
        a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    const void *outs[5] = {a->hrm_re, a->hrm_im, a->yvpad_re, a->yvpad_im, a->no};
    const void *ins[5] = {a->rx_grid_re, a->rx_grid_im, a->h_grid_re, a->h_grid_im, a->noise_var_rx};
    for (int i = 0; i < 5; ++i) {
        for (int j = 0; j < 5; ++j)
            if (outs[i] == ins[j]) return AIRAN_INVALID_ARGUMENT;
        for (int j = i + 1; j < 5; ++j)
            if (outs[i] == outs[j]) return AIRAN_INVALID_ARGUMENT;
    }
    MimoDetectIoPackKernelMetadata m{};
    return airan_mimo_detect_io_pack_BuildCurrentProfile(a->configs, a->num_configs, a->layer_plan, &m);
}

int32_t airan_mimo_detect_io_pack_Enqueue(const MimoDetectIoPackOpArgsV1 *a, const void *metadata,
                                          size_t metadata_bytes)
{
    int32_t st = airan_mimo_detect_io_pack_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    if (metadata == nullptr || metadata_bytes < sizeof(MimoDetectIoPackKernelMetadata))
        return AIRAN_RESOURCE_TOO_SMALL;
    Geom g;
    airan::DeriveGeom(a->configs[0], &g);
    const uint32_t n_re = g.nsym * g.pad;
    if (n_re % 8 != 0) return AIRAN_UNSUPPORTED_PROFILE;
    hipLaunchKernelGGL(IoPackKernel, dim3(airan::CeilDiv(static_cast<uint64_t>(n_re) * g.nr, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(a->stream), static_cast<const __half *>(a->rx_grid_re),
                       static_cast<const __half *>(a->rx_grid_im), static_cast<const __half *>(a->h_grid_re),
                       static_cast<const __half *>(a->h_grid_im), static_cast<const __half *>(a->noise_var_rx),
                       static_cast<__half *>(a->hrm_re), static_cast<__half *>(a->hrm_im),
                       static_cast<__half *>(a->yvpad_re), static_cast<__half *>(a->yvpad_im),
                       static_cast<__half *>(a->no), g.nr, g.nl, n_re);
    return airan::LaunchStatus();
}

int32_t airan_mimo_detect_lmmse_ValidateOpArgs(const MimoDetectLmmseOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(MimoDetectLmmseOpArgsV1)) || a->hrm_re == nullptr || a->hrm_im == nullptr ||
        a->yvpad_re == nullptr || a->yvpad_im == nullptr || a->no == nullptr || a->xhat_re == nullptr ||
        a->xhat_im == nullptr || a->no_eff == nullptr || a->config == nullptr || a->layer_plan == nullptr ||
        a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    const void *outs[3] = {a->xhat_re, a->xhat_im, a->no_eff};
    const void *ins[5] = {a->hrm_re, a->hrm_im, a->yvpad_re, a->yvpad_im, a->no};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 5; ++j)
            if (outs[i] == ins[j]) return AIRAN_INVALID_ARGUMENT;
        for (int j = i + 1; j < 3; ++j)
            if (outs[i] == outs[j]) return AIRAN_INVALID_ARGUMENT;
    }
    Geom g;
    int32_t st = airan::DeriveGeom(*a->config, &g);
    if (st != AIRAN_OK) return st;
    return PlanOk(a->layer_plan, g) ? AIRAN_OK : AIRAN_LAYOUT_MISMATCH;
}

int32_t airan_mimo_detect_lmmse_Enqueue(const MimoDetectLmmseOpArgsV1 *a)
{
    int32_t st = airan_mimo_detect_lmmse_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    const uint32_t n_re = g.nsym * g.pad;
    dim3 grid(airan::CeilDiv(n_re, 256));
    hipStream_t s = static_cast<hipStream_t>(a->stream);
    const auto *hr = static_cast<const __half *>(a->hrm_re);
    const auto *hi = static_cast<const __half *>(a->hrm_im);
    const auto *yr = static_cast<const __half *>(a->yvpad_re);
    const auto *yi = static_cast<const __half *>(a->yvpad_im);
    const auto *no = static_cast<const __half *>(a->no);
    auto *xr = static_cast<__half *>(a->xhat_re);
    auto *xi = static_cast<__half *>(a->xhat_im);
    auto *ne = static_cast<__half *>(a->no_eff);
    switch (g.nl) {
    case 1: hipLaunchKernelGGL(DetectPackedKernel<1>, grid, dim3(256), 0, s, hr, hi, yr, yi, no, xr, xi, ne, g.nr, n_re); break;
    case 2: hipLaunchKernelGGL(DetectPackedKernel<2>, grid, dim3(256), 0, s, hr, hi, yr, yi, no, xr, xi, ne, g.nr, n_re); break;
    case 3: hipLaunchKernelGGL(DetectPackedKernel<3>, grid, dim3(256), 0, s, hr, hi, yr, yi, no, xr, xi, ne, g.nr, n_re); break;
    default: hipLaunchKernelGGL(DetectPackedKernel<4>, grid, dim3(256), 0, s, hr, hi, yr, yi, no, xr, xi, ne, g.nr, n_re); break;
    }
    return airan::LaunchStatus();
}

}  // extern "C"
