// op_dmrs.hip -- type-1 DMRS generation (Gold -> QPSK -> FD-OCC) and DMRS
// least-squares channel observation with per-Rx noise estimation.
#include "common.h"
#include "gold.h"
#include "launch.h"
#include "ragged.h"

using airan::DevGeom;
using airan::Geom;
using airan::HAdd;
using airan::HMul;
using airan::HSub;

namespace {

constexpr uint32_t kDmrsGenMagic = 0x4d444731u;  // "MDG1"
constexpr uint32_t kDmrsLsMagic = 0x4d4c5331u;   // "MLS1"

// mimo/mimo_dmrs_gen/mimo_dmrs_gen.cpp:38-45
uint32_t DmrsCinit(const PuschMimoConfig &c, uint32_t symbol)
{
    const uint64_t term = (1ull << 17) * (14ull * c.slot_number + symbol + 1ull) *
                          (2ull * c.dmrs_scrambling_id + 1ull);
    const uint64_t tail = 2ull * c.dmrs_scrambling_id + c.n_scid;
    return static_cast<uint32_t>((term + tail) & 0x7fffffffu);
}

struct DmrsGenDev {
    uint32_t nl, ndmrs, n_dmrs_re, ref_pad, off;
    uint32_t cinit[4];
    uint32_t wf_odd[4];
    uint32_t sym[4];      // DMRS symbol of each d (c_init from a slot number)
    uint32_t nid, nscid;  // dmrs_scrambling_id, n_SCID
};

// DmrsCinit on the device.
__device__ inline uint32_t DevDmrsCinit(uint32_t slot_number, uint32_t symbol, uint32_t nid, uint32_t nscid)
{
    const uint64_t term = (1ull << 17) * (14ull * slot_number + symbol + 1ull) * (2ull * nid + 1ull);
    return static_cast<uint32_t>((term + 2ull * nid + nscid) & 0x7fffffffu);
}

// r(m) = ((1-2c(2m)) + j(1-2c(2m+1)))/sqrt(2) (mimo_dmrs_gen_kernel.cpp:163),
// times w_f = (-1)^m on ports 1001/1003.  A thread owns 16 consecutive m.
// Grid (m blocks, D, slots * L).  slot_no == nullptr: one reference from
// d.cinit; else one reference per slot ([slots][L][D][ref_pad]), its c_init
// from the slot's number slot_no[slot].
// 16 consecutive pilots m0 .. m0 + 15 of one reference [ref_pad] at `base`:
// pilot m of DMRS c_init `cinit` (Gold offset `off`), zero past n_dmrs_re.
__device__ inline void DmrsGen16(uint16_t *__restrict__ re, uint16_t *__restrict__ im, size_t base, uint32_t m0,
                                 uint32_t ref_pad, uint32_t n_dmrs_re, uint32_t cinit, uint32_t off, bool wf_odd)
{
    const uint16_t pos = 0x39a8, neg = 0xb9a8;  // fp16(+-0.70710678)
    const uint32_t bits = airan::GoldWordAt(cinit, 2ull * (m0 + off));
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
        const uint32_t m = m0 + j;
        if (m >= ref_pad) break;
        uint16_t vr = 0, vi = 0;
        if (m < n_dmrs_re) {
            vr = ((bits >> (2 * j)) & 1u) ? neg : pos;
            vi = ((bits >> (2 * j + 1)) & 1u) ? neg : pos;
            if (wf_odd && (m & 1u)) {
                vr ^= 0x8000u;
                vi ^= 0x8000u;
            }
        }
        re[base + m] = vr;
        im[base + m] = vi;
    }
}

// Grid (m blocks, D, slots): the layers of a (slot, DMRS symbol) share one
// Gold sequence (they differ in w_f only), so a thread derives the word of
// its 16 pilots once and writes every layer's 16 values; vec: the outputs
// are 16-byte aligned (ref_pad is a multiple of 128, so are all the rows),
// which allows 16-byte stores.
__global__ void __launch_bounds__(256) DmrsGenKernel(uint16_t *__restrict__ re, uint16_t *__restrict__ im,
                                                     const uint16_t *__restrict__ slot_no, DmrsGenDev d,
                                                     uint32_t vec)
{
    const uint32_t m0 = (blockIdx.x * blockDim.x + threadIdx.x) * 16;
    if (m0 >= d.ref_pad) return;
    const uint32_t dd = blockIdx.y, slot = blockIdx.z;
    const uint32_t cinit = slot_no == nullptr ? d.cinit[dd] : DevDmrsCinit(slot_no[slot], d.sym[dd], d.nid, d.nscid);
    if (!vec) {
        for (uint32_t l = 0; l < d.nl; ++l) {
            const size_t base = ((static_cast<size_t>(slot) * d.nl + l) * d.ndmrs + dd) * d.ref_pad;
            DmrsGen16(re, im, base, m0, d.ref_pad, d.n_dmrs_re, cinit, d.off, d.wf_odd[l] != 0);
        }
        return;
    }
    // Pilots m0 .. m0 + 15 (m0 is a multiple of 16 and ref_pad of 128: all
    // below ref_pad) as fp16 pairs, w_f = +1; pilots past n_dmrs_re are +0.
    const uint32_t pos = 0x39a8u, neg = 0xb9a8u;  // fp16(+-0.70710678)
    const uint32_t bits = m0 < d.n_dmrs_re ? airan::GoldWordAt(cinit, 2ull * (m0 + d.off)) : 0u;
    uint32_t wr[8], wi[8], odd[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t m = m0 + 2u * k;
        const bool v0 = m < d.n_dmrs_re, v1 = m + 1u < d.n_dmrs_re;
        const uint32_t r0 = v0 ? ((bits >> (4u * k)) & 1u ? neg : pos) : 0u;
        const uint32_t i0 = v0 ? ((bits >> (4u * k + 1u)) & 1u ? neg : pos) : 0u;
        const uint32_t r1 = v1 ? ((bits >> (4u * k + 2u)) & 1u ? neg : pos) : 0u;
        const uint32_t i1 = v1 ? ((bits >> (4u * k + 3u)) & 1u ? neg : pos) : 0u;
        wr[k] = r0 | r1 << 16;
        wi[k] = i0 | i1 << 16;
        odd[k] = v1 ? 0x80000000u : 0u;  // w_f = -1 negates the odd pilot m + 1
    }
    for (uint32_t l = 0; l < d.nl; ++l) {
        const uint32_t f = d.wf_odd[l] != 0 ? ~0u : 0u;
        const size_t base = ((static_cast<size_t>(slot) * d.nl + l) * d.ndmrs + dd) * d.ref_pad + m0;
        uint4 *pr = reinterpret_cast<uint4 *>(re + base), *pi = reinterpret_cast<uint4 *>(im + base);
        pr[0] = uint4{wr[0] ^ (odd[0] & f), wr[1] ^ (odd[1] & f), wr[2] ^ (odd[2] & f), wr[3] ^ (odd[3] & f)};
        pr[1] = uint4{wr[4] ^ (odd[4] & f), wr[5] ^ (odd[5] & f), wr[6] ^ (odd[6] & f), wr[7] ^ (odd[7] & f)};
        pi[0] = uint4{wi[0] ^ (odd[0] & f), wi[1] ^ (odd[1] & f), wi[2] ^ (odd[2] & f), wi[3] ^ (odd[3] & f)};
        pi[1] = uint4{wi[4] ^ (odd[4] & f), wi[5] ^ (odd[5] & f), wi[6] ^ (odd[6] & f), wi[7] ^ (odd[7] & f)};
    }
}

// Ragged batch: item = (grant, layer); grid (m blocks of the largest grant,
// D max, items).
__global__ void __launch_bounds__(256) RaggedDmrsGenKernel(uint16_t *__restrict__ re, uint16_t *__restrict__ im,
                                                           const airan::RaggedGrant *__restrict__ grants,
                                                           const airan::RaggedItem *__restrict__ items)
{
    const airan::RaggedItem it = items[blockIdx.z];
    const airan::RaggedGrant &gr = grants[it.grant];
    const uint32_t m0 = (blockIdx.x * blockDim.x + threadIdx.x) * 16;
    const uint32_t dd = blockIdx.y;
    if (m0 >= gr.g.ref_pad || dd >= gr.g.ndmrs) return;
    const size_t base = gr.ref_off + (static_cast<size_t>(it.a) * gr.g.ndmrs + dd) * gr.g.ref_pad;
    DmrsGen16(re, im, base, m0, gr.g.ref_pad, gr.g.n_dmrs_re, gr.dmrs_cinit[dd], gr.dmrs_off, gr.wf_odd[it.a] != 0);
}

// channel_est/mimo_dmrs_ls/mimo_dmrs_ls_kernel.cpp:146-177,301-331: h = y x*
// in the reference fp16 operation order, FD-OCC2 pair mean when two layers
// share the comb.  Grid (pilot blocks, D, slots * NR * L).  The DMRS
// references are [L][D][ref_pad], one for every slot (ref_per_slot, their
// c_init depends on the slot number, mimo_dmrs_gen.cpp:38-45) or one shared.
// Observation p of layer l (comb `comb`, FD-OCC2 pair mean when occ) from
// the grid row y and the layer's reference x.
__device__ inline void DmrsLsPilot(const __half *__restrict__ y_r, const __half *__restrict__ y_i,
                                   const __half *__restrict__ x_r, const __half *__restrict__ x_i, uint32_t comb,
                                   bool occ, uint32_t n_occ, uint32_t n_dmrs_re, uint32_t p, __half *vr, __half *vi)
{
    auto ls = [&](uint32_t m, __half *or_, __half *oi) {
        const uint32_t sc = comb + 2 * m;
        const __half a = y_r[sc], b = y_i[sc], c = x_r[m], e = x_i[m];
        *or_ = HAdd(HMul(a, c), HMul(b, e));
        *oi = HSub(HMul(b, c), HMul(a, e));
    };
    *vr = __ushort_as_half(0);
    *vi = __ushort_as_half(0);
    if (occ) {
        if (p < n_occ) {
            __half r0, i0, r1, i1;
            ls(2 * p, &r0, &i0);
            ls(2 * p + 1, &r1, &i1);
            const __half half_one = __float2half_rn(0.5f);
            *vr = HMul(HAdd(r0, r1), half_one);
            *vi = HMul(HAdd(i0, i1), half_one);
        }
    } else if (p < n_dmrs_re) {
        ls(p, vr, vi);
    }
}

__global__ void __launch_bounds__(256) DmrsLsKernel(const __half *__restrict__ yr,
                                                    const __half *__restrict__ yi,
                                                    const __half *__restrict__ xr,
                                                    const __half *__restrict__ xi, uint32_t ref_per_slot,
                                                    __half *__restrict__ hr, __half *__restrict__ hi,
                                                    DevGeom g)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.pilot_pad) return;
    const uint32_t d = blockIdx.y;
    // blockIdx.z = (slot * NR + rx) * L + l; slots are stacked grids.
    const uint32_t srx = blockIdx.z / g.nl, l = blockIdx.z - srx * g.nl;
    const size_t out = (static_cast<size_t>(blockIdx.z) * g.ndmrs + d) * g.pilot_pad + p;
    const __half *y_r = yr + (static_cast<size_t>(srx) * g.nsym + g.dmrs_sym[d]) * g.pad;
    const __half *y_i = yi + (static_cast<size_t>(srx) * g.nsym + g.dmrs_sym[d]) * g.pad;
    const uint32_t ref_slot = ref_per_slot ? srx / g.nr : 0u;
    const size_t ref = (static_cast<size_t>(ref_slot) * g.nl + l) * g.ndmrs + d;
    __half vr, vi;
    DmrsLsPilot(y_r, y_i, xr + ref * g.ref_pad, xi + ref * g.ref_pad, g.comb[l], g.occ[l] != 0, g.n_occ,
                g.n_dmrs_re, p, &vr, &vi);
    hr[out] = vr;
    hi[out] = vi;
}

// Ragged batch: item = (grant, rx, layer); grid (pilot blocks of the largest
// grant, D max, items).
__global__ void __launch_bounds__(256) RaggedDmrsLsKernel(const __half *__restrict__ yr,
                                                          const __half *__restrict__ yi,
                                                          const __half *__restrict__ xr,
                                                          const __half *__restrict__ xi, __half *__restrict__ hr,
                                                          __half *__restrict__ hi,
                                                          const airan::RaggedGrant *__restrict__ grants,
                                                          const airan::RaggedItem *__restrict__ items)
{
    const airan::RaggedItem it = items[blockIdx.z];
    const airan::RaggedGrant &gr = grants[it.grant];
    const DevGeom &g = gr.g;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x, d = blockIdx.y;
    if (p >= g.pilot_pad || d >= g.ndmrs) return;
    const uint32_t rx = it.a, l = it.b;
    const size_t out = gr.hls_off + ((static_cast<size_t>(rx) * g.nl + l) * g.ndmrs + d) * g.pilot_pad + p;
    const size_t row = gr.rx_off + (static_cast<size_t>(rx) * g.nsym + g.dmrs_sym[d]) * g.pad;
    const size_t ref = gr.ref_off + (static_cast<size_t>(l) * g.ndmrs + d) * g.ref_pad;
    __half vr, vi;
    DmrsLsPilot(yr + row, yi + row, xr + ref, xi + ref, g.comb[l], g.occ[l] != 0, g.n_occ, g.n_dmrs_re, p, &vr,
                &vi);
    hr[out] = vr;
    hi[out] = vi;
}

// Noise per Rx from layer 0 (mimo_dmrs_ls_kernel.cpp:179-196,334-337):
// sum_d sum_p |h_{p+1} - h_p|^2 / (D (cnt - 1)), halved for the comb model.
// One workgroup per (slot, Rx) = blockIdx.x, fp32 accumulation, fixed
// reduction tree.
// The noise of one Rx: `h` at its layer-0 observations [D][pilot_pad].
__device__ inline void DmrsNoise(const __half *__restrict__ hr, const __half *__restrict__ hi, size_t rx_base,
                                 const DevGeom &g, __half *out)
{
    const uint32_t cnt = g.occ[0] ? g.n_occ : g.n_dmrs_re;
    float acc = 0.0f;
    for (uint32_t d = 0; d < g.ndmrs; ++d) {
        const size_t base = rx_base + static_cast<size_t>(d) * g.pilot_pad;
        for (uint32_t p = threadIdx.x; p + 1 < cnt; p += blockDim.x) {
            const float dr = __half2float(hr[base + p + 1]) - __half2float(hr[base + p]);
            const float di = __half2float(hi[base + p + 1]) - __half2float(hi[base + p]);
            acc += dr * dr + di * di;
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float s = (part[0] + part[1]) + (part[2] + part[3]);
        float v = s / (static_cast<float>(g.ndmrs) * static_cast<float>(cnt - 1));
        if (!g.occ[0]) v *= 0.5f;
        *out = __float2half_rn(v);
    }
}

// Noise per Rx from layer 0 (mimo_dmrs_ls_kernel.cpp:179-196,334-337):
// sum_d sum_p |h_{p+1} - h_p|^2 / (D (cnt - 1)), halved for the comb model.
// One workgroup per (slot, Rx) = blockIdx.x, fp32 accumulation, fixed
// reduction tree.
__global__ void __launch_bounds__(256) DmrsNoiseKernel(const __half *__restrict__ hr,
                                                       const __half *__restrict__ hi,
                                                       __half *__restrict__ noise, DevGeom g)
{
    const uint32_t rx = blockIdx.x;
    DmrsNoise(hr, hi, static_cast<size_t>(rx) * g.nl * g.ndmrs * g.pilot_pad, g, &noise[rx]);
}

// Ragged batch: one workgroup per item (grant, rx).
__global__ void __launch_bounds__(256) RaggedDmrsNoiseKernel(const __half *__restrict__ hr,
                                                             const __half *__restrict__ hi,
                                                             __half *__restrict__ noise,
                                                             const airan::RaggedGrant *__restrict__ grants,
                                                             const airan::RaggedItem *__restrict__ items)
{
    const airan::RaggedItem it = items[blockIdx.x];
    const airan::RaggedGrant &gr = grants[it.grant];
    DmrsNoise(hr, hi, gr.hls_off + static_cast<size_t>(it.a) * gr.g.nl * gr.g.ndmrs * gr.g.pilot_pad, gr.g,
              &noise[gr.noise_off + it.a]);
}

// Pilot description (mimo_dmrs_ls_kernel.cpp:209-240): observation p of layer
// l sits at comb + 1 + 4p (FD-OCC2) or comb + 2p; pilot_count[l*D + d].
__global__ void __launch_bounds__(256) PilotDescKernel(uint16_t *__restrict__ sc,
                                                       uint16_t *__restrict__ count, DevGeom g)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y, l = blockIdx.z;
    const uint32_t cnt = g.occ[l] ? g.n_occ : g.n_dmrs_re;
    if (p < g.pilot_pad)
        sc[(static_cast<size_t>(l) * g.ndmrs + d) * g.pilot_pad + p] =
            p < cnt ? static_cast<uint16_t>(g.comb[l] + (g.occ[l] ? 1u + 4u * p : 2u * p)) : 0;
    if (blockIdx.x == 0 && d == 0 && l == 0 && threadIdx.x < 16) {
        const uint32_t t = threadIdx.x;
        const uint32_t tl = t / g.ndmrs;
        count[t] = tl < g.nl ? static_cast<uint16_t>(g.occ[tl] ? g.n_occ : g.n_dmrs_re) : 0;
    }
}

}  // namespace

struct MimoDmrsGenRuntimeV1 {
    int device;
};

namespace airan {

int32_t LaunchDmrsLs(const Geom &g, const __half *rx_re, const __half *rx_im, const __half *ref_re,
                     const __half *ref_im, __half *h_re, __half *h_im, __half *noise, uint32_t slots,
                     hipStream_t s, bool ref_per_slot)
{
    const DevGeom dg = ToDev(g);
    const dim3 grid(CeilDiv(g.pilot_pad, 256), g.ndmrs, slots * g.nr * g.nl);
    hipLaunchKernelGGL(DmrsLsKernel, grid, dim3(256), 0, s, rx_re, rx_im, ref_re, ref_im, ref_per_slot ? 1u : 0u,
                       h_re, h_im, dg);
    hipLaunchKernelGGL(DmrsNoiseKernel, dim3(slots * g.nr), dim3(256), 0, s, h_re, h_im, noise, dg);
    return LaunchStatus();
}

namespace {

DmrsGenDev GenDev(const PuschMimoConfig &cfg, const Geom &g)
{
    DmrsGenDev d{};
    d.nl = g.nl;
    d.ndmrs = g.ndmrs;
    d.n_dmrs_re = g.n_dmrs_re;
    d.ref_pad = g.ref_pad;
    d.off = 6u * cfg.rb_start;
    for (uint32_t i = 0; i < g.ndmrs; ++i) {
        d.cinit[i] = DmrsCinit(cfg, g.dmrs_sym[i]);
        d.sym[i] = g.dmrs_sym[i];
    }
    for (uint32_t l = 0; l < g.nl; ++l) d.wf_odd[l] = (cfg.dmrs_ports[l] - 1000u) & 1u;
    d.nid = cfg.dmrs_scrambling_id;
    d.nscid = cfg.n_scid;
    return d;
}

}  // namespace

uint32_t HostDmrsCinit(const PuschMimoConfig &cfg, uint32_t symbol) { return DmrsCinit(cfg, symbol); }

int32_t LaunchDmrsRagged(const RaggedGrant *grants, const RaggedItem *gen_items, uint32_t n_gen,
                         const RaggedItem *ls_items, uint32_t n_ls, const RaggedItem *noise_items, uint32_t n_noise,
                         uint32_t max_ref_pad, uint32_t max_pilot_pad, uint32_t max_dmrs, const __half *rx_re,
                         const __half *rx_im, __half *ref_re, __half *ref_im, __half *h_re, __half *h_im,
                         __half *noise, hipStream_t s)
{
    if (n_gen == 0 || n_ls == 0 || n_noise == 0) return AIRAN_INVALID_ARGUMENT;
    hipLaunchKernelGGL(RaggedDmrsGenKernel, dim3(CeilDiv(CeilDiv(max_ref_pad, 16), 256), max_dmrs, n_gen), dim3(256),
                       0, s, reinterpret_cast<uint16_t *>(ref_re), reinterpret_cast<uint16_t *>(ref_im), grants,
                       gen_items);
    hipLaunchKernelGGL(RaggedDmrsLsKernel, dim3(CeilDiv(max_pilot_pad, 256), max_dmrs, n_ls), dim3(256), 0, s, rx_re,
                       rx_im, ref_re, ref_im, h_re, h_im, grants, ls_items);
    hipLaunchKernelGGL(RaggedDmrsNoiseKernel, dim3(n_noise), dim3(256), 0, s, h_re, h_im, noise, grants,
                       noise_items);
    return LaunchStatus();
}

int32_t LaunchDmrsGenSlots(const PuschMimoConfig &cfg, const Geom &g, const uint16_t *slot_no, uint32_t slots,
                           __half *ref_re, __half *ref_im, hipStream_t s)
{
    if (slot_no == nullptr || slots == 0) return AIRAN_INVALID_ARGUMENT;
    const DmrsGenDev d = GenDev(cfg, g);
    const dim3 grid(CeilDiv(CeilDiv(g.ref_pad, 16), 256), g.ndmrs, slots);
    const uint32_t vec = ((reinterpret_cast<uintptr_t>(ref_re) | reinterpret_cast<uintptr_t>(ref_im)) & 15u) == 0;
    hipLaunchKernelGGL(DmrsGenKernel, grid, dim3(256), 0, s, reinterpret_cast<uint16_t *>(ref_re),
                       reinterpret_cast<uint16_t *>(ref_im), slot_no, d, vec);
    return LaunchStatus();
}

}  // namespace airan

extern "C" {

// mimo_dmrs_gen.cpp:49-64 (TS 38.211 Table 6.4.1.1.3-1, ports 1000..1003)
int32_t airan_mimo_dmrs_gen_GetCurrentPortOccSemantics(uint16_t port, MimoDmrsPortOccSemantics *s)
{
    if (s == nullptr) return AIRAN_INVALID_ARGUMENT;
    if (port < 1000 || port > 1003) return AIRAN_UNSUPPORTED_PROFILE;
    const uint32_t idx = port - 1000u;
    s->port_index = idx;
    s->comb_delta = idx / 2u;
    s->wf_odd_negative = idx & 1u;
    s->wt_negative[0] = 0;
    s->wt_negative[1] = 0;
    return AIRAN_OK;
}

// mimo_dmrs_gen.cpp:66-132
int32_t airan_mimo_dmrs_gen_BuildCurrentProfile(const PuschMimoConfig *config, PuschMimoLayout *layout,
                                                MimoDmrsGenKernelMetadata *m, int32_t *cinit)
{
    if (config == nullptr || layout == nullptr || m == nullptr || cinit == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (config->flags != 0 || config->codeword_index != 0 ||
        (config->num_tx_ports != 1 && config->num_tx_ports != 2 && config->num_tx_ports != 4))
        return AIRAN_UNSUPPORTED_PROFILE;
    std::memset(m, 0, sizeof(*m));
    for (int i = 0; i < 16; ++i) cinit[i] = 0;
    m->magic = kDmrsGenMagic;
    m->num_layers = g.nl;
    m->num_dmrs_symbols = g.ndmrs;
    m->output_layer_stride = g.ndmrs * g.ref_pad;
    m->output_dmrs_stride = g.ref_pad;
    for (uint32_t d = 0; d < g.ndmrs; ++d) {
        m->dmrs_symbols[d] = g.dmrs_sym[d];
        cinit[d] = static_cast<int32_t>(DmrsCinit(*config, g.dmrs_sym[d]));
    }
    for (uint32_t l = 0; l < g.nl; ++l) {
        MimoDmrsPortOccSemantics s{};
        airan_mimo_dmrs_gen_GetCurrentPortOccSemantics(config->dmrs_ports[l], &s);
        m->port_indices[l] = s.port_index;
        m->comb_delta[l] = s.comb_delta;
        m->wf_odd_negative[l] = s.wf_odd_negative;
    }
    airan::FillLayout(g, layout);
    return AIRAN_OK;
}

int32_t airan_mimo_dmrs_gen_ValidateOpArgs(const MimoDmrsGenOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(MimoDmrsGenOpArgsV1)) || a->dmrs_re == nullptr ||
        a->dmrs_im == nullptr || a->config == nullptr || a->layout == nullptr ||
        a->stream == nullptr || a->dmrs_re == a->dmrs_im)
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    MimoDmrsGenKernelMetadata m{};
    int32_t cinit[16];
    int32_t st = airan_mimo_dmrs_gen_BuildCurrentProfile(a->config, &expected, &m, cinit);
    if (st != AIRAN_OK) return st;
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_INVALID_ARGUMENT;
}

int32_t airan_mimo_dmrs_gen_CreateRuntime(MimoDmrsGenRuntimeV1 **runtime)
{
    if (runtime == nullptr) return AIRAN_INVALID_ARGUMENT;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    if (airan::InitGoldTables() != AIRAN_OK) return AIRAN_LAUNCH_FAILED;
    *runtime = new MimoDmrsGenRuntimeV1{dev};
    return AIRAN_OK;
}

int32_t airan_mimo_dmrs_gen_Enqueue(MimoDmrsGenRuntimeV1 *runtime, const MimoDmrsGenOpArgsV1 *a)
{
    if (runtime == nullptr) return AIRAN_INVALID_ARGUMENT;
    int32_t st = airan_mimo_dmrs_gen_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    const DmrsGenDev d = airan::GenDev(*a->config, g);
    dim3 grid(airan::CeilDiv(airan::CeilDiv(g.ref_pad, 16), 256), g.ndmrs, 1);
    const uint32_t vec = ((reinterpret_cast<uintptr_t>(a->dmrs_re) | reinterpret_cast<uintptr_t>(a->dmrs_im)) & 15u) == 0;
    hipLaunchKernelGGL(DmrsGenKernel, grid, dim3(256), 0, static_cast<hipStream_t>(a->stream),
                       static_cast<uint16_t *>(a->dmrs_re), static_cast<uint16_t *>(a->dmrs_im), nullptr, d, vec);
    return airan::LaunchStatus();
}

void airan_mimo_dmrs_gen_DestroyRuntime(MimoDmrsGenRuntimeV1 *runtime) { delete runtime; }

// mimo_dmrs_ls.cpp:29-108
int32_t airan_mimo_dmrs_ls_BuildCurrentProfile(const PuschMimoConfig *config, PuschMimoLayout *layout,
                                               MimoDmrsLsKernelMetadata *m, uint16_t *pilot_count,
                                               uint16_t *pilot_sc)
{
    if (config == nullptr || layout == nullptr || m == nullptr || pilot_count == nullptr ||
        pilot_sc == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    std::memset(m, 0, sizeof(*m));
    m->magic = kDmrsLsMagic;
    m->num_rx = g.nr;
    m->num_layers = g.nl;
    m->num_dmrs_symbols = g.ndmrs;
    m->num_symbols = g.nsym;
    m->used_subcarriers = g.used;
    m->padded_subcarriers = g.pad;
    m->dmrs_ref_stride = g.ref_pad;
    m->pilot_stride = g.pilot_pad;
    m->block_dim = 256;
    for (uint32_t d = 0; d < g.ndmrs; ++d) m->dmrs_symbols[d] = g.dmrs_sym[d];
    for (int i = 0; i < 16; ++i) pilot_count[i] = 0;
    std::memset(pilot_sc, 0, sizeof(uint16_t) * g.nl * g.ndmrs * g.pilot_pad);
    for (uint32_t l = 0; l < g.nl; ++l) {
        m->ports[l] = config->dmrs_ports[l];
        m->comb_delta[l] = g.comb[l];
        m->observation_model[l] = g.occ[l];
        const uint32_t cnt = g.occ[l] ? g.n_occ : g.n_dmrs_re;
        for (uint32_t d = 0; d < g.ndmrs; ++d) {
            pilot_count[l * g.ndmrs + d] = static_cast<uint16_t>(cnt);
            uint16_t *dst = pilot_sc + (static_cast<size_t>(l) * g.ndmrs + d) * g.pilot_pad;
            for (uint32_t p = 0; p < cnt; ++p)
                dst[p] = static_cast<uint16_t>(g.comb[l] + (g.occ[l] ? 1u + 4u * p : 2u * p));
        }
    }
    airan::FillLayout(g, layout);
    return AIRAN_OK;
}

int32_t airan_mimo_dmrs_ls_ValidateOpArgs(const MimoDmrsLsOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(MimoDmrsLsOpArgsV1)) || a->rx_grid_re == nullptr ||
        a->rx_grid_im == nullptr || a->dmrs_ref_re == nullptr || a->dmrs_ref_im == nullptr ||
        a->h_ls_re == nullptr || a->h_ls_im == nullptr || a->pilot_sc == nullptr ||
        a->pilot_count == nullptr || a->noise_var_rx == nullptr || a->config == nullptr ||
        a->layout == nullptr || a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    const void *outs[5] = {a->h_ls_re, a->h_ls_im, a->pilot_sc, a->pilot_count, a->noise_var_rx};
    const void *ins[4] = {a->rx_grid_re, a->rx_grid_im, a->dmrs_ref_re, a->dmrs_ref_im};
    for (int i = 0; i < 5; ++i) {
        for (int j = 0; j < 4; ++j)
            if (outs[i] == ins[j]) return AIRAN_INVALID_ARGUMENT;
        for (int j = i + 1; j < 5; ++j)
            if (outs[i] == outs[j]) return AIRAN_INVALID_ARGUMENT;
    }
    Geom g;
    int32_t st = airan::DeriveGeom(*a->config, &g);
    if (st != AIRAN_OK) return st;
    PuschMimoLayout expected{};
    airan::FillLayout(g, &expected);
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_INVALID_ARGUMENT;
}

int32_t airan_mimo_dmrs_ls_Enqueue(const MimoDmrsLsOpArgsV1 *a)
{
    int32_t st = airan_mimo_dmrs_ls_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    hipStream_t s = static_cast<hipStream_t>(a->stream);
    const DevGeom dg = airan::ToDev(g);
    hipLaunchKernelGGL(PilotDescKernel, dim3(airan::CeilDiv(g.pilot_pad, 256), g.ndmrs, g.nl), dim3(256), 0, s,
                       static_cast<uint16_t *>(a->pilot_sc), static_cast<uint16_t *>(a->pilot_count), dg);
    return airan::LaunchDmrsLs(g, static_cast<const __half *>(a->rx_grid_re),
                               static_cast<const __half *>(a->rx_grid_im),
                               static_cast<const __half *>(a->dmrs_ref_re),
                               static_cast<const __half *>(a->dmrs_ref_im), static_cast<__half *>(a->h_ls_re),
                               static_cast<__half *>(a->h_ls_im), static_cast<__half *>(a->noise_var_rx), 1, s);
}

}  // extern "C"
