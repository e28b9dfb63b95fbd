// op_integer.hip -- integer and copy stages of the PUSCH receive chain:
// RE demap, QAM max-log demap, layer demap, descramble and rate dematch.
// All are HBM-bound gathers/maps: one thread per output element (or per
// 16-symbol group for descramble), coalesced along the contiguous axis.
#include <cmath>

#include "common.h"
#include "gold.h"
#include "launch.h"
#include "qam.h"
#include "ragged.h"
#include "ratematch.h"

using airan::DevGeom;
using airan::Geom;

// =================================================================== RE demap
namespace {

constexpr uint32_t kReDemapMagic = 0x52444231u;  // "RDB1"

// mapping/re_demap_batch/re_demap_batch_kernel.cpp:129 -- gather of the used
// subcarriers from the stage-4 layout through the caller's byte-offset table.
__global__ void __launch_bounds__(256) ReDemapKernel(const uint16_t *__restrict__ fre,
                                                     const uint16_t *__restrict__ fim,
                                                     const uint32_t *__restrict__ index,
                                                     uint16_t *__restrict__ rre,
                                                     uint16_t *__restrict__ rim, uint32_t fft,
                                                     uint32_t used, uint32_t pad, uint32_t nsym)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= pad) return;
    const size_t row = static_cast<size_t>(blockIdx.z) * nsym + blockIdx.y;
    uint16_t vr = 0, vi = 0;
    if (u < used) {
        const uint32_t e = index[u] >> 1;
        if (e < fft) {
            vr = fre[row * fft + e];
            vi = fim[row * fft + e];
        }
    }
    rre[row * pad + u] = vr;
    rim[row * pad + u] = vi;
}

}  // namespace

extern "C" {

int32_t airan_re_demap_batch_BuildCurrentProfile(const PuschMimoConfig *config,
                                                 PuschMimoLayout *layout,
                                                 ReDemapKernelMetadata *meta)
{
    if (config == nullptr || layout == nullptr || meta == nullptr) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (config->flags != 0) return AIRAN_UNSUPPORTED_PROFILE;
    airan::FillLayout(g, layout);
    std::memset(meta, 0, sizeof(*meta));
    meta->magic = kReDemapMagic;
    meta->num_rx_antennas = g.nr;
    meta->num_symbols = g.nsym;
    meta->fft_size = g.fft;
    meta->used_subcarriers = g.used;
    meta->padded_subcarriers = g.pad;
    meta->block_dim = 256;
    meta->symbols_per_tile = 1;
    return AIRAN_OK;
}

int32_t airan_re_demap_batch_ValidateOpArgs(const ReDemapBatchOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(ReDemapBatchOpArgsV1)) || a->fft_grid_re == nullptr ||
        a->fft_grid_im == nullptr || a->rx_grid_re == nullptr || a->rx_grid_im == nullptr ||
        a->config == nullptr || a->layout == nullptr || a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    if (a->fft_grid_re == a->rx_grid_re || a->fft_grid_re == a->rx_grid_im ||
        a->fft_grid_im == a->rx_grid_re || a->fft_grid_im == a->rx_grid_im ||
        a->rx_grid_re == a->rx_grid_im)
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    ReDemapKernelMetadata meta{};
    int32_t st = airan_re_demap_batch_BuildCurrentProfile(a->config, &expected, &meta);
    if (st != AIRAN_OK) return st;
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_LAYOUT_MISMATCH;
}

// mapping/re_demap_batch/re_demap_batch.cpp:105-123 for fft = P * 64:
// used index u -> bin (fftshift order) -> stage-4 element (bin % P) * 64 + bin / P.
int32_t airan_re_demap_batch_BuildGatherIndex(const PuschMimoConfig *config, uint32_t *index,
                                              size_t index_elems)
{
    if (config == nullptr || index == nullptr) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (index_elems < g.pad) return AIRAN_RESOURCE_TOO_SMALL;
    for (uint32_t u = 0; u < g.pad; ++u) index[u] = 0;
    for (uint32_t u = 0; u < g.used; ++u) {
        uint32_t bin = u < g.used / 2 ? g.fft - g.used / 2 + u : u - g.used / 2;
        uint32_t element = (bin % g.P) * g.Q + bin / g.P;
        index[u] = element * 2u;
    }
    return AIRAN_OK;
}

int32_t airan_re_demap_batch_Enqueue(const ReDemapBatchOpArgsV1 *a, const void *gather_index,
                                     size_t gather_index_bytes, void *workspace,
                                     size_t workspace_bytes, const void *tiling,
                                     size_t tiling_bytes)
{
    int32_t st = airan_re_demap_batch_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    (void)workspace;
    (void)workspace_bytes;
    if (gather_index == nullptr || tiling == nullptr ||
        gather_index_bytes < g.pad * sizeof(uint32_t) || tiling_bytes < sizeof(ReDemapKernelMetadata))
        return AIRAN_RESOURCE_TOO_SMALL;
    dim3 grid(airan::CeilDiv(g.pad, 256), g.nsym, g.nr);
    hipLaunchKernelGGL(ReDemapKernel, grid, dim3(256), 0, static_cast<hipStream_t>(a->stream),
                       static_cast<const uint16_t *>(a->fft_grid_re),
                       static_cast<const uint16_t *>(a->fft_grid_im),
                       static_cast<const uint32_t *>(gather_index),
                       static_cast<uint16_t *>(a->rx_grid_re), static_cast<uint16_t *>(a->rx_grid_im),
                       g.fft, g.used, g.pad, g.nsym);
    return airan::LaunchStatus();
}

}  // extern "C"

// ================================================================ QAM demap
namespace {

using airan::Axis256;
using airan::AxisExhaustive;
using airan::QamUT;

// qam256_demod_batch_kernel.cpp:36-183: one thread per (layer, data symbol,
// subcarrier); writes Qm planes of [L,Qm,ndata,qam_row].
__global__ void __launch_bounds__(256) QamDemodKernel(const __half *__restrict__ xr,
                                                      const __half *__restrict__ xi,
                                                      const __half *__restrict__ ne,
                                                      int16_t *__restrict__ llr, DevGeom g,
                                                      uint32_t layer_stride)
{
    const uint32_t sc = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc >= g.qam_row) return;
    const uint32_t ds = blockIdx.y, l = blockIdx.z;
    const uint32_t m = g.qm / 2;
    int16_t *out = llr + static_cast<size_t>(l) * g.qm * g.qam_plane + static_cast<size_t>(ds) * g.qam_row + sc;
    if (sc >= g.used) {
        for (uint32_t q = 0; q < g.qm; ++q) out[static_cast<size_t>(q) * g.qam_plane] = 0;
        return;
    }
    const size_t in = static_cast<size_t>(l) * layer_stride + static_cast<size_t>(g.data_sym[ds]) * g.pad + sc;
    const __half n = ne[in];
    int ui, ti, uq, tq;
    QamUT(g.qm, xr[in], n, &ui, &ti);
    QamUT(g.qm, xi[in], n, &uq, &tq);
    int a[4], b[4];
    if (g.qm == 8) {
        Axis256(ui, ti, a);
        Axis256(uq, tq, b);
    } else {
        AxisExhaustive(static_cast<int>(m), ui, ti, a);
        AxisExhaustive(static_cast<int>(m), uq, tq, b);
    }
    for (uint32_t i = 0; i < m; ++i) {
        out[static_cast<size_t>(i) * g.qam_plane] = static_cast<int16_t>(a[i]);
        out[static_cast<size_t>(m + i) * g.qam_plane] = static_cast<int16_t>(b[i]);
    }
}

}  // namespace

extern "C" {

int32_t airan_qam256_demod_batch_BuildCurrentProfile(const PuschMimoConfig *config,
                                                     PuschMimoLayout *layout,
                                                     QamDemodBatchTilingData *t)
{
    if (config == nullptr || layout == nullptr || t == nullptr) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    airan::FillLayout(g, layout);
    std::memset(t, 0, sizeof(*t));
    t->n_sc_used = static_cast<int32_t>(g.used);
    t->n_sc_pad = static_cast<int32_t>(g.qam_row);
    t->n_data_symbols = static_cast<int32_t>(g.ndata);
    t->n_data_re = static_cast<int32_t>(g.data_re);
    t->llr_layer_stride = static_cast<int32_t>(g.qam_plane);
    t->block_dim = 256;
    t->llr_clip = 2560;
    t->y_scaled_clip = 1000;
    t->num_layers = static_cast<int32_t>(g.nl);
    t->grid_layer_stride = static_cast<int32_t>(g.nsym * g.pad);
    for (uint32_t i = 0; i < g.ndata; ++i) t->physical_data_symbols[i] = static_cast<int32_t>(g.data_sym[i]);
    t->group_size = static_cast<int32_t>(g.qm);
    return AIRAN_OK;
}

int32_t airan_qam256_demod_batch_ValidateOpArgs(const QamDemod256BatchOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(QamDemod256BatchOpArgsV1)) || a->x_re == nullptr ||
        a->x_im == nullptr || a->no_eff == nullptr || a->layer_llr == nullptr ||
        a->config == nullptr || a->layout == nullptr || a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    if (a->layer_llr == a->x_re || a->layer_llr == a->x_im || a->layer_llr == a->no_eff)
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    QamDemodBatchTilingData t{};
    int32_t st = airan_qam256_demod_batch_BuildCurrentProfile(a->config, &expected, &t);
    if (st != AIRAN_OK) return st;
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_LAYOUT_MISMATCH;
}

int32_t airan_qam256_demod_batch_Enqueue(const QamDemod256BatchOpArgsV1 *a, void *workspace,
                                         size_t workspace_bytes, const void *tiling,
                                         size_t tiling_bytes)
{
    int32_t st = airan_qam256_demod_batch_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    (void)workspace;
    (void)workspace_bytes;
    if (tiling == nullptr || tiling_bytes < sizeof(QamDemodBatchTilingData)) return AIRAN_RESOURCE_TOO_SMALL;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    dim3 grid(airan::CeilDiv(g.qam_row, 256), g.ndata, g.nl);
    hipLaunchKernelGGL(QamDemodKernel, grid, dim3(256), 0, static_cast<hipStream_t>(a->stream),
                       static_cast<const __half *>(a->x_re), static_cast<const __half *>(a->x_im),
                       static_cast<const __half *>(a->no_eff), static_cast<int16_t *>(a->layer_llr),
                       airan::ToDev(g), g.nsym * g.pad);
    return airan::LaunchStatus();
}

}  // extern "C"

// ============================================================== layer demap
namespace {

constexpr uint32_t kLayerDemapMagic = 0x4c444d31u;  // "LDM1"

// mapping/layer_demap_mimo/layer_demap.cpp:126-157: cw[q][re*L + l] =
// llr[l][q][ds][sc]; one thread per (q, re) writes the L-vector of its RE.
__global__ void __launch_bounds__(256) LayerDemapKernel(const int16_t *__restrict__ llr,
                                                        int16_t *__restrict__ cw, DevGeom g)
{
    const uint32_t re = blockIdx.x * blockDim.x + threadIdx.x;
    if (re >= g.data_stride) return;
    const uint32_t q = blockIdx.y;
    int16_t *out = cw + static_cast<size_t>(q) * g.cw_stride + static_cast<size_t>(re) * g.nl;
    if (re >= g.data_re) {
        for (uint32_t l = 0; l < g.nl; ++l) out[l] = 0;
        return;
    }
    const uint32_t ds = re / g.used, sc = re - ds * g.used;
    const int16_t *in = llr + static_cast<size_t>(q) * g.qam_plane + static_cast<size_t>(ds) * g.qam_row + sc;
    const size_t lstride = static_cast<size_t>(g.qm) * g.qam_plane;
    if (g.nl == 4) {
        short4 v;
        v.x = in[0];
        v.y = in[lstride];
        v.z = in[2 * lstride];
        v.w = in[3 * lstride];
        *reinterpret_cast<short4 *>(out) = v;
    } else if (g.nl == 2) {
        short2 v;
        v.x = in[0];
        v.y = in[lstride];
        *reinterpret_cast<short2 *>(out) = v;
    } else {
        for (uint32_t l = 0; l < g.nl; ++l) out[l] = in[l * lstride];
    }
}

}  // namespace

extern "C" {

int32_t airan_layer_demap_BuildCurrentProfile(const PuschMimoConfig *config, PuschMimoLayout *layout,
                                              LayerDemapKernelMetadata *m)
{
    if (config == nullptr || layout == nullptr || m == nullptr) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    airan::FillLayout(g, layout);
    std::memset(m, 0, sizeof(*m));
    m->magic = kLayerDemapMagic;
    m->num_layers = g.nl;
    m->qm = g.qm;
    m->num_data_re = g.data_re;
    m->data_stride = g.data_stride;
    m->codeword_symbols = g.cw_symbols;
    m->codeword_stride = g.cw_stride;
    m->input_symbol_stride = g.qam_row;
    m->symbols_per_group = 1;
    m->group_data_re = g.used;
    m->gather_chunk = 256;
    m->gather_index_elems = 0;
    return AIRAN_OK;
}

int32_t airan_layer_demap_ValidateOpArgs(const LayerDemapOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(LayerDemapOpArgsV1)) || a->layer_llr == nullptr ||
        a->cw_llr == nullptr || a->config == nullptr || a->layout == nullptr || a->stream == nullptr ||
        a->layer_llr == a->cw_llr)
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    LayerDemapKernelMetadata m{};
    int32_t st = airan_layer_demap_BuildCurrentProfile(a->config, &expected, &m);
    if (st != AIRAN_OK) return st;
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_LAYOUT_MISMATCH;
}

int32_t airan_layer_demap_Enqueue(const LayerDemapOpArgsV1 *a)
{
    int32_t st = airan_layer_demap_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    dim3 grid(airan::CeilDiv(g.data_stride, 256), g.qm);
    hipLaunchKernelGGL(LayerDemapKernel, grid, dim3(256), 0, static_cast<hipStream_t>(a->stream),
                       static_cast<const int16_t *>(a->layer_llr), static_cast<int16_t *>(a->cw_llr),
                       airan::ToDev(g));
    return airan::LaunchStatus();
}

}  // extern "C"

// =============================================================== descramble
namespace {

constexpr uint32_t kDescrMagic = 0x444d4d31u;  // "DMM1"
constexpr uint32_t kDescrSymbolsPerThread = 16;

__device__ inline int16_t SatNeg(int16_t v) { return v == -32768 ? static_cast<int16_t>(32767) : static_cast<int16_t>(-v); }

// fec/descramble_mimo/descramble_mimo.cpp:188-222 with the Gold sign of
// :155-186 generated in registers: a thread owns 16 codeword symbols, i.e.
// 16*Qm consecutive Gold bits (a multiple of 32 for every Qm).
__global__ void __launch_bounds__(256) DescrambleKernel(const int16_t *__restrict__ in,
                                                        const int16_t *__restrict__ sign,
                                                        int16_t *__restrict__ out, uint32_t qm,
                                                        uint32_t cw_symbols, uint32_t cw_stride,
                                                        uint32_t cinit, uint32_t restart)
{
    const uint32_t g0 = (blockIdx.x * blockDim.x + threadIdx.x) * kDescrSymbolsPerThread;
    if (g0 >= cw_stride) return;
    const uint32_t slot = blockIdx.y;
    const uint32_t m = qm / 2;
    const size_t sbase = static_cast<size_t>(slot) * qm * cw_stride;
    if (g0 >= cw_symbols) {
        for (uint32_t b = 0; b < qm; ++b)
            for (uint32_t s = g0; s < g0 + kDescrSymbolsPerThread && s < cw_stride; ++s)
                out[sbase + static_cast<size_t>(b) * cw_stride + s] = 0;
        return;
    }
    const uint32_t nsym = min(kDescrSymbolsPerThread, cw_symbols - g0);
    uint32_t bitsw[4] = {0, 0, 0, 0};
    if (sign == nullptr) {
        const uint64_t start = (restart ? 0ull : static_cast<uint64_t>(slot) * cw_symbols * qm) +
                               static_cast<uint64_t>(g0) * qm;
        uint32_t s1, s2;
        airan::GoldStateAt(start + 1600ull, cinit, &s1, &s2);
        for (uint32_t w = 0; w < m; ++w) bitsw[w] = airan::GoldNext32(&s1, &s2);
    }
    for (uint32_t b = 0; b < qm; ++b) {
        const uint32_t stream = (b & 1u) ? m + b / 2 : b / 2;
        const int16_t *src = in + sbase + static_cast<size_t>(stream) * cw_stride + g0;
        int16_t *dst = out + sbase + static_cast<size_t>(b) * cw_stride + g0;
        const int16_t *sg = sign ? sign + sbase + static_cast<size_t>(b) * cw_stride + g0 : nullptr;
        for (uint32_t s = 0; s < nsym; ++s) {
            const int16_t v = src[s];
            bool flip;
            if (sg != nullptr) {
                flip = sg[s] < 0;
            } else {
                const uint32_t bit = s * qm + b;
                flip = (bitsw[bit >> 5] >> (bit & 31u)) & 1u;
            }
            dst[s] = flip ? SatNeg(v) : v;
        }
        for (uint32_t s = nsym; s < kDescrSymbolsPerThread && g0 + s < cw_stride; ++s) dst[s] = 0;
    }
}

}  // namespace

extern "C" {

int32_t airan_descramble_mimo_BuildCurrentProfile(const PuschMimoConfig *config, uint32_t num_slots,
                                                  PuschMimoLayout *layout,
                                                  DescrambleKernelMetadata *m)
{
    if (config == nullptr || layout == nullptr || m == nullptr || num_slots == 0 || num_slots > 1024)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (config->data_scrambling_id > 1023 || config->codeword_index > 1) return AIRAN_UNSUPPORTED_PROFILE;
    airan::FillLayout(g, layout);
    std::memset(m, 0, sizeof(*m));
    m->magic = kDescrMagic;
    m->num_slots = num_slots;
    m->num_layers = g.nl;
    m->qm = g.qm;
    m->num_data_re = g.data_re;
    m->data_stride = g.data_stride;
    m->codeword_symbols = g.cw_symbols;
    m->codeword_stride = g.cw_stride;
    m->tile_elems = kDescrSymbolsPerThread;
    m->max_slots = 1024;
    return AIRAN_OK;
}

int32_t airan_descramble_mimo_ValidateOpArgs(const DescrambleMimoOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, offsetof(DescrambleMimoOpArgsV1, flags)) || a->llr_qam == nullptr ||
        a->llr_nr == nullptr || a->config == nullptr || a->layout == nullptr || a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    if (a->llr_qam == a->llr_nr) return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    DescrambleKernelMetadata m{};
    int32_t st = airan_descramble_mimo_BuildCurrentProfile(a->config, a->num_slots, &expected, &m);
    if (st != AIRAN_OK) return st;
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_LAYOUT_MISMATCH;
}

// descramble_mimo.cpp:155-186 (host sign table, kept for callers that pass one).
int32_t airan_descramble_mimo_BuildGoldSign(const PuschMimoConfig *config, const PuschMimoLayout *layout,
                                            uint32_t num_slots, int16_t *sign, size_t sign_elems)
{
    if (sign == nullptr || layout == nullptr) return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    DescrambleKernelMetadata m{};
    int32_t st = airan_descramble_mimo_BuildCurrentProfile(config, num_slots, &expected, &m);
    if (st != AIRAN_OK) return st;
    if (!airan::SameLayout(expected, *layout)) return AIRAN_LAYOUT_MISMATCH;
    const size_t required = static_cast<size_t>(num_slots) * m.qm * layout->codeword_stride;
    if (sign_elems < required) return AIRAN_RESOURCE_TOO_SMALL;
    for (size_t i = 0; i < required; ++i) sign[i] = 1;
    const uint32_t cinit = (static_cast<uint32_t>(config->rnti) << 15) |
                           (static_cast<uint32_t>(config->codeword_index & 1u) << 14) |
                           config->data_scrambling_id;
    uint8_t x1[31] = {1}, x2[31];
    for (int b = 0; b < 31; ++b) x2[b] = static_cast<uint8_t>((cinit >> b) & 1u);
    uint32_t pos = 0;
    auto step = [&]() {
        uint32_t p1 = (pos + 1) % 31, p2 = (pos + 2) % 31, p3 = (pos + 3) % 31;
        uint8_t n1 = x1[p3] ^ x1[pos];
        uint8_t n2 = x2[p3] ^ x2[p2] ^ x2[p1] ^ x2[pos];
        x1[pos] = n1;
        x2[pos] = n2;
        pos = (pos + 1) % 31;
    };
    for (int n = 0; n < 1600; ++n) step();
    for (uint32_t s = 0; s < num_slots; ++s) {
        const size_t base = static_cast<size_t>(s) * m.qm * layout->codeword_stride;
        for (uint32_t sym = 0; sym < layout->codeword_symbols; ++sym)
            for (uint32_t b = 0; b < m.qm; ++b) {
                uint8_t c = x1[pos] ^ x2[pos];
                step();
                sign[base + static_cast<size_t>(b) * layout->codeword_stride + sym] = c ? -1 : 1;
            }
    }
    return AIRAN_OK;
}

int32_t airan_descramble_mimo_Enqueue(const DescrambleMimoOpArgsV1 *a, const void *sign,
                                      size_t sign_bytes, void *workspace, size_t workspace_bytes,
                                      const void *tiling, size_t tiling_bytes)
{
    int32_t st = airan_descramble_mimo_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    (void)workspace;
    (void)workspace_bytes;
    (void)tiling;
    (void)tiling_bytes;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    const size_t required = static_cast<size_t>(a->num_slots) * g.qm * g.cw_stride * sizeof(int16_t);
    if (sign != nullptr && (sign_bytes < required || sign == a->llr_qam || sign == a->llr_nr))
        return AIRAN_RESOURCE_TOO_SMALL;
    if (sign == nullptr && (st = airan::InitGoldTables()) != AIRAN_OK) return st;
    const uint32_t restart = a->struct_size >= sizeof(DescrambleMimoOpArgsV1)
                                 ? (a->flags & AIRAN_DESCR_RESTART_PER_SLOT)
                                 : 0u;
    const uint32_t cinit = (static_cast<uint32_t>(a->config->rnti) << 15) |
                           (static_cast<uint32_t>(a->config->codeword_index & 1u) << 14) |
                           a->config->data_scrambling_id;
    dim3 grid(airan::CeilDiv(airan::CeilDiv(g.cw_stride, kDescrSymbolsPerThread), 256), a->num_slots);
    hipLaunchKernelGGL(DescrambleKernel, grid, dim3(256), 0, static_cast<hipStream_t>(a->stream),
                       static_cast<const int16_t *>(a->llr_qam), static_cast<const int16_t *>(sign),
                       static_cast<int16_t *>(a->llr_nr), g.qm, g.cw_symbols, g.cw_stride, cinit, restart);
    return airan::LaunchStatus();
}

}  // extern "C"

// ============================================================ rate dematch
namespace {

constexpr uint32_t kRateMagic = 0x52444d31u;  // "RDM1"

// One thread per decoder input position p of code block blockIdx.y.
__global__ void __launch_bounds__(256) RateDematchKernel(const int16_t *__restrict__ in,
                                                         const RateMatchDescriptor *__restrict__ desc,
                                                         int16_t *__restrict__ out, airan::RateDev r)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 68u * r.z;
    if (p >= n) return;
    const uint32_t cb = blockIdx.y;
    const airan::RateCb c = airan::RateCbSetup(desc[cb], r);
    out[static_cast<size_t>(cb) * n + p] = airan::RateDematchValue(p, c, r, in);
}

// Ragged batch: code block blockIdx.y of any grant, its geometry from the
// grant's record, its decoder input at its own offset.
__global__ void __launch_bounds__(256) RaggedRateDematchKernel(const int16_t *__restrict__ cw,
                                                               const airan::RaggedGrant *__restrict__ grants,
                                                               const airan::RaggedCb *__restrict__ cbs,
                                                               int16_t *__restrict__ out)
{
    const airan::RaggedCb c = cbs[blockIdx.y];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= 68u * c.z) return;
    const airan::RaggedGrant &gr = grants[c.grant];
    const airan::RateDev r{gr.g.qm, gr.g.cw_symbols, gr.g.cw_stride, c.z, c.kprime, c.filler, c.ncb, 1};
    const RateMatchDescriptor desc{c.e, c.k0, c.ncb, c.cw_symbol_offset};
    out[c.lam_off + p] = airan::RateDematchValue(p, airan::RateCbSetup(desc, r), r, cw + gr.cw_off);
}

int32_t BuildRateProfile(const PuschMimoConfig *config, uint32_t num_slots, const AiranTbPlanV1 &p,
                         PuschMimoLayout *layout, RateDematchKernelMetadata *m,
                         RateMatchDescriptor *desc, size_t desc_count)
{
    if (config == nullptr || layout == nullptr || m == nullptr || desc == nullptr || num_slots == 0)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (desc_count < p.C) return AIRAN_INVALID_ARGUMENT;
    airan::FillLayout(g, layout);
    const uint64_t total_bits = static_cast<uint64_t>(num_slots) * g.qm * g.cw_symbols;
    uint64_t off = 0, e_sum = 0;
    for (uint32_t cb = 0; cb < p.C; ++cb) {
        const uint32_t e = cb < p.n_low ? p.E_low : p.E_high;
        desc[cb] = {e, p.k0, p.Ncb, static_cast<uint32_t>(off)};
        off += e / g.qm;
        e_sum += e;
    }
    if (e_sum != total_bits || p.G != total_bits) return AIRAN_UNSUPPORTED_PROFILE;
    std::memset(m, 0, sizeof(*m));
    m->magic = kRateMagic;
    m->num_slots = num_slots;
    m->num_layers = g.nl;
    m->qm = g.qm;
    m->num_data_re = g.data_re;
    m->codeword_symbols = g.cw_symbols;
    m->codeword_stride = g.cw_stride;
    m->total_coded_bits = static_cast<uint32_t>(total_bits);
    m->code_blocks = p.C;
    m->ldpc_n = 68 * p.Z;
    m->n_2z = 2 * p.Z;
    m->ncb = p.Ncb;
    m->llr_scale = 8;
    m->llr_clip = 5120;
    m->tile_elems = 256;
    m->max_slots = 1024;
    m->descriptor_words = 4 * p.C;
    m->z = p.Z;
    m->kprime = p.Kprime;
    m->filler = p.filler;
    return AIRAN_OK;
}

}  // namespace

namespace airan {

int32_t LaunchRateDematch(const Geom &g, const AiranTbPlanV1 &p, const RateMatchDescriptor *desc,
                          const int16_t *in, int16_t *out, uint32_t num_cb, hipStream_t s)
{
    if (num_cb == 0 || num_cb > 65535 || p.Z == 0 || p.Ncb <= p.filler) return AIRAN_INVALID_ARGUMENT;
    const RateDev r{g.qm, g.cw_symbols, g.cw_stride, p.Z, p.Kprime, p.filler, p.Ncb, num_cb};
    hipLaunchKernelGGL(RateDematchKernel, dim3(CeilDiv(68u * p.Z, 256), num_cb), dim3(256), 0, s, in, desc, out, r);
    return LaunchStatus();
}

int32_t LaunchRateDematchRagged(const RaggedGrant *grants, const RaggedCb *cbs, uint32_t num_cb, const int16_t *cw,
                                int16_t *out, hipStream_t s)
{
    if (num_cb == 0 || num_cb > 65535) return AIRAN_INVALID_ARGUMENT;
    hipLaunchKernelGGL(RaggedRateDematchKernel, dim3(CeilDiv(68u * 384u, 256), num_cb), dim3(256), 0, s, cw, grants,
                       cbs, out);
    return LaunchStatus();
}

}  // namespace airan

extern "C" {

int32_t airan_rate_dematch_mimo_BuildCurrentProfile(const PuschMimoConfig *config, uint32_t num_slots,
                                                    PuschMimoLayout *layout,
                                                    RateDematchKernelMetadata *m,
                                                    RateMatchDescriptor *desc, size_t desc_count)
{
    AiranTbPlanV1 p{};
    int32_t st = airan_pusch_ReferenceTbPlan(config, num_slots, &p);
    if (st != AIRAN_OK) return st;
    return BuildRateProfile(config, num_slots, p, layout, m, desc, desc_count);
}

int32_t airan_rate_dematch_mimo_BuildProfileForTb(const PuschMimoConfig *config, uint32_t num_slots,
                                                  const AiranTbPlanV1 *plan, PuschMimoLayout *layout,
                                                  RateDematchKernelMetadata *m,
                                                  RateMatchDescriptor *desc, size_t desc_count)
{
    if (plan == nullptr) return AIRAN_INVALID_ARGUMENT;
    return BuildRateProfile(config, num_slots, *plan, layout, m, desc, desc_count);
}

int32_t airan_rate_dematch_mimo_ValidateOpArgs(const RateDematchMimoOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(RateDematchMimoOpArgsV1)) || a->cw_llr == nullptr ||
        a->ldpc_llr == nullptr || a->config == nullptr || a->layout == nullptr ||
        a->stream == nullptr || a->cw_llr == a->ldpc_llr || a->num_slots == 0)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*a->config, &g);
    if (st != AIRAN_OK) return st;
    PuschMimoLayout expected{};
    airan::FillLayout(g, &expected);
    return airan::SameLayout(expected, *a->layout) ? AIRAN_OK : AIRAN_LAYOUT_MISMATCH;
}

int32_t airan_rate_dematch_mimo_Enqueue(const RateDematchMimoOpArgsV1 *a,
                                        const RateDematchKernelMetadata *m, const void *descriptors,
                                        size_t descriptor_bytes, void *workspace, size_t workspace_bytes,
                                        const void *tiling, size_t tiling_bytes)
{
    int32_t st = airan_rate_dematch_mimo_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    (void)workspace;
    (void)workspace_bytes;
    if (m == nullptr || m->magic != kRateMagic || m->num_slots != a->num_slots || m->z == 0 ||
        m->z > 384 || m->ncb > 66 * m->z || m->ncb <= m->filler || m->kprime > 22 * m->z ||
        m->kprime < 2 * m->z ||
        m->filler > 22 * m->z - m->kprime || m->code_blocks == 0 || m->code_blocks > 65535)
        return AIRAN_INVALID_ARGUMENT;
    if (descriptors == nullptr || tiling == nullptr ||
        descriptor_bytes < m->code_blocks * sizeof(RateMatchDescriptor) ||
        tiling_bytes < sizeof(RateDematchKernelMetadata))
        return AIRAN_RESOURCE_TOO_SMALL;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    if (m->qm != g.qm || m->codeword_symbols != g.cw_symbols || m->codeword_stride != g.cw_stride)
        return AIRAN_LAYOUT_MISMATCH;
    airan::RateDev r{g.qm, g.cw_symbols, g.cw_stride, m->z, m->kprime, m->filler, m->ncb, m->code_blocks};
    dim3 grid(airan::CeilDiv(68u * m->z, 256), m->code_blocks);
    hipLaunchKernelGGL(RateDematchKernel, grid, dim3(256), 0, static_cast<hipStream_t>(a->stream),
                       static_cast<const int16_t *>(a->cw_llr),
                       static_cast<const RateMatchDescriptor *>(descriptors),
                       static_cast<int16_t *>(a->ldpc_llr), r);
    return airan::LaunchStatus();
}

}  // extern "C"
