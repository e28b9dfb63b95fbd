// host_reference.cpp -- the operators' bit-exact host references and the
// DMRS-LS observation-contract validators (host side of the C ABI, no GPU).
//
// Each function restates the reference's host entry point of the same name
// for the geometry derived from PuschMimoConfig (the reference fixes it per
// profile); at the reference's profile the outputs are identical.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

namespace {

// Gold sequence c(n), TS 38.211 5.2.1 (Nc = 1600), one bit per call.
class Gold {
public:
    explicit Gold(uint32_t cinit)
    {
        x1_ = 1u;
        x2_ = cinit & 0x7fffffffu;
        for (int n = 0; n < 1600; ++n) Step();
    }
    uint32_t Next()
    {
        const uint32_t c = (x1_ ^ x2_) & 1u;
        Step();
        return c;
    }

private:
    void Step()
    {
        const uint32_t n1 = ((x1_ >> 3) ^ x1_) & 1u;
        const uint32_t n2 = ((x2_ >> 3) ^ (x2_ >> 2) ^ (x2_ >> 1) ^ x2_) & 1u;
        x1_ = (x1_ >> 1) | (n1 << 30);
        x2_ = (x2_ >> 1) | (n2 << 30);
    }
    uint32_t x1_, x2_;
};

int16_t SatProduct(int16_t v, int16_t s)
{
    const int32_t p = static_cast<int32_t>(v) * s;
    return static_cast<int16_t>(std::max(-32768, std::min(32767, p)));
}

int16_t ScaleClip8(int16_t v)
{
    const int32_t x = static_cast<int32_t>(v) * 8;
    return static_cast<int16_t>(std::max(-5120, std::min(5120, x)));
}

// Reference observation-model constants of the 133-PRB profile
// (channel_est/mimo_dmrs_ls/mimo_dmrs_ls.h:27-35), used when no config is
// given.
constexpr uint32_t kRefDmrsRe = 798, kRefOccPilot = 399, kRefPilotPad = 832, kRefDmrsSymbols = 2;

struct PilotGeom {
    uint32_t n_dmrs_re, n_occ, pilot_pad, ndmrs;
};

int32_t PilotGeomOf(const PuschMimoConfig *config, PilotGeom *pg)
{
    if (config == nullptr) {
        *pg = {kRefDmrsRe, kRefOccPilot, kRefPilotPad, kRefDmrsSymbols};
        return AIRAN_OK;
    }
    airan::Geom g;
    const int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    *pg = {g.n_dmrs_re, g.n_occ, g.pilot_pad, g.ndmrs};
    return AIRAN_OK;
}

int32_t RateReference(const int16_t *cw, const PuschMimoLayout &lay, uint32_t qm, uint32_t num_slots,
                      const AiranTbPlanV1 &p, const RateMatchDescriptor *desc, int16_t *out)
{
    const uint32_t N = 68u * p.Z, two_z = 2u * p.Z;
    const uint32_t fstart = p.Kprime - two_z;
    const uint64_t total_symbols = static_cast<uint64_t>(num_slots) * lay.codeword_symbols;
    std::fill(out, out + static_cast<size_t>(p.C) * N, int16_t{0});
    std::vector<int32_t> acc(p.Ncb);
    for (uint32_t cb = 0; cb < p.C; ++cb) {
        const RateMatchDescriptor &d = desc[cb];
        const uint32_t eq = d.e / qm;
        if (static_cast<uint64_t>(d.cw_symbol_offset) + eq > total_symbols) return AIRAN_LAYOUT_MISMATCH;
        std::fill(acc.begin(), acc.end(), 0);
        // Transmitted bit kk = b * eq + j (the bit interleaver) walks the
        // circular buffer from k0, skipping filler positions.
        uint32_t pos = d.k0;
        for (uint32_t kk = 0; kk < d.e; ++kk) {
            while (pos >= fstart && pos < fstart + p.filler) pos = pos + 1 == d.ncb ? 0 : pos + 1;
            const uint32_t b = kk / eq, j = kk - b * eq;
            const uint64_t logical = static_cast<uint64_t>(d.cw_symbol_offset) + j;
            const uint64_t slot = logical / lay.codeword_symbols, sym = logical % lay.codeword_symbols;
            acc[pos] += ScaleClip8(cw[(slot * qm + b) * lay.codeword_stride + sym]);
            pos = pos + 1 == d.ncb ? 0 : pos + 1;
        }
        int16_t *row = out + static_cast<size_t>(cb) * N;
        for (uint32_t k = 0; k < d.ncb && two_z + k < N; ++k) {
            const bool filler = k >= fstart && k < fstart + p.filler;
            row[two_z + k] = filler ? int16_t{5120}
                                    : static_cast<int16_t>(std::max(-5120, std::min(5120, acc[k])));
        }
    }
    return AIRAN_OK;
}

}  // namespace

extern "C" {

// mapping/re_demap_batch/re_demap_batch.cpp:125-163: rx[rx][symbol][u] =
// fft[rx][symbol][index[u] / 2] for the used subcarriers, zero tail.
int32_t airan_re_demap_batch_Reference(const PuschMimoConfig *config, const uint16_t *fft_grid_re,
                                       const uint16_t *fft_grid_im, uint32_t num_rx_antennas,
                                       const uint32_t *gather_index, uint16_t *rx_grid_re,
                                       uint16_t *rx_grid_im)
{
    if (config == nullptr || fft_grid_re == nullptr || fft_grid_im == nullptr || gather_index == nullptr ||
        rx_grid_re == nullptr || rx_grid_im == nullptr || num_rx_antennas == 0 ||
        num_rx_antennas > AIRAN_PUSCH_MIMO_MAX_RX_ANTENNAS || fft_grid_re == rx_grid_re ||
        fft_grid_re == rx_grid_im || fft_grid_im == rx_grid_re || fft_grid_im == rx_grid_im ||
        rx_grid_re == rx_grid_im)
        return AIRAN_INVALID_ARGUMENT;
    airan::Geom g;
    const int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    for (uint32_t u = 0; u < g.used; ++u)
        if ((gather_index[u] & 1u) != 0 || gather_index[u] >= g.fft * sizeof(uint16_t))
            return AIRAN_INVALID_ARGUMENT;
    for (uint32_t rx = 0; rx < num_rx_antennas; ++rx) {
        for (uint32_t s = 0; s < g.nsym; ++s) {
            const size_t in = (static_cast<size_t>(rx) * g.nsym + s) * g.fft;
            const size_t out = (static_cast<size_t>(rx) * g.nsym + s) * g.pad;
            for (uint32_t u = 0; u < g.used; ++u) {
                rx_grid_re[out + u] = fft_grid_re[in + gather_index[u] / 2];
                rx_grid_im[out + u] = fft_grid_im[in + gather_index[u] / 2];
            }
            std::fill(rx_grid_re + out + g.used, rx_grid_re + out + g.pad, uint16_t{0});
            std::fill(rx_grid_im + out + g.used, rx_grid_im + out + g.pad, uint16_t{0});
        }
    }
    return AIRAN_OK;
}

// mapping/layer_demap_mimo/layer_demap.cpp:126-157: cw[q][re L + l] =
// llr[l][q][data symbol][sc] (the QAM demapper's [L][Qm][14-D][qam_row]
// layout), codeword tails zero.
int32_t airan_layer_demap_ReferenceDemap(const int16_t *layer_llr, const PuschMimoConfig *config,
                                         const PuschMimoLayout *layout, int16_t *cw_llr)
{
    if (layer_llr == nullptr || config == nullptr || layout == nullptr || cw_llr == nullptr ||
        layer_llr == cw_llr)
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    LayerDemapKernelMetadata m{};
    const int32_t st = airan_layer_demap_BuildCurrentProfile(config, &expected, &m);
    if (st != AIRAN_OK) return st;
    if (!airan::SameLayout(expected, *layout)) return AIRAN_LAYOUT_MISMATCH;
    airan::Geom g;
    airan::DeriveGeom(*config, &g);
    std::fill(cw_llr, cw_llr + static_cast<size_t>(g.qm) * g.cw_stride, int16_t{0});
    for (uint32_t q = 0; q < g.qm; ++q)
        for (uint32_t ds = 0; ds < g.ndata; ++ds)
            for (uint32_t sc = 0; sc < g.used; ++sc) {
                const uint32_t re = ds * g.used + sc;
                for (uint32_t l = 0; l < g.nl; ++l)
                    cw_llr[static_cast<size_t>(q) * g.cw_stride + static_cast<size_t>(re) * g.nl + l] =
                        layer_llr[(static_cast<size_t>(l) * g.qm + q) * g.qam_plane +
                                  static_cast<size_t>(ds) * g.qam_row + sc];
            }
    return AIRAN_OK;
}

// fec/descramble_mimo/descramble_mimo.cpp:188-222: llr_nr[b][i] =
// sat16(llr_qam[stream(b)][i] * sign[b][i]) per slot, stream(b) = b / 2 for
// even b (I bits) and Qm / 2 + b / 2 for odd b (Q bits); tails zero.  sign
// may be NULL: the Gold sign of airan_descramble_mimo_BuildGoldSign (one
// sequence continuing across the slots) is generated here.
int32_t airan_descramble_mimo_ReferenceDescramble(const int16_t *llr_qam, const int16_t *sign,
                                                  const PuschMimoConfig *config, const PuschMimoLayout *layout,
                                                  uint32_t num_slots, int16_t *llr_nr)
{
    if (llr_qam == nullptr || config == nullptr || layout == nullptr || llr_nr == nullptr || llr_qam == llr_nr ||
        (sign != nullptr && sign == llr_nr))
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    DescrambleKernelMetadata m{};
    const int32_t st = airan_descramble_mimo_BuildCurrentProfile(config, num_slots, &expected, &m);
    if (st != AIRAN_OK) return st;
    if (!airan::SameLayout(expected, *layout)) return AIRAN_LAYOUT_MISMATCH;
    const uint32_t qm = config->qm, half = qm / 2;
    const size_t total = static_cast<size_t>(num_slots) * qm * layout->codeword_stride;
    std::fill(llr_nr, llr_nr + total, int16_t{0});
    const uint32_t cinit = (static_cast<uint32_t>(config->rnti) << 15) |
                           ((static_cast<uint32_t>(config->codeword_index) & 1u) << 14) | config->data_scrambling_id;
    Gold gold(cinit);
    for (uint32_t slot = 0; slot < num_slots; ++slot) {
        const size_t base = static_cast<size_t>(slot) * qm * layout->codeword_stride;
        for (uint32_t i = 0; i < layout->codeword_symbols; ++i)
            for (uint32_t b = 0; b < qm; ++b) {
                const size_t o = base + static_cast<size_t>(b) * layout->codeword_stride + i;
                const int16_t s = sign != nullptr ? sign[o] : (gold.Next() != 0 ? int16_t{-1} : int16_t{1});
                const uint32_t stream = (b & 1u) ? half + b / 2 : b / 2;
                llr_nr[o] = SatProduct(llr_qam[base + static_cast<size_t>(stream) * layout->codeword_stride + i], s);
            }
    }
    return AIRAN_OK;
}

// fec/rate_dematch_mimo/rate_dematch_mimo.cpp:153-201 for the reference's
// fixed plan (airan_pusch_ReferenceTbPlan): the descriptors must equal the
// profile's; ldpc_llr [C][68 Z].
int32_t airan_rate_dematch_mimo_ReferenceRateDematch(const int16_t *cw_llr, const PuschMimoConfig *config,
                                                     const PuschMimoLayout *layout, uint32_t num_slots,
                                                     const RateMatchDescriptor *descriptors,
                                                     size_t descriptor_count, int16_t *ldpc_llr)
{
    AiranTbPlanV1 p{};
    if (config == nullptr) return AIRAN_INVALID_ARGUMENT;
    const int32_t st = airan_pusch_ReferenceTbPlan(config, num_slots, &p);
    if (st != AIRAN_OK) return st;
    return airan_rate_dematch_mimo_ReferenceRateDematchForTb(cw_llr, config, layout, num_slots, &p, descriptors,
                                                             descriptor_count, ldpc_llr);
}

// The same for a TB plan (filler bits +5120, puncturing, any C): the
// descriptors must equal airan_rate_dematch_mimo_BuildProfileForTb's.
int32_t airan_rate_dematch_mimo_ReferenceRateDematchForTb(const int16_t *cw_llr, const PuschMimoConfig *config,
                                                          const PuschMimoLayout *layout, uint32_t num_slots,
                                                          const AiranTbPlanV1 *plan,
                                                          const RateMatchDescriptor *descriptors,
                                                          size_t descriptor_count, int16_t *ldpc_llr)
{
    if (cw_llr == nullptr || config == nullptr || layout == nullptr || plan == nullptr || descriptors == nullptr ||
        ldpc_llr == nullptr || cw_llr == ldpc_llr || descriptor_count < plan->C)
        return AIRAN_INVALID_ARGUMENT;
    PuschMimoLayout expected{};
    RateDematchKernelMetadata m{};
    std::vector<RateMatchDescriptor> want(plan->C);
    const int32_t st =
        airan_rate_dematch_mimo_BuildProfileForTb(config, num_slots, plan, &expected, &m, want.data(), want.size());
    if (st != AIRAN_OK) return st;
    if (!airan::SameLayout(expected, *layout) ||
        std::memcmp(want.data(), descriptors, want.size() * sizeof(RateMatchDescriptor)) != 0)
        return AIRAN_LAYOUT_MISMATCH;
    return RateReference(cw_llr, *layout, config->qm, num_slots, *plan, descriptors, ldpc_llr);
}

// channel_est/mimo_dmrs_ls/mimo_dmrs_ls.cpp:129-148: per layer, comb model
// (AIRAN_OBS_COMB2, one observation per comb pilot) or FD-OCC2 pair model
// (AIRAN_OBS_FD_OCC2, half as many), equal over the DMRS symbols.
int32_t airan_mimo_dmrs_ls_DescribeObservationModels(const PuschMimoConfig *config, const uint16_t *pilot_count,
                                                     uint32_t num_layers, uint32_t num_dmrs_symbols,
                                                     uint32_t *models)
{
    PilotGeom pg{};
    if (pilot_count == nullptr || models == nullptr || num_layers == 0 || num_layers > 4)
        return AIRAN_INVALID_ARGUMENT;
    const int32_t st = PilotGeomOf(config, &pg);
    if (st != AIRAN_OK) return st == AIRAN_UNSUPPORTED_PROFILE ? AIRAN_INVALID_ARGUMENT : st;
    if (num_dmrs_symbols != pg.ndmrs) return AIRAN_INVALID_ARGUMENT;
    for (uint32_t l = 0; l < num_layers; ++l) {
        const uint16_t count = pilot_count[l * num_dmrs_symbols];
        if (count == pg.n_dmrs_re)
            models[l] = AIRAN_OBS_COMB2;
        else if (count == pg.n_occ)
            models[l] = AIRAN_OBS_FD_OCC2;
        else
            return AIRAN_UNSUPPORTED_PROFILE;
        for (uint32_t d = 1; d < num_dmrs_symbols; ++d)
            if (pilot_count[l * num_dmrs_symbols + d] != count) return AIRAN_UNSUPPORTED_PROFILE;
    }
    return AIRAN_OK;
}

// mimo_dmrs_ls.cpp:150-182: every layer's pilot subcarriers are first +
// step * p (step 4 for FD-OCC2 with first 1 or 2, step 2 for the comb with
// first 0 or 1) for p < count and zero up to the pilot stride.
int32_t airan_mimo_dmrs_ls_ValidateNaturalLmmseContract(const PuschMimoConfig *config, const uint16_t *pilot_count,
                                                        const uint16_t *pilot_sc, uint32_t num_layers,
                                                        uint32_t num_dmrs_symbols)
{
    if (pilot_sc == nullptr) return AIRAN_INVALID_ARGUMENT;
    uint32_t models[4] = {};
    int32_t st = airan_mimo_dmrs_ls_DescribeObservationModels(config, pilot_count, num_layers, num_dmrs_symbols,
                                                              models);
    if (st != AIRAN_OK) return st;
    PilotGeom pg{};
    PilotGeomOf(config, &pg);
    for (uint32_t l = 0; l < num_layers; ++l) {
        const uint16_t count = pilot_count[l * num_dmrs_symbols];
        const bool occ = models[l] == AIRAN_OBS_FD_OCC2;
        const uint16_t step = occ ? 4u : 2u;
        for (uint32_t d = 0; d < num_dmrs_symbols; ++d) {
            const uint16_t *sc = pilot_sc + (static_cast<size_t>(l) * num_dmrs_symbols + d) * pg.pilot_pad;
            const uint16_t first = sc[0];
            if (occ ? (first != 1u && first != 2u) : (first != 0u && first != 1u)) return AIRAN_UNSUPPORTED_PROFILE;
            for (uint32_t p = 0; p < count; ++p)
                if (sc[p] != static_cast<uint16_t>(first + step * p)) return AIRAN_UNSUPPORTED_PROFILE;
            for (uint32_t p = count; p < pg.pilot_pad; ++p)
                if (sc[p] != 0u) return AIRAN_UNSUPPORTED_PROFILE;
        }
    }
    return AIRAN_OK;
}

// mimo_dmrs_ls.cpp:184-196: every layer on the comb model, else
// AIRAN_LAYOUT_MISMATCH (the reference's LMMSE_OBSERVATION_MODEL_MISMATCH,
// the same value -3).
int32_t airan_mimo_dmrs_ls_ValidateLmmse798Compatibility(const PuschMimoConfig *config, const uint16_t *pilot_count,
                                                         uint32_t num_layers, uint32_t num_dmrs_symbols)
{
    uint32_t models[4] = {};
    const int32_t st = airan_mimo_dmrs_ls_DescribeObservationModels(config, pilot_count, num_layers,
                                                                    num_dmrs_symbols, models);
    if (st != AIRAN_OK) return st;
    for (uint32_t l = 0; l < num_layers; ++l)
        if (models[l] != AIRAN_OBS_COMB2) return AIRAN_LAYOUT_MISMATCH;
    return AIRAN_OK;
}

}  // extern "C"
