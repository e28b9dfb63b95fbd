// ref_shim.cpp -- extern "C" entry points over the reference's own host C++
// sources (compiled from /root/reference by oracle/ref/Makefile), used only
// by tests to pin the oracle against the reference at its fixed profile.
#include <cstring>
#include <string>

#include "pusch_mimo_runtime_config.h"
#include "re_demap_batch.h"
#include "layer_demap.h"
#include "mimo_dmrs_gen.h"
#include "mimo_dmrs_ls.h"
#include "channel_est_lmmse.h"

extern "C" {

int ref_validate_runtime_config(const void *cfg, char *why, int len)
{
    std::string w;
    auto st = airan::ValidateMimoRuntimeConfig(*static_cast<const airan::PuschMimoRuntimeConfig *>(cfg), &w);
    if (why && len > 0) { std::strncpy(why, w.c_str(), len - 1); why[len - 1] = 0; }
    return static_cast<int>(st);
}

int ref_derive_config(const void *rt, unsigned rx_cap, unsigned slot, void *out)
{
    std::string w;
    return static_cast<int>(airan::DerivePuschMimoConfig(
        *static_cast<const airan::PuschMimoRuntimeConfig *>(rt), static_cast<uint16_t>(rx_cap),
        static_cast<airan::PuschMimoConfig *>(out), &w, static_cast<uint16_t>(slot)));
}

int ref_re_demap_gather_index(uint32_t *index, size_t n)
{
    return airan::re_demap_batch::BuildGatherIndex(index, n);
}

int ref_re_demap_reference(const uint16_t *fre, const uint16_t *fim, unsigned nr, const uint32_t *index,
                           uint16_t *rre, uint16_t *rim)
{
    return airan::re_demap_batch::Reference(fre, fim, nr, index, rre, rim);
}

int ref_re_demap_profile(const void *cfg, void *layout)
{
    airan::re_demap_batch::KernelMetadata m{};
    return airan::re_demap_batch::BuildCurrentProfile(*static_cast<const airan::PuschMimoConfig *>(cfg),
                                                      static_cast<airan::PuschMimoLayout *>(layout), &m);
}

int ref_layer_demap_reference(const int16_t *llr, const void *cfg, int16_t *cw)
{
    const auto &c = *static_cast<const airan::PuschMimoConfig *>(cfg);
    airan::PuschMimoLayout layout{};
    airan::layer_demap::KernelMetadata m{};
    uint32_t gather[airan::layer_demap::MAX_INDEX_ELEMS];
    int st = airan::layer_demap::BuildCurrentProfile(c, &layout, &m, gather);
    if (st != 0) return st;
    return airan::layer_demap::ReferenceDemap(llr, c, layout, cw);
}

int ref_dmrs_gen_profile(const void *cfg, void *layout, int32_t *cinit)
{
    airan::mimo_dmrs_gen::KernelMetadata m{};
    return airan::mimo_dmrs_gen::BuildCurrentProfile(*static_cast<const airan::PuschMimoConfig *>(cfg),
                                                     static_cast<airan::PuschMimoLayout *>(layout), &m, cinit);
}

void ref_dmrs_gold_basis(uint16_t *gmat, uint16_t *g1)
{
    airan::mimo_dmrs_gen::BuildGoldBasis(gmat, g1);
}

int ref_dmrs_ls_profile(const void *cfg, void *layout, uint16_t *count, uint16_t *sc)
{
    airan::mimo_dmrs_ls::KernelMetadata m{};
    return airan::mimo_dmrs_ls::BuildCurrentProfile(*static_cast<const airan::PuschMimoConfig *>(cfg),
                                                    static_cast<airan::PuschMimoLayout *>(layout), &m, count, sc);
}

int ref_dmrs_ls_describe(const uint16_t *count, unsigned layers, unsigned dmrs, uint32_t *models)
{
    airan::mimo_dmrs_ls::ObservationModel m[airan::mimo_dmrs_ls::MAX_LAYERS] = {};
    const int st = airan::mimo_dmrs_ls::DescribeObservationModels(count, layers, dmrs, m);
    for (unsigned l = 0; l < layers && l < airan::mimo_dmrs_ls::MAX_LAYERS; ++l) models[l] = m[l];
    return st;
}

int ref_dmrs_ls_validate_natural(const uint16_t *count, const uint16_t *sc, unsigned layers, unsigned dmrs)
{
    return airan::mimo_dmrs_ls::ValidateNaturalLmmseContract(count, sc, layers, dmrs);
}

int ref_dmrs_ls_validate_798(const uint16_t *count, unsigned layers, unsigned dmrs)
{
    return airan::mimo_dmrs_ls::ValidateLmmse798Compatibility(count, layers, dmrs);
}

unsigned long long ref_pilot_sc_hash(const uint16_t *sc, unsigned count)
{
    return airan::channel_est_lmmse::PilotScHash(sc, count);
}

}  // extern "C"
