// abi_caller.cpp -- a C++ caller of libpuschrx_hip.so compiled against
// include/*.h (the way a host chain would bind it), exercising the host-side
// contract functions without a GPU: runtime config load/validate/derive,
// every operator's BuildCurrentProfile, the TB plan, the host references and
// the DMRS-LS validators.  Prints one "key value" line per result for
// tests/test_abi.py to compare with its own reading of the same library.
//
// usage: abi_caller <192-byte runtime config file> <rx capacity>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "airan_ce_refiner.h"
#include "airan_pusch_chain.h"
#include "airan_pusch_rx.h"
#include "airan_pusch_tx.h"

namespace {

int g_fail = 0;

void Expect(const char *what, int32_t got, int32_t want)
{
    std::printf("%s %d\n", what, got);
    if (got != want) {
        std::printf("FAIL %s: %d != %d\n", what, got, want);
        g_fail = 1;
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    PuschMimoRuntimeConfig rt{};
    char why[256] = {};
    Expect("load", airan_LoadMimoRuntimeConfig(argv[1], &rt, why, sizeof(why)), AIRAN_RTCFG_SUCCESS);
    Expect("validate", airan_ValidateMimoRuntimeConfig(&rt, why, sizeof(why)), AIRAN_RTCFG_SUCCESS);
    PuschMimoConfig cfg{};
    const uint16_t cap = static_cast<uint16_t>(std::atoi(argv[2]));
    Expect("derive", airan_DerivePuschMimoConfig(&rt, cap, &cfg, why, sizeof(why), 0), AIRAN_RTCFG_SUCCESS);
    PuschMimoLayout lay{};
    Expect("layout", airan_DerivePuschMimoLayout(&cfg, &lay), AIRAN_OK);
    std::printf("data_stride %u\ncodeword_symbols %u\ncodeword_stride %u\n", lay.data_stride, lay.codeword_symbols,
                lay.codeword_stride);

    // A broken copy of the runtime config is rejected with a reason.
    PuschMimoRuntimeConfig bad = rt;
    bad.rnti = 0;
    const int32_t bad_st = airan_ValidateMimoRuntimeConfig(&bad, why, sizeof(why));
    std::printf("bad_rnti %d %s\n", bad_st, why);
    if (bad_st == AIRAN_RTCFG_SUCCESS) g_fail = 1;

    // Operator profiles.
    PuschMimoLayout l2{};
    ReDemapKernelMetadata rd{};
    Expect("re_demap_profile", airan_re_demap_batch_BuildCurrentProfile(&cfg, &l2, &rd), AIRAN_OK);
    MimoDmrsGenKernelMetadata dg{};
    int32_t cinit[16] = {};
    Expect("dmrs_gen_profile", airan_mimo_dmrs_gen_BuildCurrentProfile(&cfg, &l2, &dg, cinit), AIRAN_OK);
    std::printf("dmrs_cinit0 %d\n", cinit[0]);
    MimoDmrsLsKernelMetadata dl{};
    uint16_t count[16] = {};
    std::vector<uint16_t> sc(4u * 4u * 2048u);
    Expect("dmrs_ls_profile", airan_mimo_dmrs_ls_BuildCurrentProfile(&cfg, &l2, &dl, count, sc.data()), AIRAN_OK);
    const uint32_t nd = dl.num_dmrs_symbols;
    uint32_t models[4] = {};
    Expect("dmrs_ls_describe",
           airan_mimo_dmrs_ls_DescribeObservationModels(&cfg, count, cfg.num_layers, nd, models), AIRAN_OK);
    Expect("dmrs_ls_natural",
           airan_mimo_dmrs_ls_ValidateNaturalLmmseContract(&cfg, count, sc.data(), cfg.num_layers, nd), AIRAN_OK);
    QamDemodBatchTilingData qt{};
    Expect("qam_profile", airan_qam256_demod_batch_BuildCurrentProfile(&cfg, &l2, &qt), AIRAN_OK);
    LayerDemapKernelMetadata ld{};
    Expect("layer_profile", airan_layer_demap_BuildCurrentProfile(&cfg, &l2, &ld), AIRAN_OK);
    DescrambleKernelMetadata dd{};
    Expect("descramble_profile", airan_descramble_mimo_BuildCurrentProfile(&cfg, 1, &l2, &dd), AIRAN_OK);

    // TB plan of one slot at R = 0.85 (or the reference's fixed plan).
    AiranTbPlanV1 plan{};
    const uint32_t tbs =
        airan_pusch_Tbs(cfg.num_rb, 12u * (14u - nd), cfg.qm, cfg.num_layers, 873.0 / 1024.0);
    Expect("tb_plan", airan_pusch_BuildTbPlan(tbs, lay.codeword_symbols * cfg.qm, cfg.qm, cfg.num_layers, 0, &plan),
           AIRAN_OK);
    std::printf("tbs %u\nC %u\nZ %u\nE_low %u\nE_high %u\nactive_rows %u\n", plan.tbs, plan.C, plan.Z, plan.E_low,
                plan.E_high, airan_ldpc_ActiveRows(&plan));
    std::vector<RateMatchDescriptor> desc(plan.C);
    RateDematchKernelMetadata rm{};
    Expect("rate_profile",
           airan_rate_dematch_mimo_BuildProfileForTb(&cfg, 1, &plan, &l2, &rm, desc.data(), desc.size()), AIRAN_OK);

    // Host references on a deterministic pattern.
    std::vector<int16_t> llr(static_cast<size_t>(cfg.num_layers) * cfg.qm * (14u - nd) * ((cfg.used_subcarriers + 63u) / 64u * 64u));
    for (size_t i = 0; i < llr.size(); ++i) llr[i] = static_cast<int16_t>((i * 2654435761u >> 7) % 5121) - 2560;
    std::vector<int16_t> cw(static_cast<size_t>(cfg.qm) * lay.codeword_stride);
    std::vector<int16_t> nr(cw.size());
    Expect("layer_reference", airan_layer_demap_ReferenceDemap(llr.data(), &cfg, &lay, cw.data()), AIRAN_OK);
    Expect("descramble_reference",
           airan_descramble_mimo_ReferenceDescramble(cw.data(), nullptr, &cfg, &lay, 1, nr.data()), AIRAN_OK);
    std::vector<int16_t> ldpc(static_cast<size_t>(plan.C) * 68u * plan.Z);
    Expect("rate_reference",
           airan_rate_dematch_mimo_ReferenceRateDematchForTb(nr.data(), &cfg, &lay, 1, &plan, desc.data(),
                                                             desc.size(), ldpc.data()),
           AIRAN_OK);
    long long sum = 0;
    for (int16_t v : ldpc) sum += v;
    std::printf("ldpc_llr_sum %lld\n", sum);

    // Argument checks fail closed.
    Expect("null_layout", airan_layer_demap_ReferenceDemap(llr.data(), &cfg, nullptr, cw.data()),
           AIRAN_INVALID_ARGUMENT);
    LdpcDecodeOpArgsV1 la{};
    Expect("ldpc_bad_header", airan_ldpc_decode_ValidateOpArgs(&la), AIRAN_INVALID_ARGUMENT);
    Expect("chain_bad_create", airan_pusch_chain_Create(nullptr, nullptr, nullptr, nullptr, nullptr),
           AIRAN_INVALID_ARGUMENT);
    std::printf("result %s\n", g_fail ? "FAIL" : "OK");
    return g_fail;
}
