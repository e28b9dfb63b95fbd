// host_config.cpp -- runtime-config contract, layout derivation and the
// transport-block plan (host side of the C ABI).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "common.h"

namespace {

int32_t Fail(int32_t status, const char *message, char *why, size_t len)
{
    if (why != nullptr && len > 0) {
        std::snprintf(why, len, "%s", message);
    }
    return status;
}

}  // namespace

extern "C" {

// common/pusch_mimo_runtime_config.cpp:23-170
int32_t airan_ValidateMimoRuntimeConfig(const PuschMimoRuntimeConfig *cp, char *why, size_t len)
{
    if (cp == nullptr) return Fail(AIRAN_RTCFG_NULL_ARGUMENT, "runtime config is null", why, len);
    const PuschMimoRuntimeConfig &c = *cp;
    if (c.abi_version != AIRAN_PUSCH_MIMO_RUNTIME_ABI_VERSION ||
        c.struct_size != sizeof(PuschMimoRuntimeConfig))
        return Fail(AIRAN_RTCFG_ABI_MISMATCH, "runtime config ABI version/size mismatch", why, len);
    if ((c.flags & ~AIRAN_RUNTIME_KNOWN_FLAGS) != 0)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "runtime config contains unknown flags", why, len);
    if ((c.flags & AIRAN_RUNTIME_FLAG_PROFILE_HASH) != 0 &&
        std::all_of(c.profile_sha256, c.profile_sha256 + 32, [](uint8_t v) { return v == 0; }))
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "profile hash flag set with an all-zero digest", why, len);
    if (c.architecture < AIRAN_MIMO_ARCH_FULL_DIGITAL || c.architecture > AIRAN_MIMO_ARCH_ANALOG)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "unsupported MIMO architecture", why, len);
    auto valid_tx = [](uint16_t v) { return v >= 1 && v <= AIRAN_PUSCH_MIMO_MAX_TX_ANTENNAS; };
    auto valid_rx = [](uint16_t v) { return v >= 1 && v <= AIRAN_PUSCH_MIMO_MAX_RX_ANTENNAS; };
    if (!valid_tx(c.num_tx_antennas) || !valid_tx(c.num_tx_rf_chains) ||
        !valid_rx(c.num_rx_antennas) || !valid_rx(c.num_rx_rf_chains) ||
        c.num_tx_rf_chains > c.num_tx_antennas || c.num_rx_rf_chains > c.num_rx_antennas)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid antenna/RF-chain topology", why, len);
    if (c.architecture == AIRAN_MIMO_ARCH_FULL_DIGITAL &&
        (c.num_tx_rf_chains != c.num_tx_antennas || c.num_rx_rf_chains != c.num_rx_antennas))
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE,
                    "full-digital topology requires one RF chain per antenna", why, len);
    if (c.num_layers < 1 || c.num_layers > AIRAN_PUSCH_MIMO_MAX_PUSCH_LAYERS ||
        c.num_layers > c.num_tx_rf_chains || c.num_layers > c.num_rx_rf_chains ||
        c.num_tx_ports < c.num_layers || c.num_tx_ports > AIRAN_PUSCH_MIMO_MAX_PUSCH_LAYERS ||
        c.num_tx_ports > c.num_tx_rf_chains)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid layer/port dimensions", why, len);
    if (c.max_rx_antennas < c.num_rx_antennas ||
        c.max_rx_antennas > AIRAN_PUSCH_MIMO_MAX_RX_ANTENNAS || c.max_layers < c.num_layers ||
        c.max_layers > AIRAN_PUSCH_MIMO_MAX_PUSCH_LAYERS || c.rx_bucket < c.num_rx_antennas ||
        c.rx_bucket > c.max_rx_antennas || c.layer_bucket < c.num_layers ||
        c.layer_bucket > AIRAN_PUSCH_MIMO_MAX_DETECT_LAYERS)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE,
                    "active dimensions exceed declared bucket/capacity", why, len);
    if ((c.num_rx_antennas != c.rx_bucket || c.num_layers != c.layer_bucket) &&
        (c.flags & AIRAN_RUNTIME_FLAG_ADAPTER_REQUIRED) == 0)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "capacity padding requires the adapter flag", why, len);
    if (c.qm != 2 && c.qm != 4 && c.qm != 6 && c.qm != 8)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "unsupported modulation order", why, len);
    if (c.num_symbols < 1 || c.num_symbols > 14 || c.fft_size < 128 || c.fft_size > 4096 ||
        c.num_rb < 1 || c.num_rb > 275 || c.rb_start + c.num_rb > 275 ||
        c.used_subcarriers != c.num_rb * 12 || c.used_subcarriers > c.padded_subcarriers ||
        c.padded_subcarriers > c.fft_size)
        return Fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "invalid waveform geometry", why, len);
    if (c.num_dmrs_symbols < 1 || c.num_dmrs_symbols >= c.num_symbols ||
        c.dmrs_port_count != c.num_layers ||
        c.dmrs_port_count > AIRAN_PUSCH_MIMO_MAX_DMRS_PORTS ||
        (c.dmrs_type != 1 && c.dmrs_type != 2) || (c.dmrs_length != 1 && c.dmrs_length != 2) ||
        c.num_cdm_groups_without_data < 1 || c.num_cdm_groups_without_data > 3 ||
        c.n_scid > 1 ||
        __builtin_popcount(static_cast<unsigned>(c.dmrs_symbol_mask)) != c.num_dmrs_symbols)
        return Fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "invalid DMRS dimensions", why, len);
    for (uint16_t i = 0; i < c.dmrs_port_count; ++i)
        for (uint16_t j = i + 1; j < c.dmrs_port_count; ++j)
            if (c.dmrs_ports[i] == c.dmrs_ports[j])
                return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "duplicate active DMRS port", why, len);
    for (uint16_t i = c.dmrs_port_count; i < AIRAN_PUSCH_MIMO_MAX_DMRS_PORTS; ++i)
        if (c.dmrs_ports[i] != 0)
            return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "inactive DMRS ports must be zero", why, len);
    if (c.precoding_mode > AIRAN_MIMO_PRECODING_NON_CODEBOOK ||
        (c.precoding_mode == AIRAN_MIMO_PRECODING_BYPASS && c.num_tx_ports != c.num_layers))
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid precoding configuration", why, len);
    if (c.slot_number > 1023 || c.data_scrambling_id > 1023 || c.rnti == 0 ||
        c.start_symbol > 13 || c.num_allocated_symbols == 0 ||
        c.start_symbol + c.num_allocated_symbols > 14 || c.dmrs_additional_position > 3 ||
        c.mapping_type > 1 || c.codeword_index > 1 || c.reserved16 != 0)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid dynamic PUSCH scheduler grant", why, len);
    if (!std::isfinite(c.channel_gain) || c.channel_gain <= 0.0f || c.channel_gain > 1.0f ||
        !std::isfinite(c.awgn_std_int16) || c.awgn_std_int16 < 0.0f || c.awgn_std_int16 > 200.0f)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "invalid channel gain/noise", why, len);
    const uint32_t data_re = static_cast<uint32_t>(c.num_symbols - c.num_dmrs_symbols) * c.used_subcarriers;
    if (c.data_re_per_layer != data_re || c.data_stride_per_layer < data_re ||
        c.qam_row_stride < c.used_subcarriers ||
        c.grid_re_per_port != static_cast<uint32_t>(c.num_symbols) * c.padded_subcarriers)
        return Fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "derived tensor geometry mismatch", why, len);
    for (uint32_t v : c.reserved)
        if (v != 0)
            return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE, "reserved runtime config fields must be zero", why, len);
    if (why != nullptr && len > 0) why[0] = '\0';
    return AIRAN_RTCFG_SUCCESS;
}

// common/pusch_mimo_runtime_config.cpp:172-190
int32_t airan_LoadMimoRuntimeConfig(const char *path, PuschMimoRuntimeConfig *config, char *why,
                                    size_t len)
{
    if (config == nullptr || path == nullptr)
        return Fail(AIRAN_RTCFG_NULL_ARGUMENT, "runtime config output is null", why, len);
    FILE *f = std::fopen(path, "rb");
    if (f == nullptr) return Fail(AIRAN_RTCFG_IO_ERROR, "runtime config file missing or wrong size", why, len);
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size != static_cast<long>(sizeof(*config))) {
        std::fclose(f);
        return Fail(AIRAN_RTCFG_IO_ERROR, "runtime config file missing or wrong size", why, len);
    }
    size_t got = std::fread(config, 1, sizeof(*config), f);
    std::fclose(f);
    if (got != sizeof(*config)) return Fail(AIRAN_RTCFG_IO_ERROR, "runtime config read failed", why, len);
    return airan_ValidateMimoRuntimeConfig(config, why, len);
}

// common/pusch_mimo_runtime_config.cpp:192-214
int32_t airan_LoadMimoRuntimeConfigFromEnv(uint16_t expected_rank, PuschMimoRuntimeConfig *config,
                                           int32_t *enabled, char *why, size_t len)
{
    if (config == nullptr || enabled == nullptr)
        return Fail(AIRAN_RTCFG_NULL_ARGUMENT, "runtime config/environment output is null", why, len);
    const char *path = std::getenv("PUSCH_MIMO_RUNTIME_CONFIG");
    if (path == nullptr || path[0] == '\0') {
        *enabled = 0;
        std::memset(config, 0, sizeof(*config));
        if (why != nullptr && len > 0) why[0] = '\0';
        return AIRAN_RTCFG_SUCCESS;
    }
    int32_t st = airan_LoadMimoRuntimeConfig(path, config, why, len);
    if (st != AIRAN_RTCFG_SUCCESS) return st;
    if (expected_rank != 0 && config->num_layers != expected_rank)
        return Fail(AIRAN_RTCFG_INCONSISTENT_SHAPE, "runtime config Rank does not match stage Rank", why, len);
    *enabled = 1;
    return AIRAN_RTCFG_SUCCESS;
}

// common/pusch_mimo_runtime_config.cpp:216-272
int32_t airan_DerivePuschMimoConfig(const PuschMimoRuntimeConfig *rt, uint16_t rx_capacity,
                                    PuschMimoConfig *config, char *why, size_t len,
                                    uint16_t slot_number)
{
    if (config == nullptr)
        return Fail(AIRAN_RTCFG_NULL_ARGUMENT, "derived PuschMimoConfig output is null", why, len);
    int32_t st = airan_ValidateMimoRuntimeConfig(rt, why, len);
    if (st != AIRAN_RTCFG_SUCCESS) return st;
    if (rt->dmrs_port_count > 4 || rx_capacity < rt->num_rx_antennas ||
        rx_capacity > AIRAN_PUSCH_MIMO_MAX_RX_ANTENNAS || rt->num_symbols > 255)
        return Fail(AIRAN_RTCFG_UNSUPPORTED_VALUE,
                    "runtime config cannot be represented by operator ABI v1", why, len);
    PuschMimoConfig r{};
    r.abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    r.struct_size = sizeof(r);
    r.num_layers = rt->num_layers;
    r.num_tx_ports = rt->num_tx_ports;
    r.num_rx_antennas = rx_capacity;
    r.qm = rt->qm;
    r.num_symbols = rt->num_symbols;
    r.fft_size = rt->fft_size;
    r.num_rb = rt->num_rb;
    r.rb_start = rt->rb_start;
    r.slot_number = slot_number == 0 ? rt->slot_number : slot_number;
    r.start_symbol = static_cast<uint8_t>(rt->start_symbol);
    r.num_allocated_symbols = static_cast<uint8_t>(rt->num_allocated_symbols);
    r.used_subcarriers = rt->used_subcarriers;
    r.padded_subcarriers = rt->padded_subcarriers;
    r.dmrs_symbol_mask = rt->dmrs_symbol_mask;
    for (uint16_t i = 0; i < rt->dmrs_port_count; ++i) r.dmrs_ports[i] = rt->dmrs_ports[i];
    r.dmrs_type = static_cast<uint8_t>(rt->dmrs_type);
    r.dmrs_length = static_cast<uint8_t>(rt->dmrs_length);
    r.dmrs_scrambling_id = rt->dmrs_scrambling_id;
    r.data_scrambling_id = rt->data_scrambling_id;
    r.rnti = rt->rnti;
    r.dmrs_additional_position = static_cast<uint8_t>(rt->dmrs_additional_position);
    r.mapping_type = static_cast<uint8_t>(rt->mapping_type);
    r.num_cdm_groups_without_data = static_cast<uint8_t>(rt->num_cdm_groups_without_data);
    r.n_scid = static_cast<uint8_t>(rt->n_scid);
    r.codeword_index = static_cast<uint8_t>(rt->codeword_index);
    r.codebook_enabled = rt->precoding_mode == AIRAN_MIMO_PRECODING_CODEBOOK;
    r.tpmi = rt->tpmi;
    r.prg_size_rb = rt->prg_size_rb;
    *config = r;
    if (why != nullptr && len > 0) why[0] = '\0';
    return AIRAN_RTCFG_SUCCESS;
}

int32_t airan_DerivePuschMimoLayout(const PuschMimoConfig *config, PuschMimoLayout *layout)
{
    if (config == nullptr || layout == nullptr) return AIRAN_INVALID_ARGUMENT;
    airan::Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    airan::FillLayout(g, layout);
    return AIRAN_OK;
}

// TS 38.214 Table 5.1.3.2-1 (TBS for N_info <= 3824).
static const uint32_t kTbsTable[] = {
    24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144, 152, 160, 168, 176,
    184, 192, 208, 224, 240, 256, 272, 288, 304, 320, 336, 352, 368, 384, 408, 432, 456, 480,
    504, 528, 552, 576, 608, 640, 672, 704, 736, 768, 808, 848, 888, 928, 984, 1032, 1064,
    1128, 1160, 1192, 1224, 1256, 1288, 1320, 1352, 1416, 1480, 1544, 1608, 1672, 1736, 1800,
    1864, 1928, 2024, 2088, 2152, 2216, 2280, 2408, 2472, 2536, 2600, 2664, 2728, 2792, 2856,
    2976, 3104, 3240, 3368, 3496, 3624, 3752, 3824};

// TS 38.214 5.1.3.2 steps 1-4 with N_oh = 0.
uint32_t airan_pusch_Tbs(uint32_t n_prb, uint32_t n_re_per_prb, uint32_t qm, uint32_t layers,
                         double rate)
{
    const uint32_t nre = std::min<uint32_t>(n_re_per_prb, 156) * n_prb;
    const double ninfo = static_cast<double>(nre) * rate * qm * layers;
    if (ninfo <= 0.0) return 0;
    if (ninfo <= 3824.0) {
        int n = std::max(3, static_cast<int>(std::floor(std::log2(ninfo))) - 6);
        double p = std::ldexp(1.0, n);
        double np = std::max(24.0, p * std::floor(ninfo / p));
        for (uint32_t v : kTbsTable)
            if (v >= np) return v;
        return 3824;
    }
    int n = static_cast<int>(std::floor(std::log2(ninfo - 24.0))) - 5;
    double p = std::ldexp(1.0, n);
    double np = std::max(3840.0, p * std::nearbyint((ninfo - 24.0) / p));
    double c;
    if (rate <= 0.25)
        c = std::ceil((np + 24.0) / 3816.0);
    else if (np > 8424.0)
        c = std::ceil((np + 24.0) / 8424.0);
    else
        c = 1.0;
    return static_cast<uint32_t>(8.0 * c * std::ceil((np + 24.0) / (8.0 * c)) - 24.0);
}

// TS 38.212 5.1, 5.2.2 (base graph 1, lifting set i_LS = 1) and 5.4.2.1.
int32_t airan_pusch_BuildTbPlan(uint32_t tbs, uint32_t G, uint32_t qm, uint32_t layers,
                                uint32_t rv, AiranTbPlanV1 *p)
{
    if (p == nullptr || tbs == 0 || qm == 0 || layers == 0 || rv > 3) return AIRAN_INVALID_ARGUMENT;
    std::memset(p, 0, sizeof(*p));
    p->tbs = tbs;
    p->crc_tb = tbs > 3824 ? 24 : 16;
    const uint32_t B = tbs + p->crc_tb;
    const uint32_t kcb = 8448;
    uint32_t bprime;
    if (B <= kcb) {
        p->C = 1;
        p->crc_cb = 0;
        bprime = B;
    } else {
        p->C = (B + (kcb - 24) - 1) / (kcb - 24);
        p->crc_cb = 24;
        bprime = B + 24 * p->C;
    }
    p->Kprime = bprime / p->C;
    uint32_t z = 0;
    for (uint32_t cand = 3; cand <= 384; cand *= 2)
        if (22 * cand >= p->Kprime) {
            z = cand;
            break;
        }
    if (z == 0) return AIRAN_UNSUPPORTED_PROFILE;
    p->Z = z;
    p->K = 22 * z;
    p->filler = p->K - p->Kprime;
    p->N = 66 * z;
    p->Ncb = p->N;
    p->rv = rv;
    p->G = G;
    const uint32_t q = layers * qm;
    if (G == 0 || G % q != 0) return AIRAN_INVALID_ARGUMENT;
    const uint32_t gp = G / q;
    p->n_low = p->C - gp % p->C;
    p->E_low = q * (gp / p->C);
    p->E_high = q * ((gp + p->C - 1) / p->C);
    static const uint32_t kNum[4] = {0, 17, 33, 56};
    p->k0 = static_cast<uint32_t>((static_cast<uint64_t>(kNum[rv]) * p->Ncb) / (66ull * z)) * z;
    return AIRAN_OK;
}

// fec/rate_dematch_mimo/rate_dematch_mimo.cpp:90-112: G split over C = 143 CBs
// of K = 8448 (Z = 384, no CRC, no filler), rv0.
int32_t airan_pusch_ReferenceTbPlan(const PuschMimoConfig *config, uint32_t num_slots,
                                    AiranTbPlanV1 *p)
{
    if (config == nullptr || p == nullptr || num_slots == 0) return AIRAN_INVALID_ARGUMENT;
    airan::Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    std::memset(p, 0, sizeof(*p));
    p->C = 143;
    p->Z = 384;
    p->K = p->Kprime = 22 * 384;
    p->tbs = p->C * p->K;
    p->N = p->Ncb = 66 * 384;
    p->G = num_slots * g.qm * g.cw_symbols;
    const uint32_t q = g.nl * g.qm;
    const uint32_t units = p->G / q;
    p->n_low = p->C - units % p->C;
    p->E_low = q * (units / p->C);
    p->E_high = q * ((units + p->C - 1) / p->C);
    return AIRAN_OK;
}

float airan_pusch_NoEffFloor(uint32_t qm)
{
    float norm;
    switch (qm) {
    case 2: norm = 2.0f; break;
    case 4: norm = 10.0f; break;
    case 6: norm = 42.0f; break;
    case 8: norm = 170.0f; break;
    default: return 0.0f;
    }
    return 0.04f * static_cast<float>((1u << (qm / 2)) - 1u) / norm;
}

}  // extern "C"
