// txgen.cpp -- synthetic PUSCH transmitter (host): the inverse of the
// receive chain's bit-level stages, used to build loopback slots.
//
// Conventions are the receive side's, stage by stage:
//   CRC            TS 38.212 5.1 (CRC24A / CRC24B / CRC16, MSB first)
//   segmentation   TS 38.212 5.2.2 (K' bits per block incl. CRC24B, filler
//                  K - K' zero bits at the end of the systematic part)
//   LDPC           base graph 1 (include/airan_ldpc_bg1.h), circulant
//                  P^s x [i] = x[(i + s) mod Z]
//   rate matching  TS 38.212 5.4.2.1 bit selection from k0 over the
//                  circular buffer d = codeword[2Z .. 2Z + Ncb), filler
//                  positions skipped; interleave f[j Qm + i] = e[i E/Qm + j]
//   scrambling     c_init = rnti 2^15 + q 2^14 + n_ID, Gold index from 0
//   modulation     the reference mapper's polarity, which is TS 38.211 5.1
//                  Gray mapping with every bit inverted (positive LLR <=> 0
//                  through the whole receive chain), unit average power
//   layer mapping  codeword symbol i -> layer i % L, layer symbol i / L
//   DMRS           type 1, r(m) = ((1-2c(2m)) + j(1-2c(2m+1)))/sqrt(2) on
//                  subcarrier comb + 2m, times (-1)^m on ports 1001/1003.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "airan_ldpc_bg1.h"
#include "airan_pusch_tx.h"
#include "common.h"

namespace {

class GoldHost {
public:
    GoldHost(uint32_t cinit)
    {
        s1_ = 1u;
        s2_ = cinit & 0x7fffffffu;
        for (int i = 0; i < 1600; ++i) Step();
    }
    uint32_t Next()
    {
        const uint32_t c = (s1_ ^ s2_) & 1u;
        Step();
        return c;
    }

private:
    void Step()
    {
        const uint32_t n1 = ((s1_ >> 3) ^ s1_) & 1u;
        const uint32_t n2 = ((s2_ >> 3) ^ (s2_ >> 2) ^ (s2_ >> 1) ^ s2_) & 1u;
        s1_ = (s1_ >> 1) | (n1 << 30);
        s2_ = (s2_ >> 1) | (n2 << 30);
    }
    uint32_t s1_, s2_;
};

uint32_t CrcBits(uint32_t kind, const uint8_t *bits, uint64_t n)
{
    uint32_t poly, len;
    switch (kind) {
    case 0: poly = 0x864CFBu; len = 24; break;
    case 1: poly = 0x800063u; len = 24; break;
    default: poly = 0x1021u; len = 16; break;
    }
    const uint32_t mask = (1u << len) - 1u;
    uint32_t reg = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t fb = ((reg >> (len - 1)) & 1u) ^ (bits[i] & 1u);
        reg = (reg << 1) & mask;
        if (fb) reg ^= poly;
    }
    return reg;
}

void ShiftXor(uint8_t *dst, const uint8_t *src, uint32_t s, uint32_t z)
{
    for (uint32_t i = 0; i < z; ++i) dst[i] ^= src[(i + s) % z];
}

int32_t Bg1Encode(uint32_t z, const uint8_t *info, uint8_t *cw)
{
    std::memset(cw, 0, 68u * z);
    std::memcpy(cw, info, 22u * z);
    // Core rows 0..3: lam_r = sum over systematic columns.
    std::vector<uint8_t> lam(4u * z, 0);
    for (int r = 0; r < 4; ++r)
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            if (AIRAN_BG1_COL[e] < 22) ShiftXor(&lam[r * z], info + AIRAN_BG1_COL[e] * z, AIRAN_BG1_SHIFT[e] % z, z);
    // Summing the four core rows cancels columns 23..25 (each appears twice
    // with shift 0) and every pair of equal shifts on column 22, leaving
    // P^b p0 = sum_r lam_r.
    std::vector<uint32_t> s22;
    for (int r = 0; r < 4; ++r)
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            if (AIRAN_BG1_COL[e] == 22) {
                const uint32_t s = AIRAN_BG1_SHIFT[e] % z;
                auto it = std::find(s22.begin(), s22.end(), s);
                if (it != s22.end()) s22.erase(it);
                else s22.push_back(s);
            }
    if (s22.size() != 1) return AIRAN_UNSUPPORTED_PROFILE;
    uint8_t *p0 = cw + 22u * z;
    for (uint32_t i = 0; i < z; ++i) p0[(i + s22[0]) % z] = lam[i] ^ lam[z + i] ^ lam[2 * z + i] ^ lam[3 * z + i];
    // Remaining core parities: solve any core row with one unknown column
    // (shift 0) until all of 23..25 are known.
    bool known[26] = {};
    for (int c = 0; c <= 22; ++c) known[c] = true;
    for (int pass = 0; pass < 4; ++pass)
        for (int r = 0; r < 4; ++r) {
            int unknown = -1, count = 0;
            for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e) {
                const int c = AIRAN_BG1_COL[e];
                if (c >= 22 && c < 26 && !known[c]) {
                    unknown = e;
                    ++count;
                }
            }
            if (count != 1) continue;
            if (AIRAN_BG1_SHIFT[unknown] % z != 0) return AIRAN_UNSUPPORTED_PROFILE;
            uint8_t *pu = cw + AIRAN_BG1_COL[unknown] * z;
            std::memcpy(pu, &lam[r * z], z);
            for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e) {
                const int c = AIRAN_BG1_COL[e];
                if (c >= 22 && c < 26 && e != unknown) ShiftXor(pu, cw + c * z, AIRAN_BG1_SHIFT[e] % z, z);
            }
            known[AIRAN_BG1_COL[unknown]] = true;
        }
    if (!known[23] || !known[24] || !known[25]) return AIRAN_UNSUPPORTED_PROFILE;
    // Extension rows: one degree-1 parity column each.
    for (int r = 4; r < AIRAN_BG1_ROWS; ++r) {
        uint8_t *pc = cw + (22u + r) * z;
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            if (AIRAN_BG1_COL[e] != 22 + r) ShiftXor(pc, cw + AIRAN_BG1_COL[e] * z, AIRAN_BG1_SHIFT[e] % z, z);
    }
    // Every check must hold.
    std::vector<uint8_t> syn(z);
    for (int r = 0; r < AIRAN_BG1_ROWS; ++r) {
        std::fill(syn.begin(), syn.end(), 0);
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            ShiftXor(syn.data(), cw + AIRAN_BG1_COL[e] * z, AIRAN_BG1_SHIFT[e] % z, z);
        for (uint32_t i = 0; i < z; ++i)
            if (syn[i]) return AIRAN_LAUNCH_FAILED;
    }
    return AIRAN_OK;
}

// Axis level d0 (2^(m-1) - d1 (2^(m-2) - ... d_{m-1})), d_i = 2 c_i - 1
// (mapping/qam_mod_256_siso/scripts/qam256_mod_ref.py:38-45 for m = 4; the
// same recursion for QPSK/16QAM/64QAM).
float AxisLevel(const uint8_t *axis_bits, uint32_t m)
{
    float e = 1.0f;
    for (int i = static_cast<int>(m) - 1; i >= 1; --i)
        e = static_cast<float>(1u << (m - i)) - (2.0f * axis_bits[i] - 1.0f) * e;
    return (2.0f * axis_bits[0] - 1.0f) * e;
}

float QamNorm(uint32_t qm)
{
    switch (qm) {
    case 2: return 1.0f / std::sqrt(2.0f);
    case 4: return 1.0f / std::sqrt(10.0f);
    case 6: return 1.0f / std::sqrt(42.0f);
    default: return 1.0f / std::sqrt(170.0f);
    }
}

bool PlanOk(const AiranTbPlanV1 &p, const airan::Geom &g, uint32_t slots_per_tb)
{
    if (p.C == 0 || p.Z == 0 || p.K != 22 * p.Z || p.Kprime > p.K || p.filler != p.K - p.Kprime ||
        p.Ncb == 0 || p.Ncb > 66 * p.Z || p.k0 >= p.Ncb || p.E_low % g.qm != 0 || p.E_high % g.qm != 0 ||
        p.n_low > p.C)
        return false;
    if (static_cast<uint64_t>(p.G) != static_cast<uint64_t>(slots_per_tb) * g.cw_symbols * g.qm) return false;
    const uint64_t e_sum = static_cast<uint64_t>(p.n_low) * p.E_low + static_cast<uint64_t>(p.C - p.n_low) * p.E_high;
    if (e_sum != p.G) return false;
    if (p.tbs + p.crc_tb + static_cast<uint64_t>(p.C) * p.crc_cb != static_cast<uint64_t>(p.C) * p.Kprime)
        return false;
    if (p.crc_tb != 0 && p.crc_tb != 16 && p.crc_tb != 24) return false;
    if (p.crc_cb != 0 && p.crc_cb != 24) return false;
    // Filler positions must not cover the whole buffer.
    return p.Ncb > p.filler;
}

// slot_numbers: the slot number of each of the TB's slots (its DMRS), or
// null for cfg.slot_number.
int32_t EncodeOneTb(const PuschMimoConfig &cfg, const airan::Geom &g, const AiranTbPlanV1 &p,
                    uint32_t slots_per_tb, const uint16_t *slot_numbers, const uint8_t *payload, float *grid,
                    uint8_t *coded_out)
{
    const uint32_t Z = p.Z;
    // TB with CRC.
    std::vector<uint8_t> b(p.tbs + p.crc_tb);
    std::memcpy(b.data(), payload, p.tbs);
    if (p.crc_tb) {
        const uint32_t rem = CrcBits(p.crc_tb == 24 ? 0 : 2, payload, p.tbs);
        for (uint32_t i = 0; i < p.crc_tb; ++i) b[p.tbs + i] = (rem >> (p.crc_tb - 1 - i)) & 1u;
    }
    // Segmentation, encoding, rate matching, interleaving.
    std::vector<uint8_t> coded(p.G);
    std::vector<uint8_t> info(22u * Z), cw(68u * Z), e_bits;
    const uint32_t per_cb = p.Kprime - p.crc_cb;
    const uint32_t fstart = p.Kprime - 2 * Z;  // filler start in circular-buffer coordinates
    uint64_t out_pos = 0;
    for (uint32_t r = 0; r < p.C; ++r) {
        std::fill(info.begin(), info.end(), 0);
        std::memcpy(info.data(), b.data() + static_cast<size_t>(r) * per_cb, per_cb);
        if (p.crc_cb) {
            const uint32_t rem = CrcBits(1, info.data(), per_cb);
            for (uint32_t i = 0; i < 24; ++i) info[per_cb + i] = (rem >> (23 - i)) & 1u;
        }
        const int32_t st = Bg1Encode(Z, info.data(), cw.data());
        if (st != AIRAN_OK) return st;
        const uint32_t E = r < p.n_low ? p.E_low : p.E_high;
        e_bits.resize(E);
        const uint8_t *d = cw.data() + 2 * Z;
        uint32_t k = 0;
        for (uint64_t j = 0; k < E; ++j) {
            const uint32_t idx = static_cast<uint32_t>((p.k0 + j) % p.Ncb);
            if (idx >= fstart && idx < fstart + p.filler) continue;
            e_bits[k++] = d[idx];
        }
        const uint32_t eq = E / g.qm;
        for (uint32_t j = 0; j < eq; ++j)
            for (uint32_t i = 0; i < g.qm; ++i) coded[out_pos + static_cast<uint64_t>(j) * g.qm + i] = e_bits[i * eq + j];
        out_pos += E;
    }
    // Scrambling.
    const uint32_t cinit = (static_cast<uint32_t>(cfg.rnti) << 15) |
                           ((static_cast<uint32_t>(cfg.codeword_index) & 1u) << 14) | cfg.data_scrambling_id;
    GoldHost gold(cinit);
    for (uint64_t n = 0; n < p.G; ++n) coded[n] ^= static_cast<uint8_t>(gold.Next());
    if (coded_out) std::memcpy(coded_out, coded.data(), p.G);
    // Modulation + layer mapping + grid.
    const uint32_t m = g.qm / 2;
    const float norm = QamNorm(g.qm);
    const size_t sym_plane = g.used;
    const size_t layer_plane = static_cast<size_t>(14) * sym_plane;
    const size_t slot_plane = static_cast<size_t>(g.nl) * layer_plane;
    std::memset(grid, 0, sizeof(float) * 2 * slot_plane * slots_per_tb);
    uint8_t ib[4], qb[4];
    for (uint32_t s = 0; s < slots_per_tb; ++s) {
        float *sg = grid + 2 * slot_plane * s;
        for (uint32_t i = 0; i < g.cw_symbols; ++i) {
            const uint8_t *bits = coded.data() + (static_cast<uint64_t>(s) * g.cw_symbols + i) * g.qm;
            for (uint32_t a = 0; a < m; ++a) {
                ib[a] = bits[2 * a];
                qb[a] = bits[2 * a + 1];
            }
            const uint32_t l = i % g.nl, re = i / g.nl;
            const uint32_t ds = re / g.used, sc = re - ds * g.used;
            float *o = sg + 2 * (l * layer_plane + g.data_sym[ds] * sym_plane + sc);
            o[0] = AxisLevel(ib, m) * norm;
            o[1] = AxisLevel(qb, m) * norm;
        }
        // DMRS of this slot's number (mimo_dmrs_gen.cpp:38-45).
        const float h = 0.70710678118654752f;
        const uint64_t slot_number = slot_numbers != nullptr ? slot_numbers[s] : cfg.slot_number;
        for (uint32_t d = 0; d < g.ndmrs; ++d) {
            const uint64_t term = (1ull << 17) * (14ull * slot_number + g.dmrs_sym[d] + 1ull) *
                                  (2ull * cfg.dmrs_scrambling_id + 1ull);
            const uint32_t c_dmrs =
                static_cast<uint32_t>((term + 2ull * cfg.dmrs_scrambling_id + cfg.n_scid) & 0x7fffffffu);
            GoldHost dg(c_dmrs);
            for (uint32_t skip = 0; skip < 12u * cfg.rb_start; ++skip) dg.Next();
            std::vector<uint8_t> c(2 * g.n_dmrs_re);
            for (auto &v : c) v = static_cast<uint8_t>(dg.Next());
            for (uint32_t l = 0; l < g.nl; ++l) {
                const bool wf_odd = cfg.dmrs_ports[l] == 1001 || cfg.dmrs_ports[l] == 1003;
                for (uint32_t mm = 0; mm < g.n_dmrs_re; ++mm) {
                    float vr = c[2 * mm] ? -h : h, vi = c[2 * mm + 1] ? -h : h;
                    if (wf_odd && (mm & 1u)) {
                        vr = -vr;
                        vi = -vi;
                    }
                    float *o = sg + 2 * (l * layer_plane + g.dmrs_sym[d] * sym_plane + g.comb[l] + 2 * mm);
                    o[0] = vr;
                    o[1] = vi;
                }
            }
        }
    }
    return AIRAN_OK;
}

}  // namespace

extern "C" {

uint32_t airan_crc_Compute(uint32_t kind, const uint8_t *bits, uint64_t nbits)
{
    if (bits == nullptr && nbits != 0) return 0;
    return CrcBits(kind, bits, nbits);
}

int32_t airan_ldpc_bg1_Encode(uint32_t z, const uint8_t *info, uint8_t *codeword)
{
    if (info == nullptr || codeword == nullptr || z == 0 || z > 384) return AIRAN_INVALID_ARGUMENT;
    return Bg1Encode(z, info, codeword);
}

int32_t airan_pusch_tx_EncodeTbs(const PuschMimoConfig *config, const AiranTbPlanV1 *plan, uint32_t num_tbs,
                                 uint32_t slots_per_tb, const uint8_t *payload, float *grid, uint8_t *coded_bits)
{
    return airan_pusch_tx_EncodeTbsSlots(config, plan, num_tbs, slots_per_tb, nullptr, payload, grid, coded_bits);
}

int32_t airan_pusch_tx_EncodeTbsSlots(const PuschMimoConfig *config, const AiranTbPlanV1 *plan, uint32_t num_tbs,
                                      uint32_t slots_per_tb, const uint16_t *slot_numbers, const uint8_t *payload,
                                      float *grid, uint8_t *coded_bits)
{
    if (config == nullptr || plan == nullptr || payload == nullptr || grid == nullptr || num_tbs == 0 ||
        slots_per_tb == 0)
        return AIRAN_INVALID_ARGUMENT;
    airan::Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (!PlanOk(*plan, g, slots_per_tb)) return AIRAN_INVALID_ARGUMENT;
    const size_t slot_floats = static_cast<size_t>(2) * g.nl * 14 * g.used;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nthreads = std::min<unsigned>(hw, num_tbs);
    std::vector<int32_t> status(num_tbs, AIRAN_OK);
    auto work = [&](unsigned t) {
        for (uint32_t tb = t; tb < num_tbs; tb += nthreads)
            status[tb] = EncodeOneTb(*config, g, *plan, slots_per_tb,
                                     slot_numbers ? slot_numbers + static_cast<size_t>(tb) * slots_per_tb : nullptr,
                                     payload + static_cast<size_t>(tb) * plan->tbs,
                                     grid + slot_floats * slots_per_tb * tb,
                                     coded_bits ? coded_bits + static_cast<size_t>(tb) * plan->G : nullptr);
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    for (int32_t s : status)
        if (s != AIRAN_OK) return s;
    return AIRAN_OK;
}

}  // extern "C"
