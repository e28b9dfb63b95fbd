// grants.hip -- ragged multi-grant batches (include/airan_pusch_grants.h).
//
// N grants of different shapes, each its own cell and slot with one TB, run
// through the receive chain with one launch per stage on work lists built
// at Create (ragged.h).  The per-grant arithmetic is that of the
// single-grant chain: the same device functions, only the indexing differs.
// The CE filter of each distinct grant shape is built once per process and
// reused (keyed by the shape and the CE prior); grants of one shape are
// laid out next to each other in the channel buffers so that one CE job
// serves them, and the jobs of all shapes run in one launch per CE step.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "airan_pusch_grants.h"
#include "common.h"
#include "crc.h"
#include "detect.h"
#include "detect_demap.h"
#include "gold.h"
#include "launch.h"
#include "ragged.h"

using airan::Cf;
using airan::DevGeom;
using airan::Geom;
using airan::RaggedCb;
using airan::RaggedGrant;
using airan::RaggedItem;
using airan::RaggedStream;
using airan::RaggedTile;

namespace {

// Gold bits of one grant's data scrambling sequence: word w = c(32 w ..).
__global__ void __launch_bounds__(256) GrantGoldKernel(uint32_t *__restrict__ bits, uint32_t words, uint32_t cinit)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t s1, s2;
    airan::GoldStateAt(32ull * w + 1600ull, cinit, &s1, &s2);
    bits[w] = airan::GoldNext32(&s1, &s2);
}

// Detection + QAM demap + layer demap + descramble of one RE of one grant:
// the single-grant chain's fused detector (per-RE LMMSE on the vector
// units, no_eff floor, the qam256_demod arithmetic, Gold descrambling).
template <int NL>
__device__ inline void RaggedDetectRe(const RaggedGrant &gr, uint32_t ds, uint32_t sc, const __half *__restrict__ rx_re,
                                      const __half *__restrict__ rx_im, const __half *__restrict__ h_re,
                                      const __half *__restrict__ h_im, const __half *__restrict__ noise,
                                      const uint32_t *__restrict__ gold, int16_t *__restrict__ cw)
{
    const DevGeom &g = gr.g;
    const uint32_t nr = g.nr;
    float s = 0.0f;
    for (uint32_t r = 0; r < nr; ++r) s += __half2float(noise[gr.noise_off + r]);
    const float no = __half2float(__float2half_rn(s * (1.0f / static_cast<float>(nr))));
    const size_t plane = static_cast<size_t>(g.nsym) * g.pad;
    const size_t off = static_cast<size_t>(g.data_sym[ds]) * g.pad + sc;
    auto loadh = [&](int r, int l) {
        const size_t i = gr.h_off + (static_cast<size_t>(r) * NL + l) * plane + off;
        return Cf{__half2float(h_re[i]), __half2float(h_im[i])};
    };
    auto loady = [&](int r) {
        const size_t i = gr.rx_off + static_cast<size_t>(r) * plane + off;
        return Cf{__half2float(rx_re[i]), __half2float(rx_im[i])};
    };
    Cf xh[4];
    float ne[4];
    airan::Detect<NL>(static_cast<int>(nr), no, loadh, loady, xh, ne);
    const uint32_t re = ds * g.used + sc;
    const uint64_t bit0 = static_cast<uint64_t>(re) * NL * g.qm;
    const uint32_t *gb = gold + gr.gold_off;
    const uint32_t w0 = static_cast<uint32_t>(bit0 >> 5);
    const uint64_t gw = (static_cast<uint64_t>(gb[w0]) | (static_cast<uint64_t>(gb[w0 + 1]) << 32)) >> (bit0 & 31u);
    airan::DemapStore<NL, 0>(xh, ne, g.qm, __ushort_as_half(static_cast<uint16_t>(gr.floor16)), gw,
                             cw + gr.cw_off + static_cast<size_t>(re) * NL, g.cw_stride);
}

// One workgroup per tile (grant, data symbol, blockDim.x subcarriers).
__global__ void __launch_bounds__(256) RaggedDetectKernel(const __half *__restrict__ rx_re,
                                                          const __half *__restrict__ rx_im,
                                                          const __half *__restrict__ h_re,
                                                          const __half *__restrict__ h_im,
                                                          const __half *__restrict__ noise,
                                                          const uint32_t *__restrict__ gold,
                                                          int16_t *__restrict__ cw,
                                                          const RaggedGrant *__restrict__ grants,
                                                          const RaggedTile *__restrict__ tiles)
{
    const RaggedTile t = tiles[blockIdx.x];
    const RaggedGrant &gr = grants[t.grant];
    const uint32_t sc = t.sc0 + threadIdx.x;
    if (sc >= gr.g.used) return;
    switch (gr.g.nl) {
    case 1: RaggedDetectRe<1>(gr, t.ds, sc, rx_re, rx_im, h_re, h_im, noise, gold, cw); break;
    case 2: RaggedDetectRe<2>(gr, t.ds, sc, rx_re, rx_im, h_re, h_im, noise, gold, cw); break;
    case 3: RaggedDetectRe<3>(gr, t.ds, sc, rx_re, rx_im, h_re, h_im, noise, gold, cw); break;
    default: RaggedDetectRe<4>(gr, t.ds, sc, rx_re, rx_im, h_re, h_im, noise, gold, cw); break;
    }
}

// Host fp32 -> fp16 bits (normal range).
uint16_t HostHalf(float f) { return __half_as_ushort(__float2half_rn(f)); }

// ---------------------------------------------------------- filter cache
// Key: the geometry the filter depends on (CeFilterMatches) plus the prior.
struct ShapeKey {
    int device;
    uint32_t nr, nl, used, pad, ndmrs, max_rank;
    uint32_t dmrs_sym[4], comb[4], occ[4];
    double tau, fd, scs;
    bool operator<(const ShapeKey &o) const { return std::memcmp(this, &o, sizeof(ShapeKey)) < 0; }
};

struct CacheEntry {
    LmmseFilter *filter = nullptr;
    uint32_t users = 0;
};

std::mutex g_cache_mutex;
std::map<ShapeKey, CacheEntry> g_cache;

ShapeKey KeyOf(int device, const Geom &g, const AiranGrantBatchOptionsV1 &o)
{
    ShapeKey k;
    std::memset(&k, 0, sizeof(k));  // padding too: the key is compared bytewise
    k.device = device;
    k.nr = g.nr;
    k.nl = g.nl;
    k.used = g.used;
    k.pad = g.pad;
    k.ndmrs = g.ndmrs;
    k.max_rank = o.ce_max_rank;
    for (int i = 0; i < 4; ++i) {
        k.dmrs_sym[i] = g.dmrs_sym[i];
        k.comb[i] = g.comb[i];
        k.occ[i] = g.occ[i];
    }
    k.tau = o.tau_s;
    k.fd = o.fd_hz;
    k.scs = o.scs_hz;
    return k;
}

template <typename T>
bool DevAlloc(T **p, size_t count, size_t *total)
{
    *p = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    if (hipMalloc(reinterpret_cast<void **>(p), bytes) != hipSuccess) return false;
    *total += bytes;
    return true;
}

template <typename T>
bool Upload(T **p, const std::vector<T> &v, size_t *total)
{
    if (!DevAlloc(p, v.size(), total)) return false;
    return v.empty() || hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

}  // namespace

struct AiranGrantBatch {
    int device = 0;
    uint32_t n = 0;
    AiranGrantBatchOptionsV1 opts{};
    std::vector<AiranGrantLayoutV1> layout;
    size_t bytes = 0;
    uint64_t iq_total = 0, num_cb = 0;
    // Work lists (device).
    RaggedGrant *grants = nullptr;
    struct FftGroup {
        uint32_t fft = 0, count = 0;
        RaggedStream *streams = nullptr;
    };
    std::vector<FftGroup> ffts;
    RaggedItem *gen_items = nullptr, *ls_items = nullptr, *noise_items = nullptr;
    uint32_t n_gen = 0, n_ls = 0, n_noise = 0, max_ref_pad = 0, max_pilot_pad = 0, max_dmrs = 0;
    RaggedTile *tiles = nullptr;
    uint32_t n_tiles = 0, tile = 256;
    std::vector<RaggedTile> host_tiles_src;  // (grant, ds, used) per row, for rebuilding tiles
    RaggedCb *cbs = nullptr;
    uint32_t *reg_list = nullptr, *ws_list = nullptr, *tb_first = nullptr, *tb_count = nullptr;
    uint32_t n_reg = 0, n_ws = 0;
    // CE groups: grants of one shape, adjacent in the channel buffers.
    struct CeGroup {
        ShapeKey key;
        LmmseFilter *filter = nullptr;
        uint32_t count = 0, nl = 0;
        uint64_t hls_off = 0, noise_off = 0, h_off = 0;
    };
    std::vector<CeGroup> groups;
    // The CE groups run together, one launch per CE step over all of them
    // (airan::LaunchCeGrouped), each group with its own workspace.
    airan::CeGrouped *ce_grouped = nullptr;
    uint8_t *ce_ws_all = nullptr;
    std::vector<size_t> ce_ws_off;
    // Buffers.
    __half *rx_re = nullptr, *rx_im = nullptr, *ref_re = nullptr, *ref_im = nullptr;
    __half *hls_re = nullptr, *hls_im = nullptr, *noise_ls = nullptr, *noise_ce = nullptr;
    __half *h_re = nullptr, *h_im = nullptr;
    size_t ce_ws_bytes = 0;
    uint32_t *gold = nullptr;
    int16_t *cw = nullptr, *lam = nullptr;
    uint2 *ldpc_ws = nullptr;
    int8_t *bits = nullptr;
    uint8_t *iters = nullptr, *cb_ok = nullptr, *tb_ok = nullptr;
    uint32_t *tb_part = nullptr;
    // Graph of the last (iq, stream).
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *graph_iq = nullptr;
    void *graph_stream = nullptr;
};

namespace {

void DropGraph(AiranGrantBatch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_iq = nullptr;
    b->graph_stream = nullptr;
}

int32_t BuildTiles(AiranGrantBatch *b, uint32_t tile)
{
    std::vector<RaggedTile> t;
    for (const RaggedTile &row : b->host_tiles_src)  // sc0 holds the grant's used subcarriers
        for (uint32_t sc0 = 0; sc0 < row.sc0; sc0 += tile) t.push_back({row.grant, row.ds, sc0});
    if (b->tiles) (void)hipFree(b->tiles);
    b->tiles = nullptr;
    size_t dummy = 0;
    if (!Upload(&b->tiles, t, &dummy)) return AIRAN_RESOURCE_TOO_SMALL;
    b->n_tiles = static_cast<uint32_t>(t.size());
    b->tile = tile;
    return AIRAN_OK;
}

int32_t EnqueueStages(AiranGrantBatch *b, const void *iq, hipStream_t s)
{
    int32_t st;
    for (const auto &f : b->ffts)
        if ((st = airan::LaunchOfdmRagged(f.fft, b->opts.iq_format, f.streams, f.count, iq, b->rx_re, b->rx_im, s)) !=
            AIRAN_OK)
            return st;
    if ((st = airan::LaunchDmrsRagged(b->grants, b->gen_items, b->n_gen, b->ls_items, b->n_ls, b->noise_items,
                                      b->n_noise, b->max_ref_pad, b->max_pilot_pad, b->max_dmrs, b->rx_re, b->rx_im,
                                      b->ref_re, b->ref_im, b->hls_re, b->hls_im, b->noise_ls, s)) != AIRAN_OK)
        return st;
    if ((st = airan::LaunchCeGrouped(b->ce_grouped, s)) != AIRAN_OK) return st;
    hipLaunchKernelGGL(RaggedDetectKernel, dim3(b->n_tiles), dim3(b->tile), 0, s, b->rx_re, b->rx_im, b->h_re,
                       b->h_im, b->noise_ce, b->gold, b->cw, b->grants, b->tiles);
    if ((st = airan::LaunchStatus()) != AIRAN_OK) return st;
    const uint32_t ncb = static_cast<uint32_t>(b->num_cb);
    if ((st = airan::LaunchRateDematchRagged(b->grants, b->cbs, ncb, b->cw, b->lam, s)) != AIRAN_OK) return st;
    if ((st = airan::LaunchLdpcRagged(b->cbs, b->reg_list, b->n_reg, b->ws_list, b->n_ws, b->opts.ldpc_iterations,
                                      b->opts.early_stop != 0, b->lam, b->bits, b->iters, b->ldpc_ws, s)) != AIRAN_OK)
        return st;
    return airan::LaunchCrcRagged(b->cbs, ncb, b->bits, b->cb_ok, b->tb_part, b->tb_first, b->tb_count, b->tb_ok,
                                  b->n, s);
}

bool OptionsOk(const AiranGrantBatchOptionsV1 &o)
{
    return o.abi_version == AIRAN_PUSCH_MIMO_ABI_VERSION && o.struct_size >= sizeof(AiranGrantBatchOptionsV1) &&
           (o.iq_format == AIRAN_IQ_INT16_Q8_8 || o.iq_format == AIRAN_IQ_COMPLEX64) && o.ldpc_iterations >= 1 &&
           o.ldpc_iterations <= 64 && o.ce_max_rank >= 1 && o.ce_max_rank <= 512 && o.tau_s >= 0.0 &&
           o.fd_hz >= 0.0 && o.scs_hz > 0.0 && (o.detect_tile == 64 || o.detect_tile == 128 || o.detect_tile == 256);
}

}  // namespace

extern "C" {

int32_t airan_pusch_grants_Create(const PuschMimoConfig *configs, const AiranTbPlanV1 *plans, uint32_t num_grants,
                                  const AiranGrantBatchOptionsV1 *options, void *stream, AiranGrantBatch **out)
{
    if (configs == nullptr || plans == nullptr || options == nullptr || stream == nullptr || out == nullptr ||
        num_grants == 0 || num_grants > 4096 || !OptionsOk(*options))
        return AIRAN_INVALID_ARGUMENT;
    const uint32_t n = num_grants;
    std::vector<Geom> geom(n);
    for (uint32_t i = 0; i < n; ++i) {
        Geom &g = geom[i];
        int32_t st = airan::DeriveGeom(configs[i], &g);
        if (st != AIRAN_OK) return st;
        const AiranTbPlanV1 &p = plans[i];
        if (configs[i].flags != 0 || g.ndmrs > 2 || (g.pad & 1u) != 0 || g.fft < 128) return AIRAN_UNSUPPORTED_PROFILE;
        // One slot per TB, with CRCs (the ragged CRC stage needs them).
        if (p.C == 0 || p.C > airan::kCrcMaxCb || p.Z < 24 || p.Z > 384 || p.K != 22 * p.Z || p.crc_tb == 0 ||
            p.G != static_cast<uint64_t>(g.cw_symbols) * g.qm || p.E_low % g.qm != 0 || p.E_high % g.qm != 0 ||
            static_cast<uint64_t>(p.n_low) * p.E_low + static_cast<uint64_t>(p.C - p.n_low) * p.E_high != p.G ||
            p.Ncb <= p.filler || (p.C > 1) != (p.crc_cb == 24))
            return AIRAN_INVALID_ARGUMENT;
    }
    auto *b = new AiranGrantBatch{};
    b->n = n;
    b->opts = *options;
    auto fail = [&](int32_t code) {
        airan_pusch_grants_Destroy(b);
        return code;
    };
    if (hipGetDevice(&b->device) != hipSuccess || airan_InitDevice() != AIRAN_OK) return fail(AIRAN_LAUNCH_FAILED);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Shapes: sort grants by shape so that a shape's grants are adjacent in
    // the channel buffers (h_ls, noise, H).
    std::vector<ShapeKey> keys(n);
    for (uint32_t i = 0; i < n; ++i) keys[i] = KeyOf(b->device, geom[i], *options);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return keys[x] < keys[y]; });
    std::vector<RaggedGrant> rec(n);
    b->layout.resize(n);
    uint64_t iq = 0, rx = 0, ref = 0, cwo = 0, gold = 0, bits = 0, lam = 0, ws = 0, cbi = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Geom &g = geom[i];
        const AiranTbPlanV1 &p = plans[i];
        RaggedGrant &r = rec[i];
        r.g = airan::ToDev(g);
        for (uint32_t d = 0; d < g.ndmrs; ++d) r.dmrs_cinit[d] = airan::HostDmrsCinit(configs[i], g.dmrs_sym[d]);
        for (uint32_t l = 0; l < g.nl; ++l) r.wf_odd[l] = (configs[i].dmrs_ports[l] - 1000u) & 1u;
        r.dmrs_off = 6u * configs[i].rb_start;
        r.floor16 = HostHalf(airan_pusch_NoEffFloor(g.qm));
        r.iq_off = iq;
        r.rx_off = rx;
        r.ref_off = ref;
        r.cw_off = cwo;
        r.gold_off = static_cast<uint32_t>(gold);
        AiranGrantLayoutV1 &lo = b->layout[i];
        lo.iq_offset = iq;
        lo.iq_samples = static_cast<uint64_t>(g.nr) * g.samples;
        lo.bits_offset = bits;
        lo.cw_offset = cwo;
        lo.first_cb = static_cast<uint32_t>(cbi);
        lo.num_cb = p.C;
        lo.K = p.K;
        lo.fft = g.fft;
        iq += lo.iq_samples;
        rx += static_cast<uint64_t>(g.nr) * g.nsym * g.pad;
        ref += static_cast<uint64_t>(g.nl) * g.ndmrs * g.ref_pad;
        cwo += static_cast<uint64_t>(g.qm) * g.cw_stride;
        gold += (static_cast<uint64_t>(g.cw_symbols) * g.qm + 31) / 32 + 2;
        bits += static_cast<uint64_t>(p.C) * p.K;
        cbi += p.C;
    }
    // Channel buffers in shape order.
    uint64_t hls = 0, noise = 0, h = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t i = order[j];
        const Geom &g = geom[i];
        if (b->groups.empty() || std::memcmp(&b->groups.back().key, &keys[i], sizeof(ShapeKey)) != 0) {
            AiranGrantBatch::CeGroup gp;
            gp.key = keys[i];
            gp.nl = g.nl;
            gp.hls_off = hls;
            gp.noise_off = noise;
            gp.h_off = h;
            b->groups.push_back(gp);
        }
        ++b->groups.back().count;
        rec[i].hls_off = hls;
        rec[i].noise_off = static_cast<uint32_t>(noise);
        rec[i].h_off = h;
        hls += static_cast<uint64_t>(g.nr) * g.nl * g.ndmrs * g.pilot_pad;
        noise += g.nr;
        h += static_cast<uint64_t>(g.nr) * g.nl * g.nsym * g.pad;
    }
    // Filters from the cache (built on a miss).
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        for (auto &gp : b->groups) {
            CacheEntry &e = g_cache[gp.key];
            if (e.filter == nullptr) {
                uint32_t first = 0;
                while (std::memcmp(&keys[first], &gp.key, sizeof(ShapeKey)) != 0) ++first;
                LmmseFilterParamsV1 fp{};
                fp.abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
                fp.struct_size = sizeof(fp);
                fp.max_rank = options->ce_max_rank;
                fp.tau_s = options->tau_s;
                fp.fd_hz = options->fd_hz;
                fp.scs_hz = options->scs_hz;
                fp.eig_floor = 1e-12;
                const int32_t st = airan_channel_est_lmmse_CreateFilter(&configs[first], &fp, stream, &e.filter);
                if (st != AIRAN_OK) {
                    g_cache.erase(gp.key);
                    return fail(st);
                }
            }
            ++e.users;
            gp.filter = e.filter;
        }
    }
    size_t ce_ws = 0;
    for (uint32_t j = 0, k = 0; k < b->groups.size(); ++k) {
        const Geom &g = geom[order[j]];
        b->ce_ws_off.push_back(ce_ws);
        ce_ws += (airan::CeWorkspaceBytes(g, b->groups[k].count) + 255) / 256 * 256;
        j += b->groups[k].count;
    }
    b->ce_ws_bytes = ce_ws;
    // Work lists.
    std::map<uint32_t, std::vector<RaggedStream>> streams;
    std::vector<RaggedItem> gen, ls, nz;
    std::vector<RaggedCb> cbs;
    std::vector<uint32_t> reg, wsl, tb_first(n), tb_count(n);
    for (uint32_t i = 0; i < n; ++i) {
        const Geom &g = geom[i];
        const AiranTbPlanV1 &p = plans[i];
        for (uint32_t r = 0; r < g.nr; ++r)
            streams[g.fft].push_back({rec[i].iq_off + static_cast<uint64_t>(r) * g.samples,
                                      rec[i].rx_off + static_cast<uint64_t>(r) * g.nsym * g.pad, g.used, g.pad});
        for (uint32_t l = 0; l < g.nl; ++l) gen.push_back({i, l, 0});
        for (uint32_t r = 0; r < g.nr; ++r)
            for (uint32_t l = 0; l < g.nl; ++l) ls.push_back({i, r, l});
        for (uint32_t r = 0; r < g.nr; ++r) nz.push_back({i, r, 0});
        for (uint32_t ds = 0; ds < g.ndata; ++ds) b->host_tiles_src.push_back({i, ds, g.used});
        b->max_ref_pad = std::max(b->max_ref_pad, g.ref_pad);
        b->max_pilot_pad = std::max(b->max_pilot_pad, g.pilot_pad);
        b->max_dmrs = std::max(b->max_dmrs, g.ndmrs);
        const uint32_t rows = airan_ldpc_ActiveRows(&p);
        tb_first[i] = b->layout[i].first_cb;
        tb_count[i] = p.C;
        uint64_t sym_off = 0;
        for (uint32_t r = 0; r < p.C; ++r) {
            RaggedCb c{};
            c.grant = i;
            c.e = r < p.n_low ? p.E_low : p.E_high;
            c.k0 = p.k0;
            c.ncb = p.Ncb;
            c.cw_symbol_offset = static_cast<uint32_t>(sym_off);
            sym_off += c.e / g.qm;
            c.z = p.Z;
            c.kprime = p.Kprime;
            c.filler = p.filler;
            c.k = p.K;
            c.rows = rows;
            c.lam_off = lam;
            lam += 68ull * p.Z;
            c.bits_off = b->layout[i].bits_offset + static_cast<uint64_t>(r) * p.K;
            c.ws_off = ws;
            if (rows > 8) ws += static_cast<uint64_t>(rows) * p.Z;
            c.crc_cb = p.crc_cb;
            c.len_tb = p.crc_tb;
            c.poly_tb = airan::CrcTbPoly(p.crc_tb);
            c.xpow = airan::CrcXpowHost(p, r);
            (rows > 8 ? wsl : reg).push_back(static_cast<uint32_t>(cbs.size()));
            cbs.push_back(c);
        }
    }
    b->num_cb = cbs.size();
    b->iq_total = iq;
    if (b->num_cb > 65535) return fail(AIRAN_INVALID_ARGUMENT);
    size_t &by = b->bytes;
    for (auto &kv : streams) {
        AiranGrantBatch::FftGroup fg;
        fg.fft = kv.first;
        fg.count = static_cast<uint32_t>(kv.second.size());
        if (!Upload(&fg.streams, kv.second, &by)) return fail(AIRAN_RESOURCE_TOO_SMALL);
        b->ffts.push_back(fg);
    }
    b->n_gen = static_cast<uint32_t>(gen.size());
    b->n_ls = static_cast<uint32_t>(ls.size());
    b->n_noise = static_cast<uint32_t>(nz.size());
    b->n_reg = static_cast<uint32_t>(reg.size());
    b->n_ws = static_cast<uint32_t>(wsl.size());
    if (!Upload(&b->grants, rec, &by) || !Upload(&b->gen_items, gen, &by) || !Upload(&b->ls_items, ls, &by) ||
        !Upload(&b->noise_items, nz, &by) || !Upload(&b->cbs, cbs, &by) || !Upload(&b->reg_list, reg, &by) ||
        !Upload(&b->ws_list, wsl, &by) || !Upload(&b->tb_first, tb_first, &by) ||
        !Upload(&b->tb_count, tb_count, &by) || !DevAlloc(&b->rx_re, rx, &by) || !DevAlloc(&b->rx_im, rx, &by) ||
        !DevAlloc(&b->ref_re, ref, &by) || !DevAlloc(&b->ref_im, ref, &by) || !DevAlloc(&b->hls_re, hls, &by) ||
        !DevAlloc(&b->hls_im, hls, &by) || !DevAlloc(&b->noise_ls, noise, &by) ||
        !DevAlloc(&b->noise_ce, noise, &by) || !DevAlloc(&b->h_re, h, &by) || !DevAlloc(&b->h_im, h, &by) ||
        !DevAlloc(&b->gold, gold, &by) ||
        !DevAlloc(&b->cw, cwo, &by) || !DevAlloc(&b->lam, lam, &by) || !DevAlloc(&b->ldpc_ws, ws, &by) ||
        !DevAlloc(&b->bits, bits, &by) || !DevAlloc(&b->iters, b->num_cb, &by) ||
        !DevAlloc(&b->cb_ok, b->num_cb, &by) || !DevAlloc(&b->tb_ok, n, &by) ||
        !DevAlloc(&b->tb_part, b->num_cb, &by))
        return fail(AIRAN_RESOURCE_TOO_SMALL);
    if (!DevAlloc(&b->ce_ws_all, b->ce_ws_bytes, &by)) return fail(AIRAN_RESOURCE_TOO_SMALL);
    {
        std::vector<airan::CeGroupJob> jobs;
        for (size_t k = 0; k < b->groups.size(); ++k) {
            const auto &gp = b->groups[k];
            const size_t end = k + 1 < b->groups.size() ? b->ce_ws_off[k + 1] : b->ce_ws_bytes;
            jobs.push_back(airan::CeGroupJob{gp.filter, b->hls_re + gp.hls_off, b->hls_im + gp.hls_off,
                                             b->noise_ls + gp.noise_off, b->noise_ce + gp.noise_off,
                                             b->h_re + gp.h_off, b->h_im + gp.h_off, b->ce_ws_all + b->ce_ws_off[k],
                                             end - b->ce_ws_off[k], gp.nl, gp.count});
        }
        const int32_t cst = airan::CreateCeGrouped(jobs.data(), static_cast<uint32_t>(jobs.size()), &b->ce_grouped);
        if (cst != AIRAN_OK) return fail(cst);
    }
    int32_t st = BuildTiles(b, options->detect_tile);
    if (st != AIRAN_OK) return fail(st);
    // Gold bits of every grant's data scrambling sequence.
    for (uint32_t i = 0; i < n; ++i) {
        const Geom &g = geom[i];
        const uint32_t words = static_cast<uint32_t>((static_cast<uint64_t>(g.cw_symbols) * g.qm + 31) / 32 + 2);
        const uint32_t cinit = (static_cast<uint32_t>(configs[i].rnti) << 15) |
                               ((static_cast<uint32_t>(configs[i].codeword_index) & 1u) << 14) |
                               configs[i].data_scrambling_id;
        hipLaunchKernelGGL(GrantGoldKernel, dim3(airan::CeilDiv(words, 256)), dim3(256), 0, s,
                           b->gold + rec[i].gold_off, words, cinit);
    }
    if (hipMemsetAsync(b->cw, 0, cwo * sizeof(int16_t), s) != hipSuccess) return fail(AIRAN_LAUNCH_FAILED);
    // One eager pass on zero IQ loads every per-device table before any
    // graph capture.
    void *zero = nullptr;
    const size_t iq_bytes = iq * (options->iq_format == AIRAN_IQ_COMPLEX64 ? 8 : 4);
    if (hipMalloc(&zero, std::max<size_t>(iq_bytes, 256)) != hipSuccess) return fail(AIRAN_RESOURCE_TOO_SMALL);
    if (hipMemsetAsync(zero, 0, iq_bytes, s) != hipSuccess) {
        (void)hipFree(zero);
        return fail(AIRAN_LAUNCH_FAILED);
    }
    st = EnqueueStages(b, zero, s);
    const bool synced = hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(zero);
    if (st != AIRAN_OK) return fail(st);
    if (!synced) return fail(AIRAN_LAUNCH_FAILED);
    *out = b;
    return AIRAN_OK;
}

void airan_pusch_grants_Destroy(AiranGrantBatch *b)
{
    if (b == nullptr) return;
    DropGraph(b);
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        for (auto &gp : b->groups) {
            auto it = g_cache.find(gp.key);
            if (gp.filter != nullptr && it != g_cache.end() && it->second.users > 0) --it->second.users;
        }
    }
    for (auto &f : b->ffts)
        if (f.streams) (void)hipFree(f.streams);
    airan::DestroyCeGrouped(b->ce_grouped);
    if (b->ce_ws_all) (void)hipFree(b->ce_ws_all);
    void *bufs[] = {b->grants, b->gen_items, b->ls_items, b->noise_items, b->tiles, b->cbs, b->reg_list, b->ws_list,
                    b->tb_first, b->tb_count, b->rx_re, b->rx_im, b->ref_re, b->ref_im, b->hls_re, b->hls_im,
                    b->noise_ls, b->noise_ce, b->h_re, b->h_im, b->gold, b->cw, b->lam, b->ldpc_ws,
                    b->bits, b->iters, b->cb_ok, b->tb_ok, b->tb_part};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    delete b;
}

int32_t airan_pusch_grants_Layout(const AiranGrantBatch *b, AiranGrantLayoutV1 *layouts, uint32_t num_grants,
                                  uint64_t *iq_samples_total)
{
    if (b == nullptr || layouts == nullptr || num_grants < b->n) return AIRAN_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < b->n; ++i) layouts[i] = b->layout[i];
    if (iq_samples_total) *iq_samples_total = b->iq_total;
    return AIRAN_OK;
}

int32_t airan_pusch_grants_Enqueue(AiranGrantBatch *b, const void *iq, void *stream)
{
    if (b == nullptr || iq == nullptr || stream == nullptr) return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!b->opts.use_graph) return EnqueueStages(b, iq, s);
    if (b->exec == nullptr || b->graph_iq != iq || b->graph_stream != stream) {
        DropGraph(b);
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return AIRAN_LAUNCH_FAILED;
        const int32_t st = EnqueueStages(b, iq, s);
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(s, &graph);
        if (st != AIRAN_OK || end != hipSuccess || graph == nullptr) {
            if (graph) (void)hipGraphDestroy(graph);
            return st != AIRAN_OK ? st : AIRAN_LAUNCH_FAILED;
        }
        b->graph = graph;
        if (hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            DropGraph(b);
            return AIRAN_LAUNCH_FAILED;
        }
        b->graph_iq = iq;
        b->graph_stream = stream;
    }
    return hipGraphLaunch(b->exec, s) == hipSuccess ? AIRAN_OK : AIRAN_LAUNCH_FAILED;
}

int32_t airan_pusch_grants_Outputs(const AiranGrantBatch *b, AiranGrantBatchOutputsV1 *o)
{
    if (b == nullptr || o == nullptr || o->abi_version != AIRAN_PUSCH_MIMO_ABI_VERSION ||
        o->struct_size < sizeof(AiranGrantBatchOutputsV1))
        return AIRAN_INVALID_ARGUMENT;
    o->num_grants = b->n;
    o->bits = b->bits;
    o->tb_ok = b->tb_ok;
    o->cb_ok = b->cb_ok;
    o->iterations = b->iters;
    o->cw_llr = b->cw;
    o->num_cb = b->num_cb;
    return AIRAN_OK;
}

int32_t airan_pusch_grants_SetDetectTile(AiranGrantBatch *b, uint32_t tile)
{
    if (b == nullptr || (tile != 64 && tile != 128 && tile != 256)) return AIRAN_INVALID_ARGUMENT;
    if (tile == b->tile) return AIRAN_OK;
    DropGraph(b);
    return BuildTiles(b, tile);
}

uint32_t airan_pusch_grants_NumShapes(const AiranGrantBatch *b)
{
    return b == nullptr ? 0 : static_cast<uint32_t>(b->groups.size());
}

uint32_t airan_pusch_grants_FilterCacheSize(void)
{
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    return static_cast<uint32_t>(g_cache.size());
}

void airan_pusch_grants_ClearFilterCache(void)
{
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    for (auto it = g_cache.begin(); it != g_cache.end();) {
        if (it->second.users == 0) {
            airan_channel_est_lmmse_DestroyFilter(it->second.filter);
            it = g_cache.erase(it);
        } else {
            ++it;
        }
    }
}

}  // extern "C"
