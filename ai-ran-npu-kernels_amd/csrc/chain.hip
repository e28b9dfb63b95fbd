// chain.hip -- the batched PUSCH receive chain (include/airan_pusch_chain.h).
//
// One pipeline of batched launches over `slots` slots (transport blocks of
// one or more consecutive slots); the only kernel specific to the chain is
// the fused detector/demapper, the rest are the per-operator kernels launched
// over all slots or code blocks at once.
#include <algorithm>
#include <cmath>
#include <vector>

#include "airan_ce_refiner.h"
#include "airan_pusch_chain.h"
#include "common.h"
#include "cgemm.h"
#include "crc.h"
#include "detect.h"
#include "detect_demap.h"
#include "gold.h"
#include "launch.h"
#include "qam.h"

using airan::Cf;
using airan::Geom;

namespace {

struct DemapArgs {
    uint32_t nr, qm, used, pad, nsym, cw_symbols, cw_stride;
    uint32_t slots_per_tb;  // the Gold sequence runs on across a TB's slots
    uint32_t data_sym[14];
    uint16_t floor16;  // fp16 bits of the no_eff floor
};

// Packed Gold bits of the data scrambling sequence of one transport block:
// word w holds c(32 w .. 32 w + 31), bit j = c(32 w + j).
__global__ void __launch_bounds__(256) GoldBitsKernel(uint32_t *__restrict__ bits, uint32_t words, uint32_t cinit)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    uint32_t s1, s2;
    airan::GoldStateAt(32ull * w + 1600ull, cinit, &s1, &s2);
    bits[w] = airan::GoldNext32(&s1, &s2);
}

// Detection + QAM demap + layer demap + descramble for one data RE of one
// slot (grid (used / 256, ndata, slots)).  The arithmetic is that of the
// operators, step by step:
//   no     = fp16(sum_rx fp32(noise_rx) * (1 / NR))     (detector I/O pack)
//   x, ne  = Detect<L>(...) rounded to fp16              (detect_lmmse)
//   ne     = max(ne, fp16(floor))                        (no_eff floor)
//   llr    = QamSymbolLlrs(x, ne)                        (qam256_demod)
//   cw[b][re L + l] = +-llr[stream(b)] by c(n0 + (re L + l) Qm + b)
//                                                        (layer_demap, descramble)
// where n0 = (slot mod slots_per_tb) * codeword_symbols * Qm: the Gold index
// continues across the slots of a transport block over transmitted symbols
// only (descramble_mimo.cpp:176-184).
// QM = 8 fixes the modulation at compile time (the bit -> stream mapping then
// indexes registers); QM = 0 takes it from the arguments.  ZF (one Rx, one
// layer): single-tap zero forcing with n0 = the slot's noise (the
// equalize_zf_siso arithmetic) in place of the LMMSE solve.
typedef _Float16 h16x4_t __attribute__((ext_vector_type(4)));
typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <uint32_t GRAM>
__device__ inline f32x4_t Mfma4(h16x4_t a, h16x4_t b, f32x4_t c)
{
    if constexpr (GRAM == AIRAN_GRAM_FP16) {
        return __builtin_amdgcn_mfma_f32_4x4x4f16(a, b, c, 0, 0, 0);
    } else {
        // bf16 operands: the fp16 values rounded to 8 significand bits.
        s16x4_t ab, bb;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ab[k] = static_cast<short>(airan::Bf16Bits(static_cast<float>(a[k])));
            bb[k] = static_cast<short>(airan::Bf16Bits(static_cast<float>(b[k])));
        }
        return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(ab, bb, c, 0, 0, 0);
    }
}

// Gram matrix and matched filter of the 64 REs (subcarriers) of this wave on
// the matrix cores.  v_mfma_f32_4x4x4_16b: lane 4 b + i holds row i of block
// b's A and column i of its B (K = the 4 items), D[r][i] of block b lands in
// lane 4 b + i, register r.  Block b is one RE: lane 4 b + j loads H[rx][j]
// for 4 Rx at a time, which serves as A (rows = layers: A = H^T) and as B,
// so D = H^T H per part; G = H^H H = (Hr^T Hr + Hi^T Hi) + i (Hr^T Hi -
// Hi^T Hr), m = H^H y likewise with y in every column of B.  Four groups of
// 16 REs per wave; the columns go through LDS to the thread that owns the
// RE (lane r = subcarrier base + r), which gets A (lower triangle) and m.
template <int NL, uint32_t GRAM>
__device__ inline void GramMfma(const __half *__restrict__ rx_re, const __half *__restrict__ rx_im,
                                const __half *__restrict__ h_re, const __half *__restrict__ h_im, size_t rx_slot,
                                size_t h_slot, size_t plane, uint32_t sym_off, uint32_t sc_base, uint32_t used,
                                uint32_t nr, float (*gl)[40], Cf (&A)[4][4], Cf (&m)[4])
{
    const uint32_t lane = threadIdx.x & 63u, b = lane >> 2, j = lane & 3u;
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) {
        const uint32_t scg = sc_base + 16 * g + b;
        const bool re_ok = scg < used;
        const bool h_ok = re_ok && j < static_cast<uint32_t>(NL);
        f32x4_t gr = {0, 0, 0, 0}, gi = gr, mr = gr, mi = gr;
        for (uint32_t r0 = 0; r0 < nr; r0 += 4) {
            h16x4_t hr, hi, yr, yi;
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t r = r0 + k;
                const bool rk = r < nr;
                const size_t hidx = h_slot + (static_cast<size_t>(r) * NL + j) * plane + sym_off + scg;
                const size_t yidx = rx_slot + static_cast<size_t>(r) * plane + sym_off + scg;
                hr[k] = h_ok && rk ? __builtin_bit_cast(_Float16, h_re[hidx]) : static_cast<_Float16>(0.0f);
                hi[k] = h_ok && rk ? __builtin_bit_cast(_Float16, h_im[hidx]) : static_cast<_Float16>(0.0f);
                yr[k] = re_ok && rk ? __builtin_bit_cast(_Float16, rx_re[yidx]) : static_cast<_Float16>(0.0f);
                yi[k] = re_ok && rk ? __builtin_bit_cast(_Float16, rx_im[yidx]) : static_cast<_Float16>(0.0f);
            }
            const h16x4_t nhi = -hi;
            gr = Mfma4<GRAM>(hr, hr, gr);
            gr = Mfma4<GRAM>(hi, hi, gr);
            gi = Mfma4<GRAM>(hr, hi, gi);
            gi = Mfma4<GRAM>(nhi, hr, gi);
            mr = Mfma4<GRAM>(hr, yr, mr);
            mr = Mfma4<GRAM>(hi, yi, mr);
            mi = Mfma4<GRAM>(hr, yi, mi);
            mi = Mfma4<GRAM>(nhi, yr, mi);
        }
        float *dst = gl[16 * g + b];
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            dst[(i * 4 + j) * 2] = gr[i];
            dst[(i * 4 + j) * 2 + 1] = gi[i];
        }
        if (j == 0) {
#pragma unroll
            for (uint32_t i = 0; i < 4; ++i) {
                dst[32 + 2 * i] = mr[i];
                dst[33 + 2 * i] = mi[i];
            }
        }
    }
    __syncthreads();
    const float *src = gl[lane];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        m[i] = {src[32 + 2 * i], src[33 + 2 * i]};
#pragma unroll
        for (int jj = 0; jj <= i; ++jj) A[i][jj] = {src[(i * 4 + jj) * 2], src[(i * 4 + jj) * 2 + 1]};
    }
}

template <int NL, int QM, bool ZF = false, uint32_t GRAM = AIRAN_GRAM_FP32>
__global__ void __launch_bounds__(256) DetectDemapKernel(const __half *__restrict__ rx_re,
                                                         const __half *__restrict__ rx_im,
                                                         const __half *__restrict__ h_re,
                                                         const __half *__restrict__ h_im,
                                                         const __half *__restrict__ noise,
                                                         const uint32_t *__restrict__ gold,
                                                         int16_t *__restrict__ cw, DemapArgs a)
{
    const uint32_t sc = blockIdx.x * blockDim.x + threadIdx.x;
    // The matrix-core Gram needs every lane of the workgroup (the MFMAs and
    // the barrier); its lanes past the last subcarrier leave after it.
    if (GRAM == AIRAN_GRAM_FP32 && sc >= a.used) return;
    const uint32_t ds = blockIdx.y, slot = blockIdx.z;
    const uint32_t sym = a.data_sym[ds];
    const uint32_t nr = a.nr;
    float s = 0.0f;
    for (uint32_t r = 0; r < nr; ++r) s += __half2float(noise[static_cast<size_t>(slot) * nr + r]);
    const float no = __half2float(__float2half_rn(s * (1.0f / static_cast<float>(nr))));
    const size_t plane = static_cast<size_t>(a.nsym) * a.pad;
    const size_t rx_slot = static_cast<size_t>(slot) * nr * plane;
    const size_t h_slot = rx_slot * NL;
    // The slot's planes from uniform base pointers with 32-bit element
    // offsets (NR NL nsym pad < 2^32 for every geometry).
    const __half *hr_s = h_re + h_slot, *hi_s = h_im + h_slot, *yr_s = rx_re + rx_slot, *yi_s = rx_im + rx_slot;
    const uint32_t plane32 = a.nsym * a.pad, off32 = sym * a.pad + sc;
    auto loadh = [&](int r, int l) {
        const uint32_t i = (static_cast<uint32_t>(r) * NL + static_cast<uint32_t>(l)) * plane32 + off32;
        return Cf{__half2float(hr_s[i]), __half2float(hi_s[i])};
    };
    auto loady = [&](int r) {
        const uint32_t i = static_cast<uint32_t>(r) * plane32 + off32;
        return Cf{__half2float(yr_s[i]), __half2float(yi_s[i])};
    };
    Cf xh[4];
    float ne[4];
    if constexpr (ZF) {
        static_assert(NL == 1, "zero forcing is the SISO equaliser");
        const Cf h = loadh(0, 0), y = loady(0);
        airan::ZfSolve(h.r, h.i, y.r, y.i, no, &xh[0].r, &xh[0].i, &ne[0]);
    } else if constexpr (GRAM != AIRAN_GRAM_FP32) {
        __shared__ float gl[4][64][40];  // per wave: 64 REs x (G 4x4 + m 4) complex
        Cf A[4][4], m[4];
        GramMfma<NL, GRAM>(rx_re, rx_im, h_re, h_im, rx_slot, h_slot, plane, static_cast<uint32_t>(sym) * a.pad,
                           sc - (threadIdx.x & 63u), a.used, nr, gl[threadIdx.x >> 6], A, m);
        if (sc >= a.used) return;
        airan::DetectFromGram<NL>(no, A, m, xh, ne);
    } else {
        airan::Detect<NL>(static_cast<int>(nr), no, loadh, loady, xh, ne);
    }
    const uint32_t qm = QM > 0 ? static_cast<uint32_t>(QM) : a.qm, m = qm / 2;
    const uint32_t re = ds * a.used + sc;
    const uint32_t slot_in_tb = slot - (slot / a.slots_per_tb) * a.slots_per_tb;
    const uint64_t bit0 = (static_cast<uint64_t>(slot_in_tb) * a.cw_symbols + static_cast<uint64_t>(re) * NL) * qm;
    const uint32_t w0 = static_cast<uint32_t>(bit0 >> 5);
    const uint64_t gw = (static_cast<uint64_t>(gold[w0]) | (static_cast<uint64_t>(gold[w0 + 1]) << 32)) >> (bit0 & 31u);
    int16_t *out = cw + static_cast<size_t>(slot) * qm * a.cw_stride + static_cast<size_t>(re) * NL;
    airan::DemapStore<NL, QM>(xh, ne, qm, __ushort_as_half(a.floor16), gw, out, a.cw_stride);
}

using DdKernel = void (*)(const __half *, const __half *, const __half *, const __half *, const __half *,
                          const uint32_t *, int16_t *, DemapArgs);

template <uint32_t GRAM>
DdKernel PickDetectGram(uint32_t nl, bool q8)
{
    switch (nl) {
    case 1: return q8 ? DetectDemapKernel<1, 8, false, GRAM> : DetectDemapKernel<1, 0, false, GRAM>;
    case 2: return q8 ? DetectDemapKernel<2, 8, false, GRAM> : DetectDemapKernel<2, 0, false, GRAM>;
    case 3: return q8 ? DetectDemapKernel<3, 8, false, GRAM> : DetectDemapKernel<3, 0, false, GRAM>;
    default: return q8 ? DetectDemapKernel<4, 8, false, GRAM> : DetectDemapKernel<4, 0, false, GRAM>;
    }
}

// nl = 0: the SISO zero-forcing equaliser.
DdKernel PickDetect(uint32_t nl, bool q8, uint32_t gram)
{
    if (nl == 0) return q8 ? DetectDemapKernel<1, 8, true> : DetectDemapKernel<1, 0, true>;
    if (gram == AIRAN_GRAM_FP16) return PickDetectGram<AIRAN_GRAM_FP16>(nl, q8);
    if (gram == AIRAN_GRAM_BF16) return PickDetectGram<AIRAN_GRAM_BF16>(nl, q8);
    return PickDetectGram<AIRAN_GRAM_FP32>(nl, q8);
}

// Host fp32 -> fp16 bits, round to nearest even (normal range only).
uint16_t HostF2H(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int32_t exp = static_cast<int32_t>((x >> 23) & 0xff) - 127 + 15;
    uint32_t mant = x & 0x7fffffu;
    if (exp <= 0 || exp >= 31) return static_cast<uint16_t>(sign | (exp >= 31 ? 0x7c00u : 0u));
    uint32_t h = sign | (static_cast<uint32_t>(exp) << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return static_cast<uint16_t>(h);
}

// keep[slot] = 1 when the slot's SNR estimate, 10 log10(1 / mean_rx noise),
// lies in [snr_min_db, snr_max_db] (the refiner's training range).
__global__ void __launch_bounds__(64) RefinerKeepKernel(const __half *__restrict__ noise, uint8_t *__restrict__ keep,
                                                        uint32_t nr, uint32_t slots, float snr_min_db,
                                                        float snr_max_db)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    float acc = 0.0f;
    for (uint32_t r = 0; r < nr; ++r) acc += __half2float(noise[static_cast<size_t>(s) * nr + r]);
    const float snr = -10.0f * log10f(fmaxf(acc / static_cast<float>(nr), 1e-30f));
    keep[s] = snr >= snr_min_db && snr <= snr_max_db ? 1 : 0;
}

template <typename T>
bool Alloc(T **p, size_t count, size_t *total)
{
    *p = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    if (hipMalloc(reinterpret_cast<void **>(p), bytes) != hipSuccess) return false;
    *total += bytes;
    return true;
}

void Free(void *p)
{
    if (p != nullptr) (void)hipFree(p);
}

}  // namespace

struct AiranPuschChain {
    int device = 0;
    PuschMimoConfig cfg{};
    AiranTbPlanV1 plan{};
    AiranPuschChainOptionsV1 opts{};
    Geom g{};
    uint32_t slots_per_tb = 1;   // plan.G / (codeword_symbols * Qm)
    bool crc = true;             // the plan carries a TB CRC (the reference's fixed plan has none)
    bool ls_zf = false;          // AIRAN_CHAIN_RX_LS_ZF: LS + linear interpolation, zero forcing
    uint16_t *slot_default = nullptr;  // [max_slots] config->slot_number (airan_pusch_chain_Enqueue)
    uint32_t active_rows = 0;
    bool fused_ldpc = false;  // decoder reads the codeword LLRs (no lam buffer)
    uint32_t last_slots = 0;
    size_t bytes = 0;
    LmmseFilter *filter = nullptr;
    __half *dmrs_re = nullptr, *dmrs_im = nullptr;
    __half *rx_re = nullptr, *rx_im = nullptr;
    __half *hls_re = nullptr, *hls_im = nullptr;
    __half *noise_ls = nullptr, *noise_ce = nullptr;
    uint8_t *ce_ws = nullptr;
    size_t ce_ws_bytes = 0;
    __half *h_re = nullptr, *h_im = nullptr;
    uint32_t *gold = nullptr;
    int16_t *cw = nullptr;
    RateMatchDescriptor *desc = nullptr;
    int16_t *lam = nullptr;
    uint8_t *ldpc_ws = nullptr;
    size_t ldpc_ws_bytes = 0;
    int8_t *bits = nullptr;
    uint8_t *iters = nullptr, *cb_ok = nullptr, *tb_ok = nullptr;
    uint8_t *crc_ws = nullptr;
    size_t crc_ws_bytes = 0;
    void *zero_iq = nullptr;
    // Learned refiner (optional): replaces the LMMSE estimate of the slots
    // whose SNR estimate lies in [snr_min, snr_max].
    AiranCeRefiner *refiner = nullptr;
    float refiner_snr_min = 0.0f, refiner_snr_max = 0.0f;
    uint8_t *refiner_keep = nullptr;
    uint32_t *crc_xpow = nullptr;  // [C] code-block shifts of the TB CRC remainder
    uint32_t *ldpc_queue = nullptr;  // decoder work counter (2 words, left zero by each decode)
    // Stage timing (events before the first stage and after each stage).
    bool timing = false;
    hipEvent_t ev[AIRAN_CHAIN_STAGES + 1] = {};
    // Cached graph for the last (iq, slots, stream).
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *graph_iq = nullptr;
    const void *graph_slot_no = nullptr;
    uint32_t graph_slots = 0;
    void *graph_stream = nullptr;
};

namespace {

// The CRC check's tables: the part of crc_ws after the remainder parts
// (airan_crc_check_WorkspaceBytes lays it out so).
void *CrcTables(AiranPuschChain *c)
{
    return c->crc_ws + airan::CrcTableOffset(c->opts.max_slots);
}

void DropGraph(AiranPuschChain *c)
{
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->exec = nullptr;
    c->graph = nullptr;
    c->graph_iq = nullptr;
    c->graph_slot_no = nullptr;
    c->graph_slots = 0;
    c->graph_stream = nullptr;
}

int32_t EnqueueStages(AiranPuschChain *c, const void *iq, const uint16_t *slot_no, uint32_t slots, hipStream_t s,
                      bool timed)
{
    const Geom &g = c->g;
    int stage = 0;
    auto mark = [&]() {
        if (timed) (void)hipEventRecord(c->ev[stage++], s);
    };
    mark();
    OfdmDemodBatchOpArgsV1 oa{};
    oa.abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    oa.struct_size = sizeof(oa);
    oa.iq = iq;
    oa.iq_format = c->opts.iq_format;
    oa.batch = slots * g.nr;
    oa.precision = c->opts.dft_precision;
    oa.config = &c->cfg;
    oa.stream = s;
    int32_t st = airan_ofdm_demod_redemap_Enqueue(&oa, c->rx_re, c->rx_im);
    if (st != AIRAN_OK) return st;
    mark();
    // Each slot's DMRS from its own slot number, then the LS observations.
    st = airan::LaunchDmrsGenSlots(c->cfg, g, slot_no, slots, c->dmrs_re, c->dmrs_im, s);
    if (st != AIRAN_OK) return st;
    st = airan::LaunchDmrsLs(g, c->rx_re, c->rx_im, c->dmrs_re, c->dmrs_im, c->hls_re, c->hls_im, c->noise_ls,
                             slots, s, true);
    if (st != AIRAN_OK) return st;
    mark();
    if (c->ls_zf)
        st = airan::LaunchLsInterpSiso(c->cfg, g, c->rx_re, c->rx_im, c->dmrs_re, c->dmrs_im, true, c->h_re, c->h_im,
                                       slots, s);
    else
        st = airan::LaunchCeApply(c->filter, c->hls_re, c->hls_im, c->noise_ls, c->noise_ce, c->ce_ws,
                                  c->ce_ws_bytes, c->h_re, c->h_im, g.nl, slots, s);
    if (st != AIRAN_OK) return st;
    if (c->refiner != nullptr) {
        hipLaunchKernelGGL(RefinerKeepKernel, dim3(airan::CeilDiv(slots, 64)), dim3(64), 0, s, c->noise_ce,
                           c->refiner_keep, g.nr, slots, c->refiner_snr_min, c->refiner_snr_max);
        st = airan_ce_refiner_EnqueueWithNoise(c->refiner, c->hls_re, c->hls_im, c->noise_ce, slots, c->refiner_keep,
                                               c->h_re, c->h_im, s);
        if (st != AIRAN_OK) return st;
    }
    mark();
    DemapArgs da{};
    da.nr = g.nr;
    da.qm = g.qm;
    da.used = g.used;
    da.pad = g.pad;
    da.nsym = g.nsym;
    da.cw_symbols = g.cw_symbols;
    da.cw_stride = g.cw_stride;
    da.slots_per_tb = c->slots_per_tb;
    for (uint32_t i = 0; i < g.ndata; ++i) da.data_sym[i] = g.data_sym[i];
    da.floor16 = HostF2H(airan_pusch_NoEffFloor(g.qm));
    const dim3 grid(airan::CeilDiv(g.used, 256), g.ndata, slots);
    const DdKernel dd = PickDetect(c->ls_zf ? 0u : g.nl, g.qm == 8, c->opts.gram_precision);
    hipLaunchKernelGGL(dd, grid, dim3(256), 0, s, c->rx_re, c->rx_im, c->h_re, c->h_im,
                       c->ls_zf ? c->noise_ls : c->noise_ce, c->gold, c->cw, da);
    if ((st = airan::LaunchStatus()) != AIRAN_OK) return st;
    mark();
    const uint32_t tbs = slots / c->slots_per_tb;
    const uint32_t ncb = tbs * c->plan.C;
    if (c->fused_ldpc) {
        // Rate dematch runs inside the decoder's load (the stage is empty).
        mark();
        st = airan::LaunchLdpcFromCodeword(g, c->plan, c->desc, c->cw, ncb, c->active_rows, c->opts.ldpc_iterations,
                                           c->opts.early_stop != 0, c->bits, c->iters, c->ldpc_queue, s);
        if (st != AIRAN_OK) return st;
        mark();
        if (c->crc) {
            airan::CrcJob crc{};
            if ((st = airan::CrcJobSetup(c->plan, c->cb_ok, reinterpret_cast<uint32_t *>(c->crc_ws), &crc)) !=
                AIRAN_OK)
                return st;
            crc.xpow = c->crc_xpow;
            airan::LaunchCbCrc(c->bits, crc, c->plan.K, ncb, CrcTables(c), s);
            airan::LaunchCrcFinish(reinterpret_cast<const uint32_t *>(c->crc_ws), c->tb_ok, tbs, c->plan.C, s);
            if ((st = airan::LaunchStatus()) != AIRAN_OK) return st;
        }
        mark();
        return AIRAN_OK;
    }
    st = airan::LaunchRateDematch(g, c->plan, c->desc, c->cw, c->lam, ncb, s);
    if (st != AIRAN_OK) return st;
    mark();
    LdpcDecodeOpArgsV1 la{};
    la.abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    la.struct_size = sizeof(la);
    la.lam_in = c->lam;
    la.bits = c->bits;
    la.iterations_out = c->iters;
    la.num_cb = ncb;
    la.z = c->plan.Z;
    la.k = c->plan.K;
    la.max_iterations = c->opts.ldpc_iterations;
    la.active_rows = c->active_rows;
    la.mode = c->opts.ldpc_mode;
    la.early_stop = c->opts.early_stop ? 1u : 0u;
    la.stream = s;
    st = airan_ldpc_decode_Enqueue(&la, c->ldpc_ws, c->ldpc_ws_bytes);
    if (st != AIRAN_OK) return st;
    mark();
    if (c->crc) {
        st = airan_crc_check_Enqueue(&c->plan, tbs, c->bits, c->cb_ok, c->tb_ok, nullptr, c->crc_ws,
                                     c->crc_ws_bytes, s);
        if (st != AIRAN_OK) return st;
    }
    mark();
    return AIRAN_OK;
}

bool OptionsOk(const AiranPuschChainOptionsV1 &o)
{
    return o.abi_version == AIRAN_PUSCH_MIMO_ABI_VERSION && o.struct_size >= sizeof(AiranPuschChainOptionsV1) &&
           o.max_slots >= 1 && o.max_slots <= 4096 &&
           (o.iq_format == AIRAN_IQ_INT16_Q8_8 || o.iq_format == AIRAN_IQ_COMPLEX64) && o.ldpc_iterations >= 1 &&
           o.ldpc_iterations <= 64 && (o.ldpc_mode == AIRAN_LDPC_MODE_REF || o.ldpc_mode == AIRAN_LDPC_MODE_FP16) &&
           o.ce_max_rank >= 1 && o.ce_max_rank <= 512 && o.tau_s >= 0.0 && o.fd_hz >= 0.0 && o.scs_hz > 0.0 &&
           o.eig_floor >= 0.0 && (o.receiver == AIRAN_CHAIN_RX_LMMSE || o.receiver == AIRAN_CHAIN_RX_LS_ZF) &&
           o.dft_precision <= AIRAN_DFT_FP16 && o.gram_precision <= AIRAN_GRAM_FP16;
}

}  // namespace

extern "C" {

int32_t airan_pusch_chain_Create(const PuschMimoConfig *config, const AiranTbPlanV1 *plan,
                                 const AiranPuschChainOptionsV1 *options, void *stream, AiranPuschChain **chain)
{
    if (config == nullptr || plan == nullptr || options == nullptr || stream == nullptr || chain == nullptr ||
        !OptionsOk(*options))
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (config->flags != 0 || g.ndmrs > 2 || (g.pad & 1u) != 0) return AIRAN_UNSUPPORTED_PROFILE;
    const AiranTbPlanV1 &p = *plan;
    const uint64_t slot_bits = static_cast<uint64_t>(g.cw_symbols) * g.qm;
    const uint32_t T = slot_bits > 0 && p.G % slot_bits == 0 ? static_cast<uint32_t>(p.G / slot_bits) : 0;
    if (p.C == 0 || p.C > airan::kCrcMaxCb || p.Z < 24 || p.K != 22 * p.Z || T == 0 ||
        options->max_slots % T != 0 ||
        static_cast<uint64_t>(p.n_low) * p.E_low + static_cast<uint64_t>(p.C - p.n_low) * p.E_high != p.G ||
        p.E_low % g.qm != 0 || p.E_high % g.qm != 0 || p.Ncb <= p.filler ||
        static_cast<uint64_t>(options->max_slots / T) * p.C > 65535 ||
        static_cast<uint64_t>(options->max_slots) * g.cw_symbols > 0xffffffffull)
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto *c = new AiranPuschChain{};
    c->cfg = *config;
    c->plan = p;
    c->opts = *options;
    c->g = g;
    c->slots_per_tb = T;
    c->crc = p.crc_tb != 0;
    c->active_rows = airan_ldpc_ActiveRows(plan);
    c->fused_ldpc = options->ldpc_mode == AIRAN_LDPC_MODE_FP16 && c->active_rows <= 8 && p.Z % 24 == 0;
    c->ls_zf = options->receiver == AIRAN_CHAIN_RX_LS_ZF;
    auto fail = [&](int32_t code) {
        airan_pusch_chain_Destroy(c);
        return code;
    };
    if (c->ls_zf && (g.nr != 1 || g.nl != 1 || config->dmrs_ports[0] != 1000)) return fail(AIRAN_UNSUPPORTED_PROFILE);
    if (options->dft_precision != AIRAN_DFT_FP32 && !airan::MfmaDftSupported(g.fft))
        return fail(AIRAN_UNSUPPORTED_PROFILE);
    if (hipGetDevice(&c->device) != hipSuccess || airan_InitDevice() != AIRAN_OK) return fail(AIRAN_LAUNCH_FAILED);
    LmmseFilterParamsV1 fp{};
    fp.abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    fp.struct_size = sizeof(fp);
    fp.max_rank = options->ce_max_rank;
    fp.tau_s = options->tau_s;
    fp.fd_hz = options->fd_hz;
    fp.scs_hz = options->scs_hz;
    fp.eig_floor = options->eig_floor;
    if (!c->ls_zf && (st = airan_channel_est_lmmse_CreateFilter(config, &fp, stream, &c->filter)) != AIRAN_OK)
        return fail(st);
    const size_t S = options->max_slots;
    const size_t NT = S / T;  // transport blocks per Enqueue at most
    const size_t plane = static_cast<size_t>(g.nsym) * g.pad;
    const size_t ncb = NT * p.C;
    const size_t gold_words = (static_cast<size_t>(T) * g.cw_symbols * g.qm + 31) / 32 + 2;
    const size_t dmrs_elems = S * g.nl * g.ndmrs * g.ref_pad;  // one DMRS reference per slot
    c->ce_ws_bytes = c->ls_zf ? 0 : airan::CeWorkspaceBytes(g, static_cast<uint32_t>(S));
    c->ldpc_ws_bytes =
        c->fused_ldpc ? 0 : airan_ldpc_decode_WorkspaceBytes(static_cast<uint32_t>(ncb), p.Z, c->active_rows);
    c->crc_ws_bytes = c->crc ? airan_crc_check_WorkspaceBytes(static_cast<uint32_t>(NT)) : 0;
    const size_t iq_elem = options->iq_format == AIRAN_IQ_COMPLEX64 ? 8 : 4;
    size_t &b = c->bytes;
    if (!Alloc(&c->dmrs_re, dmrs_elems, &b) || !Alloc(&c->dmrs_im, dmrs_elems, &b) ||
        !Alloc(&c->slot_default, S, &b) ||
        !Alloc(&c->rx_re, S * g.nr * plane, &b) || !Alloc(&c->rx_im, S * g.nr * plane, &b) ||
        !Alloc(&c->hls_re, S * g.nr * g.nl * g.ndmrs * g.pilot_pad, &b) ||
        !Alloc(&c->hls_im, S * g.nr * g.nl * g.ndmrs * g.pilot_pad, &b) ||
        !Alloc(&c->noise_ls, S * g.nr, &b) || !Alloc(&c->noise_ce, S * g.nr, &b) ||
        !Alloc(&c->ce_ws, c->ce_ws_bytes, &b) || !Alloc(&c->h_re, S * g.nr * g.nl * plane, &b) ||
        !Alloc(&c->h_im, S * g.nr * g.nl * plane, &b) || !Alloc(&c->gold, gold_words, &b) ||
        !Alloc(&c->cw, S * g.qm * g.cw_stride, &b) || !Alloc(&c->desc, ncb, &b) ||
        !Alloc(&c->lam, c->fused_ldpc ? 0 : ncb * 68 * p.Z, &b) || !Alloc(&c->ldpc_ws, c->ldpc_ws_bytes, &b) ||
        !Alloc(&c->bits, ncb * p.K, &b) || !Alloc(&c->iters, ncb, &b) || !Alloc(&c->cb_ok, ncb, &b) ||
        !Alloc(&c->tb_ok, NT, &b) || !Alloc(&c->refiner_keep, S, &b) || !Alloc(&c->crc_xpow, p.C, &b) ||
        !Alloc(&c->crc_ws, c->crc_ws_bytes, &b) ||
        !Alloc(&c->ldpc_queue, 2, &b) ||
        !Alloc(reinterpret_cast<uint8_t **>(&c->zero_iq), T * g.nr * g.samples * iq_elem, &b))
        return fail(AIRAN_RESOURCE_TOO_SMALL);
    // The decoder's block queue starts (and stays) zero.
    if (hipMemset(c->ldpc_queue, 0, 2 * sizeof(uint32_t)) != hipSuccess) return fail(AIRAN_LAUNCH_FAILED);
    if (c->crc) {
        // Per-code-block shifts of the transport-block CRC remainder.
        std::vector<uint32_t> xp(p.C);
        for (uint32_t r = 0; r < p.C; ++r) xp[r] = airan::CrcXpowHost(p, r);
        if (hipMemcpy(c->crc_xpow, xp.data(), p.C * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
            return fail(AIRAN_LAUNCH_FAILED);
        // The CRC check's tables, after the TB remainder parts in crc_ws.
        airan::CrcJob job{};
        if ((st = airan::CrcJobSetup(p, nullptr, nullptr, &job)) != AIRAN_OK) return fail(st);
        airan::LaunchCrcTables(job, p.K, CrcTables(c), nullptr);
        if (hipDeviceSynchronize() != hipSuccess || airan::LaunchStatus() != AIRAN_OK)
            return fail(AIRAN_LAUNCH_FAILED);
    }
    // config->slot_number for every slot of a plain Enqueue.
    {
        const std::vector<uint16_t> sn(S, config->slot_number);
        if (hipMemcpy(c->slot_default, sn.data(), S * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess)
            return fail(AIRAN_LAUNCH_FAILED);
    }
    // Gold bits of one transport block's codeword (the sequence restarts with
    // every TB and runs on across its slots).
    const uint32_t cinit = (static_cast<uint32_t>(config->rnti) << 15) |
                           ((static_cast<uint32_t>(config->codeword_index) & 1u) << 14) | config->data_scrambling_id;
    hipLaunchKernelGGL(GoldBitsKernel, dim3(airan::CeilDiv(gold_words, 256)), dim3(256), 0, s, c->gold,
                       static_cast<uint32_t>(gold_words), cinit);
    // Rate-match descriptors of every code block of every transport block
    // (symbol offsets into the stacked codeword buffer; a block may span a
    // slot boundary).
    std::vector<RateMatchDescriptor> desc(ncb);
    for (size_t tb = 0; tb < NT; ++tb) {
        uint64_t off = tb * T * static_cast<uint64_t>(g.cw_symbols);
        for (uint32_t r = 0; r < p.C; ++r) {
            const uint32_t e = r < p.n_low ? p.E_low : p.E_high;
            desc[tb * p.C + r] = {e, p.k0, p.Ncb, static_cast<uint32_t>(off)};
            off += e / g.qm;
        }
    }
    if (hipMemcpyAsync(c->desc, desc.data(), ncb * sizeof(RateMatchDescriptor), hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemsetAsync(c->cw, 0, S * g.qm * g.cw_stride * sizeof(int16_t), s) != hipSuccess ||
        hipMemsetAsync(c->zero_iq, 0, T * g.nr * g.samples * iq_elem, s) != hipSuccess)
        return fail(AIRAN_LAUNCH_FAILED);
    // One eager transport block (T slots of zero IQ: zero_iq holds T slots)
    // loads every per-device table before any graph capture.
    for (hipEvent_t &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(AIRAN_LAUNCH_FAILED);
    if ((st = EnqueueStages(c, c->zero_iq, c->slot_default, T, s, false)) != AIRAN_OK) return fail(st);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(AIRAN_LAUNCH_FAILED);
    c->last_slots = T;
    *chain = c;
    return AIRAN_OK;
}

void airan_pusch_chain_Destroy(AiranPuschChain *c)
{
    if (c == nullptr) return;
    DropGraph(c);
    for (hipEvent_t e : c->ev)
        if (e != nullptr) (void)hipEventDestroy(e);
    airan_channel_est_lmmse_DestroyFilter(c->filter);
    void *bufs[] = {c->dmrs_re, c->dmrs_im, c->slot_default, c->rx_re, c->rx_im, c->hls_re, c->hls_im, c->noise_ls, c->noise_ce,
                    c->ce_ws, c->h_re, c->h_im, c->gold, c->cw, c->desc, c->lam, c->ldpc_ws, c->bits,
                    c->iters, c->cb_ok, c->tb_ok, c->crc_ws, c->zero_iq, c->refiner_keep, c->crc_xpow,
                    c->ldpc_queue};
    for (void *p : bufs) Free(p);
    delete c;
}

int32_t airan_pusch_chain_EnqueueSlots(AiranPuschChain *c, const void *iq, uint32_t num_slots,
                                       const uint16_t *slot_numbers, void *stream)
{
    if (c == nullptr || iq == nullptr || slot_numbers == nullptr || stream == nullptr || num_slots == 0 ||
        num_slots > c->opts.max_slots || num_slots % c->slots_per_tb != 0)
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    c->last_slots = num_slots;
    if (c->timing || !c->opts.use_graph) return EnqueueStages(c, iq, slot_numbers, num_slots, s, c->timing);
    if (c->exec == nullptr || c->graph_iq != iq || c->graph_slot_no != slot_numbers || c->graph_slots != num_slots ||
        c->graph_stream != stream) {
        DropGraph(c);
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return AIRAN_LAUNCH_FAILED;
        const int32_t st = EnqueueStages(c, iq, slot_numbers, num_slots, s, false);
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(s, &graph);
        if (st != AIRAN_OK || end != hipSuccess || graph == nullptr) {
            if (graph) (void)hipGraphDestroy(graph);
            return st != AIRAN_OK ? st : AIRAN_LAUNCH_FAILED;
        }
        c->graph = graph;
        if (hipGraphInstantiate(&c->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            DropGraph(c);
            return AIRAN_LAUNCH_FAILED;
        }
        c->graph_iq = iq;
        c->graph_slot_no = slot_numbers;
        c->graph_slots = num_slots;
        c->graph_stream = stream;
    }
    return hipGraphLaunch(c->exec, s) == hipSuccess ? AIRAN_OK : AIRAN_LAUNCH_FAILED;
}

int32_t airan_pusch_chain_Enqueue(AiranPuschChain *c, const void *iq, uint32_t num_slots, void *stream)
{
    if (c == nullptr) return AIRAN_INVALID_ARGUMENT;
    return airan_pusch_chain_EnqueueSlots(c, iq, num_slots, c->slot_default, stream);
}

uint32_t airan_pusch_chain_SlotsPerTb(const AiranPuschChain *c) { return c == nullptr ? 0 : c->slots_per_tb; }

int32_t airan_pusch_chain_Outputs(const AiranPuschChain *c, AiranPuschChainOutputsV1 *o)
{
    if (c == nullptr || o == nullptr || o->abi_version != AIRAN_PUSCH_MIMO_ABI_VERSION ||
        o->struct_size < sizeof(AiranPuschChainOutputsV1))
        return AIRAN_INVALID_ARGUMENT;
    o->num_slots = c->last_slots;
    o->bits = c->bits;
    o->tb_ok = c->crc ? c->tb_ok : nullptr;
    o->cb_ok = c->crc ? c->cb_ok : nullptr;
    o->iterations = c->iters;
    o->rx_grid_re = c->rx_re;
    o->rx_grid_im = c->rx_im;
    o->h_ls_re = c->hls_re;
    o->h_ls_im = c->hls_im;
    o->noise_ls = c->noise_ls;
    o->noise_ce = c->ls_zf ? nullptr : c->noise_ce;
    o->h_re = c->h_re;
    o->h_im = c->h_im;
    o->cw_llr = c->cw;
    o->ldpc_llr = c->fused_ldpc ? nullptr : c->lam;
    return AIRAN_OK;
}

size_t airan_pusch_chain_DeviceBytes(const AiranPuschChain *c) { return c == nullptr ? 0 : c->bytes; }

int32_t airan_pusch_chain_SetRefiner(AiranPuschChain *c, AiranCeRefiner *refiner, float snr_min_db,
                                     float snr_max_db)
{
    if (c == nullptr || (refiner != nullptr && !(snr_min_db <= snr_max_db))) return AIRAN_INVALID_ARGUMENT;
    if (refiner != nullptr && c->ls_zf) return AIRAN_UNSUPPORTED_PROFILE;
    DropGraph(c);
    c->refiner = refiner;
    c->refiner_snr_min = snr_min_db;
    c->refiner_snr_max = snr_max_db;
    return AIRAN_OK;
}

int32_t airan_pusch_chain_SetTiming(AiranPuschChain *c, uint32_t enabled)
{
    if (c == nullptr) return AIRAN_INVALID_ARGUMENT;
    c->timing = enabled != 0;
    return AIRAN_OK;
}

int32_t airan_pusch_chain_StageTimes(const AiranPuschChain *c, float *ms, uint32_t count)
{
    if (c == nullptr || ms == nullptr || count < AIRAN_CHAIN_STAGES || !c->timing) return AIRAN_INVALID_ARGUMENT;
    if (hipEventSynchronize(c->ev[AIRAN_CHAIN_STAGES]) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    for (uint32_t i = 0; i < AIRAN_CHAIN_STAGES; ++i)
        if (hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    return AIRAN_OK;
}

}  // extern "C"
