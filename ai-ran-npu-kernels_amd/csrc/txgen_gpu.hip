// txgen_gpu.hip -- the synthetic transmitter and fading channel on the
// device (include/airan_pusch_tx.h, airan_pusch_txgpu_*): the bit-level
// stages of txgen.cpp for a batch of transport blocks, bit for bit, then a
// tapped-delay-line channel, OFDM modulation and AWGN straight to IQ.
//
//   TbCrcKernel / CbCrcKernel  CRC attach.  A remainder is linear in the
//       message bits, so each thread reduces a 64-bit chunk of the message
//       (counted from its end: chunk k holds bits n - 64 (k + 1) .. n - 64 k
//       - 1) to rho_k mod g and contributes rho_k x^(64 k + L) mod g (a
//       host table); a wave XOR-reduces, then one atomicXor per wave.
//   EncodeKernel   one workgroup per code block: info bits (TB bits, CB CRC,
//       filler zeros) in LDS, BG1 core parities (lambda sums, p0 through the
//       unmatched column-22 shift, then the core rows with one unknown in a
//       host-derived order), extension parities row by row, then rate
//       matching from k0 skipping filler and the bit interleaver, written to
//       the TB's coded-bit stream.
//   ModulateKernel  one thread per 32 codeword symbols: Gold scrambling
//       (one jump-ahead per thread, then 32 bits at a time), QAM mapping with
//       the reference mapper's polarity, layer mapping onto the data REs.
//   DmrsGridKernel  the DMRS symbols of every layer: the slot's DMRS on the
//       layer's comb (FD-OCC sign on ports 1001/1003), zero on the other.
//   ChannelOfdmKernel  one workgroup per (slot, Rx, symbol): y = sum_l H x_l
//       with H = sum_taps g e^(-j 2 pi f tau) on the used subcarriers, a
//       4096-point (or smaller) radix-2 IFFT in LDS, the cyclic prefix and
//       AWGN from a counter-based generator, written as IQ.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "airan_ldpc_bg1.h"
#include "airan_pusch_tx.h"
#include "common.h"
#include "crc.h"
#include "gold.h"

using airan::DevGeom;
using airan::Geom;

namespace {

constexpr uint32_t kPolyA = 0x864CFBu;   // CRC24A
constexpr uint32_t kPolyB = 0x800063u;   // CRC24B
constexpr uint32_t kPoly16 = 0x1021u;    // CRC16
constexpr uint32_t kChunkBits = 64;
constexpr uint32_t kMaxZ = 384;
constexpr uint32_t kMaxFft = 4096;

struct TxPlan {
    uint32_t Z, C, K, kprime, per_cb, crc_cb, crc_tb, tbs, filler, ncb, q0, G, e_low, e_high, n_low;
    uint32_t s22;                 // shift of column 22 left after summing the four core rows
    uint32_t nsteps;              // core solve steps
    uint32_t step_row[4], step_col[4];
    uint32_t tb_poly, tb_len;     // TB CRC polynomial
};

// Host: x^e mod g (len-bit residues, poly without the x^len term).
uint32_t HostMulMod(uint32_t a, uint32_t b, uint32_t poly, uint32_t len)
{
    const uint32_t mask = (1u << len) - 1u;
    uint32_t r = 0;
    for (int i = static_cast<int>(len) - 1; i >= 0; --i) {
        const uint32_t top = (r >> (len - 1)) & 1u;
        r = (r << 1) & mask;
        if (top) r ^= poly;
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

uint32_t HostXpow(uint64_t e, uint32_t poly, uint32_t len)
{
    uint32_t r = 1u, base = 2u;
    while (e != 0) {
        if (e & 1u) r = HostMulMod(r, base, poly, len);
        base = HostMulMod(base, base, poly, len);
        e >>= 1;
    }
    return r;
}

// Table x^(64 k + len) mod g for k < chunks.
std::vector<uint32_t> ChunkPowers(uint32_t chunks, uint32_t poly, uint32_t len)
{
    std::vector<uint32_t> t(chunks);
    uint32_t p = HostXpow(len, poly, len);
    const uint32_t step = HostXpow(kChunkBits, poly, len);
    for (uint32_t k = 0; k < chunks; ++k) {
        t[k] = p;
        p = HostMulMod(p, step, poly, len);
    }
    return t;
}

// Derives the core-parity schedule of the BG1 encoder (the same order
// txgen.cpp's encoder solves in).
bool CoreSchedule(uint32_t z, TxPlan *tp)
{
    std::vector<uint32_t> s22;
    for (int r = 0; r < 4; ++r)
        for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e)
            if (AIRAN_BG1_COL[e] == 22) {
                const uint32_t s = AIRAN_BG1_SHIFT[e] % z;
                bool found = false;
                for (size_t k = 0; k < s22.size(); ++k)
                    if (s22[k] == s) {
                        s22.erase(s22.begin() + static_cast<long>(k));
                        found = true;
                        break;
                    }
                if (!found) s22.push_back(s);
            }
    if (s22.size() != 1) return false;
    tp->s22 = s22[0];
    bool known[26] = {};
    for (int c = 0; c <= 22; ++c) known[c] = true;
    tp->nsteps = 0;
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
            if (AIRAN_BG1_SHIFT[unknown] % z != 0 || tp->nsteps >= 4) return false;
            tp->step_row[tp->nsteps] = static_cast<uint32_t>(r);
            tp->step_col[tp->nsteps] = AIRAN_BG1_COL[unknown];
            ++tp->nsteps;
            known[AIRAN_BG1_COL[unknown]] = true;
        }
    return known[23] && known[24] && known[25];
}

// ------------------------------------------------------------------ CRC
// Residue of `nbits` message bits (first bit the highest power) mod g.
__device__ inline uint32_t ShiftIn(uint32_t r, uint32_t bit, uint32_t poly, uint32_t len)
{
    const uint32_t hi = (r >> (len - 1)) & 1u;
    r = ((r << 1) & ((1u << len) - 1u)) ^ bit;
    return hi ? r ^ poly : r;
}

__device__ inline uint32_t WaveXor(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v ^= __shfl_xor(v, o, 64);
    return v;
}

// TB CRC: thread (chunk k of TB tb); crc[tb] ^= rho_k x^(64 k + L).
__global__ void __launch_bounds__(256) TbCrcKernel(const uint8_t *__restrict__ payload, TxPlan p,
                                                   const uint32_t *__restrict__ pw, uint32_t chunks,
                                                   uint32_t *__restrict__ crc)
{
    const uint32_t tb = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t contrib = 0;
    if (k < chunks) {
        const uint8_t *m = payload + static_cast<size_t>(tb) * p.tbs;
        const int64_t end = static_cast<int64_t>(p.tbs) - static_cast<int64_t>(kChunkBits) * k;
        const int64_t beg = end - static_cast<int64_t>(kChunkBits);
        uint32_t r = 0;
        for (int64_t j = beg < 0 ? 0 : beg; j < end; ++j) r = ShiftIn(r, m[j] & 1u, p.tb_poly, p.tb_len);
        contrib = airan::CrcMulMod(r, pw[k], p.tb_poly, p.tb_len);
    }
    contrib = WaveXor(contrib);
    if ((threadIdx.x & 63u) == 0 && contrib != 0) atomicXor(&crc[tb], contrib);
}

// Bit idx of TB tb's CRC-extended bit string b = payload || TB CRC.
__device__ inline uint32_t TbBit(const uint8_t *__restrict__ payload, const uint32_t *__restrict__ tb_crc,
                                 const TxPlan &p, uint32_t tb, uint32_t idx)
{
    if (idx < p.tbs) return payload[static_cast<size_t>(tb) * p.tbs + idx] & 1u;
    return (tb_crc[tb] >> (p.crc_tb - 1u - (idx - p.tbs))) & 1u;
}

// CB CRC24B: thread (chunk k of code block cb = tb * C + r).
__global__ void __launch_bounds__(256) CbCrcKernel(const uint8_t *__restrict__ payload,
                                                   const uint32_t *__restrict__ tb_crc, TxPlan p,
                                                   const uint32_t *__restrict__ pw, uint32_t chunks,
                                                   uint32_t *__restrict__ crc)
{
    const uint32_t cb = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tb = cb / p.C, r = cb % p.C;
    uint32_t contrib = 0;
    if (k < chunks) {
        const int64_t end = static_cast<int64_t>(p.per_cb) - static_cast<int64_t>(kChunkBits) * k;
        const int64_t beg = end - static_cast<int64_t>(kChunkBits);
        uint32_t res = 0;
        for (int64_t j = beg < 0 ? 0 : beg; j < end; ++j)
            res = ShiftIn(res, TbBit(payload, tb_crc, p, tb, r * p.per_cb + static_cast<uint32_t>(j)), kPolyB, 24);
        contrib = airan::CrcMulMod(res, pw[k], kPolyB, 24);
    }
    contrib = WaveXor(contrib);
    if ((threadIdx.x & 63u) == 0 && contrib != 0) atomicXor(&crc[cb], contrib);
}

__global__ void __launch_bounds__(256) ZeroWordsKernel(uint32_t *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// -------------------------------------------------------------- encode
__device__ inline uint32_t Rot(uint32_t i, uint32_t s, uint32_t z)
{
    const uint32_t v = i + s;
    return v >= z ? v - z : v;
}

// XOR over the edges of row r whose column satisfies `take`, at check i.
template <typename Take>
__device__ inline uint32_t RowXor(const uint8_t *cw, int r, uint32_t i, uint32_t z, Take take)
{
    uint32_t v = 0;
    for (int e = AIRAN_BG1_ROW_START[r]; e < AIRAN_BG1_ROW_START[r + 1]; ++e) {
        const uint32_t c = AIRAN_BG1_COL[e];
        if (take(c)) v ^= cw[c * z + Rot(i, AIRAN_BG1_SHIFT[e] % z, z)];
    }
    return v;
}

__global__ void __launch_bounds__(384) EncodeKernel(const uint8_t *__restrict__ payload,
                                                    const uint32_t *__restrict__ tb_crc,
                                                    const uint32_t *__restrict__ cb_crc, TxPlan p, uint32_t qm,
                                                    uint8_t *__restrict__ coded)
{
    __shared__ uint8_t cw[68 * kMaxZ];
    __shared__ uint8_t lam[4][kMaxZ];
    const uint32_t cb = blockIdx.x, tb = cb / p.C, r = cb % p.C;
    const uint32_t z = p.Z, t = threadIdx.x;
    const uint32_t crc = cb_crc[cb];
    for (uint32_t k = t; k < 68u * z; k += blockDim.x) {
        uint32_t v = 0;
        if (k < p.per_cb)
            v = TbBit(payload, tb_crc, p, tb, r * p.per_cb + k);
        else if (k < p.kprime)
            v = (crc >> (23u - (k - p.per_cb))) & 1u;
        cw[k] = static_cast<uint8_t>(v);
    }
    __syncthreads();
    if (t < z)
        for (int row = 0; row < 4; ++row) lam[row][t] = static_cast<uint8_t>(RowXor(cw, row, t, z, [](uint32_t c) { return c < 22u; }));
    __syncthreads();
    if (t < z) cw[22u * z + Rot(t, p.s22, z)] = lam[0][t] ^ lam[1][t] ^ lam[2][t] ^ lam[3][t];
    __syncthreads();
    for (uint32_t s = 0; s < p.nsteps; ++s) {
        const uint32_t row = p.step_row[s], u = p.step_col[s];
        if (t < z)
            cw[u * z + t] = lam[row][t] ^ static_cast<uint8_t>(RowXor(
                                              cw, static_cast<int>(row), t, z,
                                              [u](uint32_t c) { return c >= 22u && c < 26u && c != u; }));
        __syncthreads();
    }
    for (int row = 4; row < AIRAN_BG1_ROWS; ++row) {
        const uint32_t own = 22u + static_cast<uint32_t>(row);
        if (t < z) cw[own * z + t] = static_cast<uint8_t>(RowXor(cw, row, t, z, [own](uint32_t c) { return c != own; }));
        __syncthreads();
    }
    // Rate matching (rv from k0, filler skipped) and the bit interleaver.
    const uint32_t E = r < p.n_low ? p.e_low : p.e_high;
    const uint64_t out = r < p.n_low ? static_cast<uint64_t>(r) * p.e_low
                                     : static_cast<uint64_t>(p.n_low) * p.e_low + static_cast<uint64_t>(r - p.n_low) * p.e_high;
    const uint32_t eq = E / qm, nf = p.ncb - p.filler, fstart = p.kprime - 2u * z;
    uint8_t *dst = coded + static_cast<size_t>(tb) * p.G + out;
    for (uint32_t k = t; k < E; k += blockDim.x) {
        const uint32_t q = static_cast<uint32_t>((static_cast<uint64_t>(p.q0) + k) % nf);
        const uint32_t idx = q < fstart ? q : q + p.filler;
        const uint32_t i = k / eq, j = k - i * eq;
        dst[static_cast<size_t>(j) * qm + i] = cw[2u * z + idx];
    }
}

// ------------------------------------------------------------ modulate
// Axis level d0 (2^(m-1) - d1 (2^(m-2) - ... d_{m-1})), d_i = 2 c_i - 1.
__device__ inline float AxisLevel(uint32_t bits, uint32_t m)
{
    float e = 1.0f;
    for (int i = static_cast<int>(m) - 1; i >= 1; --i)
        e = static_cast<float>(1u << (m - static_cast<uint32_t>(i))) - (2.0f * static_cast<float>((bits >> i) & 1u) - 1.0f) * e;
    return (2.0f * static_cast<float>(bits & 1u) - 1.0f) * e;
}

struct ModArgs {
    DevGeom g;
    uint32_t G, T, cinit, num_tbs;
    float norm;
};

// Thread: 32 codeword symbols of one TB (32 Qm bits = Qm Gold words).
__global__ void __launch_bounds__(256) ModulateKernel(const uint8_t *__restrict__ coded, ModArgs a,
                                                      float2 *__restrict__ grid, uint8_t *__restrict__ coded_out)
{
    const DevGeom &g = a.g;
    const uint32_t per_tb = a.T * g.cw_symbols;
    const uint32_t groups = (per_tb + 31u) / 32u;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= groups * a.num_tbs) return;
    const uint32_t tb = gid / groups, grp = gid % groups;
    const uint32_t qm = g.qm, m = qm / 2u;
    uint32_t s1, s2;
    airan::GoldStateAt(static_cast<uint64_t>(grp) * 32u * qm + 1600u, a.cinit, &s1, &s2);
    uint32_t words[8];
    for (uint32_t w = 0; w < qm; ++w) words[w] = airan::GoldNext32(&s1, &s2);
    const uint8_t *src = coded + static_cast<size_t>(tb) * a.G;
    const size_t slot_plane = static_cast<size_t>(g.nl) * airan::kSymbols * g.used;
    for (uint32_t k = 0; k < 32u; ++k) {
        const uint32_t sym = grp * 32u + k;
        if (sym >= per_tb) break;
        uint32_t ib = 0, qb = 0;
        for (uint32_t b = 0; b < qm; ++b) {
            const uint32_t bit_local = k * qm + b;
            const uint64_t n = static_cast<uint64_t>(sym) * qm + b;
            const uint32_t bit = (src[n] ^ (words[bit_local >> 5] >> (bit_local & 31u))) & 1u;
            if (coded_out != nullptr) coded_out[static_cast<size_t>(tb) * a.G + n] = static_cast<uint8_t>(bit);
            if (b & 1u)
                qb |= bit << (b >> 1);
            else
                ib |= bit << (b >> 1);
        }
        const uint32_t s = sym / g.cw_symbols, i = sym - s * g.cw_symbols;
        const uint32_t l = i % g.nl, re = i / g.nl;
        const uint32_t ds = re / g.used, sc = re - ds * g.used;
        const size_t o = (static_cast<size_t>(tb) * a.T + s) * slot_plane +
                         (static_cast<size_t>(l) * airan::kSymbols + g.data_sym[ds]) * g.used + sc;
        grid[o] = make_float2(AxisLevel(ib, m) * a.norm, AxisLevel(qb, m) * a.norm);
    }
}

struct DmrsArgs {
    DevGeom g;
    uint32_t slots, rb_start, n_id, n_scid, default_slot;
    uint32_t wf_odd[4];
};

// Thread: 16 DMRS REs (one Gold word) of (slot, layer, DMRS symbol), plus
// the zeros of the other comb at the same 16 positions.
__global__ void __launch_bounds__(256) DmrsGridKernel(DmrsArgs a, const uint16_t *__restrict__ slot_numbers,
                                                      float2 *__restrict__ grid)
{
    const DevGeom &g = a.g;
    const uint32_t blocks = (g.n_dmrs_re + 15u) / 16u;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = a.slots * g.nl * g.ndmrs * blocks;
    if (gid >= total) return;
    const uint32_t blk = gid % blocks, rest = gid / blocks;
    const uint32_t d = rest % g.ndmrs, rest2 = rest / g.ndmrs;
    const uint32_t l = rest2 % g.nl, s = rest2 / g.nl;
    const uint64_t slot_number = slot_numbers != nullptr ? slot_numbers[s] : a.default_slot;
    const uint64_t term = (1ull << 17) * (14ull * slot_number + g.dmrs_sym[d] + 1ull) * (2ull * a.n_id + 1ull);
    const uint32_t cdmrs = static_cast<uint32_t>((term + 2ull * a.n_id + a.n_scid) & 0x7fffffffull);
    const uint32_t word = airan::GoldWordAt(cdmrs, 12ull * a.rb_start + 32ull * blk);
    const float h = 0.70710678118654752f;
    const size_t base = (static_cast<size_t>(s) * g.nl + l) * airan::kSymbols * g.used +
                        static_cast<size_t>(g.dmrs_sym[d]) * g.used;
    for (uint32_t j = 0; j < 16u; ++j) {
        const uint32_t mm = blk * 16u + j;
        if (mm >= g.n_dmrs_re) break;
        float vr = ((word >> (2u * j)) & 1u) ? -h : h, vi = ((word >> (2u * j + 1u)) & 1u) ? -h : h;
        if (a.wf_odd[l] && (mm & 1u)) {
            vr = -vr;
            vi = -vi;
        }
        const uint32_t sc = g.comb[l] + 2u * mm, other = (1u - g.comb[l]) + 2u * mm;
        if (sc < g.used) grid[base + sc] = make_float2(vr, vi);
        if (other < g.used) grid[base + other] = make_float2(0.0f, 0.0f);
    }
}

// ------------------------------------------------- channel + OFDM + noise
struct ChanArgs {
    uint32_t nr, nl, used, fft, log2n, taps, samples, iq_float;
    uint32_t cp[14], start[14];
    float scs, sigma, inv_sqrt_n;
    uint64_t seed;
};

__device__ inline uint64_t Mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Two independent N(0, 1) values (Box-Muller) from a counter.
__device__ inline float2 Gauss2(uint64_t key)
{
    const uint64_t r = Mix64(key);
    const float u1 = (static_cast<float>(r >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = static_cast<float>((r >> 16) & 0xffffffull) * (1.0f / 16777216.0f);
    const float rad = sqrtf(-2.0f * logf(u1));
    float sn, cs;
    sincospif(2.0f * u2, &sn, &cs);
    return make_float2(rad * cs, rad * sn);
}

// Workgroup (slot, rx, symbol).  grid: complex64 [slots][L][14][used];
// gains complex64 [slots][NR][L][taps][14]; delays [slots][taps].
__global__ void __launch_bounds__(256) ChannelOfdmKernel(const float2 *__restrict__ grid,
                                                         const float2 *__restrict__ gains,
                                                         const float *__restrict__ delays,
                                                         const float2 *__restrict__ twiddle, ChanArgs a,
                                                         void *__restrict__ iq)
{
    __shared__ float2 buf[kMaxFft];
    __shared__ float2 gs[4][32];  // this symbol's tap gains per layer
    __shared__ double tau[32];
    const uint32_t sym = blockIdx.x % 14u, rest = blockIdx.x / 14u;
    const uint32_t rx = rest % a.nr, slot = rest / a.nr;
    const uint32_t t = threadIdx.x, n = a.fft;
    for (uint32_t k = t; k < a.nl * a.taps; k += blockDim.x) {
        const uint32_t l = k / a.taps, tp = k % a.taps;
        gs[l][tp] = gains[((((static_cast<size_t>(slot) * a.nr + rx) * a.nl + l) * a.taps + tp) * 14u) + sym];
    }
    for (uint32_t k = t; k < a.taps; k += blockDim.x) tau[k] = delays[static_cast<size_t>(slot) * a.taps + k];
    __syncthreads();
    const uint32_t half = a.used / 2u;
    const size_t gplane = static_cast<size_t>(14) * a.used;
    const float2 *gx = grid + static_cast<size_t>(slot) * a.nl * gplane + static_cast<size_t>(sym) * a.used;
    for (uint32_t k = t; k < n; k += blockDim.x) {
        // Bin k carries subcarrier u = (k + used/2) mod N when u < used.
        const uint32_t u = (k + half) & (n - 1u);
        float2 y = make_float2(0.0f, 0.0f);
        if (u < a.used) {
            const double f = (static_cast<double>(u) - static_cast<double>(half)) * static_cast<double>(a.scs);
            float hr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, hi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t tp = 0; tp < a.taps; ++tp) {
                // e^(-j 2 pi f tau), the phase reduced to (-1/2, 1/2] cycles in fp64.
                const double cyc = f * tau[tp];
                const float frac = static_cast<float>(cyc - rint(cyc));
                float sn, cs;
                sincospif(2.0f * frac, &sn, &cs);
#pragma unroll
                for (uint32_t l = 0; l < 4; ++l) {
                    if (l < a.nl) {
                        const float2 gv = gs[l][tp];
                        hr[l] += gv.x * cs + gv.y * sn;
                        hi[l] += gv.y * cs - gv.x * sn;
                    }
                }
            }
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l) {
                if (l < a.nl) {
                    const float2 x = gx[static_cast<size_t>(l) * gplane + u];
                    y.x += hr[l] * x.x - hi[l] * x.y;
                    y.y += hr[l] * x.y + hi[l] * x.x;
                }
            }
        }
        buf[__brev(k) >> (32u - a.log2n)] = y;
    }
    __syncthreads();
    // Radix-2 decimation-in-time inverse DFT: e^(+j 2 pi k n / N).
    for (uint32_t h = 1; h < n; h <<= 1) {
        const uint32_t stride = n / (2u * h);
        for (uint32_t b = t; b < n / 2u; b += blockDim.x) {
            const uint32_t grp = b / h, pos = b - grp * h;
            const uint32_t i0 = grp * 2u * h + pos, i1 = i0 + h;
            const float2 w = twiddle[pos * stride];
            const float2 v = buf[i1], uu = buf[i0];
            const float2 vw = make_float2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            buf[i0] = make_float2(uu.x + vw.x, uu.y + vw.y);
            buf[i1] = make_float2(uu.x - vw.x, uu.y - vw.y);
        }
        __syncthreads();
    }
    const uint32_t cp = a.cp[sym], first = a.start[sym] - cp;
    const size_t row = (static_cast<size_t>(slot) * a.nr + rx) * a.samples;
    for (uint32_t p = t; p < cp + n; p += blockDim.x) {
        const float2 v = buf[p < cp ? n - cp + p : p - cp];
        const uint64_t pos = row + first + p;
        const float2 z = Gauss2(a.seed ^ (pos * 0x632be59bd9b4e019ull));
        const float re = v.x * a.inv_sqrt_n + a.sigma * z.x, im = v.y * a.inv_sqrt_n + a.sigma * z.y;
        if (a.iq_float) {
            static_cast<float2 *>(iq)[pos] = make_float2(re, im);
        } else {
            const float qr = fminf(fmaxf(rintf(re * 256.0f), -32768.0f), 32767.0f);
            const float qi = fminf(fmaxf(rintf(im * 256.0f), -32768.0f), 32767.0f);
            static_cast<short2 *>(iq)[pos] = make_short2(static_cast<short>(qr), static_cast<short>(qi));
        }
    }
}

void FreeDev(void *p)
{
    if (p != nullptr) (void)hipFree(p);
}

}  // namespace

struct AiranPuschTxGpu {
    PuschMimoConfig cfg{};
    Geom g{};
    TxPlan tp{};
    uint32_t max_tbs = 0, T = 1;
    uint32_t tb_chunks = 0, cb_chunks = 0;
    uint32_t *tb_crc = nullptr, *cb_crc = nullptr, *pow_tb = nullptr, *pow_cb = nullptr;
    uint8_t *coded = nullptr;
    float2 *twiddle = nullptr;
};

extern "C" {

int32_t airan_pusch_txgpu_Create(const PuschMimoConfig *config, const AiranTbPlanV1 *plan, uint32_t max_tbs,
                                 uint32_t slots_per_tb, AiranPuschTxGpu **tx)
{
    if (config == nullptr || plan == nullptr || tx == nullptr || max_tbs == 0 || slots_per_tb == 0)
        return AIRAN_INVALID_ARGUMENT;
    *tx = nullptr;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    const AiranTbPlanV1 &p = *plan;
    if (p.C == 0 || p.Z == 0 || p.Z > kMaxZ || p.K != 22u * p.Z || p.Kprime > p.K || p.filler != p.K - p.Kprime ||
        p.Kprime < 2u * p.Z || p.Ncb == 0 || p.Ncb > 66u * p.Z || p.Ncb <= p.filler || p.k0 >= p.Ncb ||
        p.E_low % g.qm != 0 || p.E_high % g.qm != 0 || p.n_low > p.C ||
        static_cast<uint64_t>(p.G) != static_cast<uint64_t>(slots_per_tb) * g.cw_symbols * g.qm ||
        static_cast<uint64_t>(p.n_low) * p.E_low + static_cast<uint64_t>(p.C - p.n_low) * p.E_high != p.G ||
        p.tbs + p.crc_tb + static_cast<uint64_t>(p.C) * p.crc_cb != static_cast<uint64_t>(p.C) * p.Kprime ||
        (p.crc_tb != 0 && p.crc_tb != 16 && p.crc_tb != 24) || (p.crc_cb != 0 && p.crc_cb != 24))
        return AIRAN_UNSUPPORTED_PROFILE;
    auto *t = new AiranPuschTxGpu{};
    t->cfg = *config;
    t->g = g;
    t->max_tbs = max_tbs;
    t->T = slots_per_tb;
    TxPlan &tp = t->tp;
    tp.Z = p.Z;
    tp.C = p.C;
    tp.K = p.K;
    tp.kprime = p.Kprime;
    tp.crc_cb = p.crc_cb;
    tp.crc_tb = p.crc_tb;
    tp.per_cb = p.Kprime - p.crc_cb;
    tp.tbs = p.tbs;
    tp.filler = p.filler;
    tp.ncb = p.Ncb;
    const uint32_t fstart = p.Kprime - 2u * p.Z;
    tp.q0 = p.k0 < fstart ? p.k0 : (p.k0 >= fstart + p.filler ? p.k0 - p.filler : fstart);
    tp.G = p.G;
    tp.e_low = p.E_low;
    tp.e_high = p.E_high;
    tp.n_low = p.n_low;
    tp.tb_poly = p.crc_tb == 16 ? kPoly16 : kPolyA;
    tp.tb_len = p.crc_tb == 16 ? 16u : 24u;
    auto fail = [&](int32_t code) {
        airan_pusch_txgpu_Destroy(t);
        return code;
    };
    if (!CoreSchedule(p.Z, &tp)) return fail(AIRAN_UNSUPPORTED_PROFILE);
    t->tb_chunks = (p.tbs + kChunkBits - 1) / kChunkBits;
    t->cb_chunks = (tp.per_cb + kChunkBits - 1) / kChunkBits;
    const std::vector<uint32_t> ptb = ChunkPowers(t->tb_chunks, tp.tb_poly, tp.tb_len);
    const std::vector<uint32_t> pcb = ChunkPowers(t->cb_chunks, kPolyB, 24);
    std::vector<float2> tw(g.fft / 2);
    for (uint32_t k = 0; k < g.fft / 2; ++k) {
        const double ph = 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(g.fft);
        tw[k] = make_float2(static_cast<float>(std::cos(ph)), static_cast<float>(std::sin(ph)));
    }
    auto alloc = [](auto **ptr, size_t bytes) { return hipMalloc(reinterpret_cast<void **>(ptr), bytes) == hipSuccess; };
    if (!alloc(&t->tb_crc, sizeof(uint32_t) * max_tbs) || !alloc(&t->cb_crc, sizeof(uint32_t) * max_tbs * p.C) ||
        !alloc(&t->pow_tb, sizeof(uint32_t) * t->tb_chunks) || !alloc(&t->pow_cb, sizeof(uint32_t) * t->cb_chunks) ||
        !alloc(&t->coded, static_cast<size_t>(max_tbs) * p.G) || !alloc(&t->twiddle, sizeof(float2) * tw.size()))
        return fail(AIRAN_RESOURCE_TOO_SMALL);
    if (hipMemcpy(t->pow_tb, ptb.data(), sizeof(uint32_t) * ptb.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->pow_cb, pcb.data(), sizeof(uint32_t) * pcb.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->twiddle, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(AIRAN_LAUNCH_FAILED);
    if ((st = airan::InitGoldTables()) != AIRAN_OK) return fail(st);
    *tx = t;
    return AIRAN_OK;
}

void airan_pusch_txgpu_Destroy(AiranPuschTxGpu *t)
{
    if (t == nullptr) return;
    FreeDev(t->tb_crc);
    FreeDev(t->cb_crc);
    FreeDev(t->pow_tb);
    FreeDev(t->pow_cb);
    FreeDev(t->coded);
    FreeDev(t->twiddle);
    delete t;
}

int32_t airan_pusch_txgpu_EncodeTbs(AiranPuschTxGpu *t, const uint8_t *payload, uint32_t num_tbs,
                                    const uint16_t *slot_numbers, float *grid, uint8_t *coded_bits, void *stream)
{
    if (t == nullptr || payload == nullptr || grid == nullptr || stream == nullptr || num_tbs == 0 ||
        num_tbs > t->max_tbs)
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Geom &g = t->g;
    const TxPlan &tp = t->tp;
    const uint32_t ncb = num_tbs * tp.C;
    hipLaunchKernelGGL(ZeroWordsKernel, dim3(airan::CeilDiv(num_tbs, 256)), dim3(256), 0, s, t->tb_crc, num_tbs);
    hipLaunchKernelGGL(ZeroWordsKernel, dim3(airan::CeilDiv(ncb, 256)), dim3(256), 0, s, t->cb_crc, ncb);
    if (tp.crc_tb)
        hipLaunchKernelGGL(TbCrcKernel, dim3(airan::CeilDiv(t->tb_chunks, 256), num_tbs), dim3(256), 0, s, payload,
                           tp, t->pow_tb, t->tb_chunks, t->tb_crc);
    if (tp.crc_cb)
        hipLaunchKernelGGL(CbCrcKernel, dim3(airan::CeilDiv(t->cb_chunks, 256), ncb), dim3(256), 0, s, payload,
                           t->tb_crc, tp, t->pow_cb, t->cb_chunks, t->cb_crc);
    hipLaunchKernelGGL(EncodeKernel, dim3(ncb), dim3(384), 0, s, payload, t->tb_crc, t->cb_crc, tp, g.qm, t->coded);
    ModArgs ma{};
    ma.g = airan::ToDev(g);
    ma.G = tp.G;
    ma.T = t->T;
    ma.cinit = (static_cast<uint32_t>(t->cfg.rnti) << 15) | ((static_cast<uint32_t>(t->cfg.codeword_index) & 1u) << 14) |
               t->cfg.data_scrambling_id;
    ma.num_tbs = num_tbs;
    ma.norm = g.qm == 2 ? 1.0f / std::sqrt(2.0f)
                        : (g.qm == 4 ? 1.0f / std::sqrt(10.0f) : (g.qm == 6 ? 1.0f / std::sqrt(42.0f) : 1.0f / std::sqrt(170.0f)));
    const uint32_t groups = airan::CeilDiv(static_cast<uint64_t>(t->T) * g.cw_symbols, 32) * num_tbs;
    hipLaunchKernelGGL(ModulateKernel, dim3(airan::CeilDiv(groups, 256)), dim3(256), 0, s, t->coded, ma,
                       reinterpret_cast<float2 *>(grid), coded_bits);
    DmrsArgs da{};
    da.g = ma.g;
    da.slots = num_tbs * t->T;
    da.rb_start = t->cfg.rb_start;
    da.n_id = t->cfg.dmrs_scrambling_id;
    da.n_scid = t->cfg.n_scid;
    da.default_slot = t->cfg.slot_number;
    for (uint32_t l = 0; l < g.nl; ++l) da.wf_odd[l] = t->cfg.dmrs_ports[l] == 1001 || t->cfg.dmrs_ports[l] == 1003;
    const uint32_t dthreads = da.slots * g.nl * g.ndmrs * airan::CeilDiv(g.n_dmrs_re, 16);
    hipLaunchKernelGGL(DmrsGridKernel, dim3(airan::CeilDiv(dthreads, 256)), dim3(256), 0, s, da, slot_numbers,
                       reinterpret_cast<float2 *>(grid));
    return airan::LaunchStatus();
}

int32_t airan_pusch_txgpu_Channel(AiranPuschTxGpu *t, const float *grid, uint32_t slots, const float *gains,
                                  const float *delays, uint32_t taps, float scs_hz, float noise_var, uint64_t seed,
                                  uint32_t iq_format, void *iq, void *stream)
{
    if (t == nullptr || grid == nullptr || gains == nullptr || delays == nullptr || iq == nullptr ||
        stream == nullptr || slots == 0 || taps == 0 || taps > 32 || !(scs_hz > 0.0f) || !(noise_var >= 0.0f) ||
        (iq_format != AIRAN_IQ_INT16_Q8_8 && iq_format != AIRAN_IQ_COMPLEX64))
        return AIRAN_INVALID_ARGUMENT;
    const Geom &g = t->g;
    if (g.fft > kMaxFft) return AIRAN_UNSUPPORTED_PROFILE;
    ChanArgs a{};
    a.nr = g.nr;
    a.nl = g.nl;
    a.used = g.used;
    a.fft = g.fft;
    a.log2n = 0;
    while ((1u << a.log2n) < g.fft) ++a.log2n;
    a.taps = taps;
    a.samples = g.samples;
    a.iq_float = iq_format == AIRAN_IQ_COMPLEX64;
    for (uint32_t k = 0; k < 14; ++k) {
        a.cp[k] = k == 0 ? g.cp_long : g.cp_short;
        a.start[k] = g.sym_start[k];
    }
    a.scs = scs_hz;
    a.sigma = std::sqrt(noise_var * 0.5f);
    a.inv_sqrt_n = 1.0f / std::sqrt(static_cast<float>(g.fft));
    a.seed = seed;
    hipLaunchKernelGGL(ChannelOfdmKernel, dim3(slots * g.nr * 14u), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const float2 *>(grid), reinterpret_cast<const float2 *>(gains), delays,
                       t->twiddle, a, iq);
    return airan::LaunchStatus();
}

}  // extern "C"
