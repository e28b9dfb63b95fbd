// op_refiner.hip -- learned channel-estimation refiner (include/airan_ce_refiner.h;
// no counterpart in the reference).
//
// Tensors are NHWC bf16: [slots][14 symbols][width subcarriers][channels],
// channels contiguous (32 or 64 = one 64/128-byte pixel vector).
//
// Convolution (forward, and the data gradient as a convolution with the
// flipped, transposed weights): implicit GEMM on MFMA
// v_mfma_f32_32x32x16_bf16 with the weights as A (32 output channels x 16
// input channels, in registers for the whole persistent kernel) and 32
// pixels as B.  A workgroup (4 waves) takes work items of 7 output symbol
// rows x 64 pixels (x 64 channels; 128 x 32 for the last layer) and walks
// the rows through a 4-row LDS ring of input rows with their halo, so each
// input row is read once per item; rows are XOR-swizzled in 16-byte chunks
// so the operand reads are free of bank conflicts.  Each layer may dilate
// its taps along the subcarrier axis (taps at subcarrier offsets -dil, 0,
// +dil; symbol offsets stay -1, 0, +1), which widens the network's view of
// the band: dilations 1, 2, 4, 8, 16, 1 see 65 subcarriers.  Weight
// gradients are fp32 VALU outer products over LDS
// tiles, reduced across workgroups in a fixed order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "airan_ce_refiner.h"
#include "airan_pusch_rx.h"
#include "common.h"
#include "launch.h"

using airan::Geom;

namespace {

using bf16 = __bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kLayers = 6;
constexpr int kSym = 14;
constexpr int kCh = 32;
constexpr int kHid = 64;
constexpr int kCin[kLayers] = {32, 64, 64, 64, 64, 64};
constexpr int kCout[kLayers] = {64, 64, 64, 64, 64, 32};

constexpr size_t ParamOffset(int layer)
{
    size_t off = 0;
    for (int i = 0; i < layer; ++i) off += size_t(9) * kCout[i] * kCin[i] + kCout[i];
    return off;
}
constexpr size_t kParams = ParamOffset(kLayers);
static_assert(kParams == AIRAN_REFINER_PARAMS, "parameter count");
constexpr size_t WeightCount(int layer) { return size_t(9) * kCout[layer] * kCin[layer]; }

// ---------------------------------------------------------------- RNG
__device__ inline uint64_t Mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Two independent N(0, 1) values (Box-Muller) from a counter.
__device__ inline float2 Gauss2(uint64_t seed, uint64_t ctr)
{
    const uint64_t r = Mix64(seed ^ Mix64(ctr));
    const float u1 = (static_cast<float>(r >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = static_cast<float>((r >> 16) & 0xffffffu) * (1.0f / 16777216.0f);
    const float rad = sqrtf(-2.0f * logf(u1));
    float sn, cs;
    sincospif(2.0f * u2, &sn, &cs);
    return make_float2(rad * cs, rad * sn);
}

// ----------------------------------------------------------- synthesis
struct SynthArgs {
    const float2 *gains;   // [slots][NR][L][taps][14]
    const float *tau;      // [slots][taps]
    const float *snr_db;   // [slots]
    const __half *dmrs_re; // [L][D][ref_pad]
    const __half *dmrs_im;
    float *target;         // [slots][14][used][32]
    __half *rx_re;         // [slots][NR][14][pad] (DMRS symbols written)
    __half *rx_im;
    uint32_t taps, used, pad, nr, nl, ndmrs, ref_pad;
    uint32_t dmrs_sym[4], comb[4];
    float scs;
    uint64_t seed;
};

// True channel H[r][l][k][u] = sum_t g[r][l][t][k] exp(-j 2 pi (u - used/2) scs tau_t)
// (the tapped-delay-line frequency response, subcarrier u at baseband
// (u - used/2) scs) and the received DMRS symbols
// Y[r][k][u] = sum_l H[r][l][k][u] x_l(u) + n.  Grid (used / 256, slots).
__global__ void __launch_bounds__(256) SynthKernel(SynthArgs a)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = blockIdx.y;
    if (u >= a.used) return;
    float2 st[32];
    for (uint32_t t = 0; t < a.taps; ++t) {
        // Phase in cycles, reduced in double before the fp32 sincos.
        const double cyc = (static_cast<double>(u) - 0.5 * a.used) * a.scs * a.tau[s * a.taps + t];
        const float fr = static_cast<float>(cyc - floor(cyc));
        float sn, cs;
        sincospif(-2.0f * fr, &sn, &cs);
        st[t] = make_float2(cs, sn);
    }
    const float sigma = sqrtf(0.5f * exp10f(-0.1f * a.snr_db[s]));
    float *tg = a.target + (static_cast<size_t>(s) * kSym * a.used + u) * kCh;
    for (uint32_t k = 0; k < kSym; ++k) {
        int d = -1;
        for (uint32_t j = 0; j < a.ndmrs; ++j)
            if (a.dmrs_sym[j] == k) d = static_cast<int>(j);
        float2 y[4] = {};
        for (uint32_t r = 0; r < a.nr; ++r)
            for (uint32_t l = 0; l < a.nl; ++l) {
                const float2 *g = a.gains + ((((static_cast<size_t>(s) * a.nr + r) * a.nl + l) * a.taps) * kSym) + k;
                float hr = 0.0f, hi = 0.0f;
                for (uint32_t t = 0; t < a.taps; ++t) {
                    const float2 gv = g[static_cast<size_t>(t) * kSym];
                    hr += gv.x * st[t].x - gv.y * st[t].y;
                    hi += gv.x * st[t].y + gv.y * st[t].x;
                }
                const uint32_t c = 2 * (r * a.nl + l);
                tg[static_cast<size_t>(k) * a.used * kCh + c] = hr;
                tg[static_cast<size_t>(k) * a.used * kCh + c + 1] = hi;
                if (d >= 0 && u >= a.comb[l] && ((u - a.comb[l]) & 1u) == 0) {
                    const size_t m = (static_cast<size_t>(l) * a.ndmrs + d) * a.ref_pad + (u - a.comb[l]) / 2;
                    const float xr = __half2float(a.dmrs_re[m]), xi = __half2float(a.dmrs_im[m]);
                    y[r].x += hr * xr - hi * xi;
                    y[r].y += hr * xi + hi * xr;
                }
            }
        if (d < 0) continue;
        for (uint32_t r = 0; r < a.nr; ++r) {
            const float2 n = Gauss2(a.seed, ((static_cast<uint64_t>(s) * 4 + r) * 4 + d) * 65536ull + u);
            const size_t o = ((static_cast<size_t>(s) * a.nr + r) * kSym + k) * a.pad + u;
            a.rx_re[o] = __float2half_rn(y[r].x + sigma * n.x);
            a.rx_im[o] = __float2half_rn(y[r].y + sigma * n.y);
        }
    }
}

// ----------------------------------------------------------------- input
struct InputArgs {
    const __half *h_re;  // [slots][NR][L][D][pilot_pad]
    const __half *h_im;
    bf16 *x;             // [slots][14][used][32]
    float *x32;          // the same in fp32 (the residual the network adds to)
    uint32_t used, nr, nl, ndmrs, pilot_pad;
    uint32_t dmrs_sym[4], comb[4], count[4];
    int32_t partner[4];  // FD-OCC partner layer, -1 for a layer alone on its comb
    float sign[4];       // +1: FD-OCC signs (+,+) (ports 1000 / 1002), -1: (+,-)
};

// Linear interpolation at subcarrier u of samples v(p) at first + step p
// (p < cnt), constant beyond the outermost samples.
template <typename F>
__device__ inline float2 Interp(F v, float first, float step, uint32_t cnt, uint32_t u)
{
    const float fp = (static_cast<float>(u) - first) / step;
    if (fp <= 0.0f) return v(0);
    if (fp >= static_cast<float>(cnt - 1)) return v(cnt - 1);
    const uint32_t p0 = static_cast<uint32_t>(fp);
    const float w = fp - static_cast<float>(p0);
    const float2 a = v(p0), b = v(p0 + 1);
    return make_float2(a.x + w * (b.x - a.x), a.y + w * (b.y - a.y));
}

// Network input: per link (rx, layer) and DMRS symbol, the least-squares
// estimate linearly interpolated in frequency, then linearly interpolated /
// extrapolated in time through the DMRS symbols (constant for one).  A layer
// alone on its comb interpolates h_ls between its pilots (comb + 2p).  An
// FD-OCC pair (layers a: signs +,+ and b: +,-) is first split into its exact
// observations z(comb + 4p) = h_ls_a[p] + h_ls_b[p] = (h_a + h_b)(comb + 4p)
// and z(comb + 4p + 2) = h_ls_a[p] - h_ls_b[p] = (h_a - h_b)(comb + 4p + 2);
// g+ = h_a + h_b and g- = h_a - h_b are interpolated on their own combs and
// h_a,b = (g+ +- g-) / 2, which avoids the pair mean's flat-channel
// assumption.  Grid (used / 256, slots).
__global__ void __launch_bounds__(256) InputKernel(InputArgs a)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = blockIdx.y;
    if (u >= a.used) return;
    bf16 *xo = a.x + (static_cast<size_t>(s) * kSym * a.used + u) * kCh;
    float *x32o = a.x32 + (static_cast<size_t>(s) * kSym * a.used + u) * kCh;
    for (uint32_t r = 0; r < a.nr; ++r)
        for (uint32_t l = 0; l < a.nl; ++l) {
            float2 hd[4];
            const uint32_t cnt = a.count[l];
            for (uint32_t d = 0; d < a.ndmrs; ++d) {
                auto row = [&](uint32_t layer) {
                    return (((static_cast<size_t>(s) * a.nr + r) * a.nl + layer) * a.ndmrs + d) * a.pilot_pad;
                };
                const size_t bl = row(l);
                auto hl = [&](size_t base, uint32_t p) {
                    return make_float2(__half2float(a.h_re[base + p]), __half2float(a.h_im[base + p]));
                };
                if (a.partner[l] < 0) {
                    hd[d] = Interp([&](uint32_t p) { return hl(bl, p); }, static_cast<float>(a.comb[l]), 2.0f, cnt,
                                   u);
                } else {
                    const size_t bq = row(static_cast<uint32_t>(a.partner[l]));
                    // h_ls of the (+,+) layer and of the (+,-) layer.
                    const size_t bp = a.sign[l] > 0.0f ? bl : bq, bm = a.sign[l] > 0.0f ? bq : bl;
                    const float2 gp = Interp(
                        [&](uint32_t p) {
                            const float2 x = hl(bp, p), y = hl(bm, p);
                            return make_float2(x.x + y.x, x.y + y.y);
                        },
                        static_cast<float>(a.comb[l]), 4.0f, cnt, u);
                    const float2 gm = Interp(
                        [&](uint32_t p) {
                            const float2 x = hl(bp, p), y = hl(bm, p);
                            return make_float2(x.x - y.x, x.y - y.y);
                        },
                        static_cast<float>(a.comb[l] + 2), 4.0f, cnt, u);
                    hd[d] = make_float2(0.5f * (gp.x + a.sign[l] * gm.x), 0.5f * (gp.y + a.sign[l] * gm.y));
                }
            }
            const uint32_t c = 2 * (r * a.nl + l);
            for (uint32_t k = 0; k < kSym; ++k) {
                float2 v = hd[0];
                if (a.ndmrs >= 2) {
                    const float t0 = static_cast<float>(a.dmrs_sym[0]);
                    const float t1 = static_cast<float>(a.dmrs_sym[a.ndmrs - 1]);
                    const float wt = (static_cast<float>(k) - t0) / (t1 - t0);
                    const float2 v1 = hd[a.ndmrs - 1];
                    v = make_float2(hd[0].x + wt * (v1.x - hd[0].x), hd[0].y + wt * (v1.y - hd[0].y));
                }
                const size_t o = static_cast<size_t>(k) * a.used * kCh + c;
                xo[o] = static_cast<bf16>(v.x);
                xo[o + 1] = static_cast<bf16>(v.y);
                x32o[o] = v.x;
                x32o[o + 1] = v.y;
            }
        }
}

// Network input from a channel estimate grid h fp16 [slots][16 links][14][pad]
// (the LMMSE estimator's output; link r L + l): x[s][k][u][2 link + {re, im}]
// = bf16(h) and, with X32, x32 = h exactly (fp16 -> fp32).  With zero_tail
// (inference, where the network's last layer then writes the refined grid in
// place over subcarriers 0 .. used - 1) it also zeroes the grid's subcarriers
// used .. pad - 1 of the slots it will write (keep[s] != 0, or all with keep
// null).  Block: 64 subcarriers of one symbol row, transposed through LDS:
// thread t first reads 8 subcarriers of channel t / 8 (one 16-byte load),
// then writes 8 channels of pixel t / 4 (16 bytes; four threads per 64-byte
// pixel, so each wave writes 1 KB contiguous).  Grid (ceil(pad / 64), 14,
// slots).
template <bool X32>
__global__ void __launch_bounds__(256) EstimateInputKernel(__half *__restrict__ h_re, __half *__restrict__ h_im,
                                                           bf16 *__restrict__ x, float *__restrict__ x32,
                                                           uint32_t used, uint32_t pad, bool zero_tail,
                                                           const uint8_t *__restrict__ keep)
{
    constexpr int kPx = 64, kStride = kCh + 1;  // LDS pixel row: 32 channels + 1 (fp32 words)
    __shared__ float tile[kPx * kStride];
    const uint32_t u0 = blockIdx.x * kPx;
    const uint32_t k = blockIdx.y, s = blockIdx.z;
    const int t = static_cast<int>(threadIdx.x);
    {
        const int c = t >> 3, q = (t & 7) * 8;
        __half *plane = (c & 1 ? h_im : h_re) + ((static_cast<size_t>(s) * (kCh / 2) + (c >> 1)) * kSym + k) * pad;
        const uint32_t u = u0 + q;
        float v[8];
        if (u + 8 <= used && pad % 8 == 0) {
            const uint4 w = *reinterpret_cast<const uint4 *>(plane + u);
            const __half *hv = reinterpret_cast<const __half *>(&w);
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = __half2float(hv[j]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = u + j < used ? __half2float(plane[u + j]) : 0.0f;
            if (zero_tail && (keep == nullptr || keep[s] != 0))
                for (uint32_t j = max(u, used); j < min(u + 8, pad); ++j) plane[j] = __float2half_rn(0.0f);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) tile[(q + j) * kStride + c] = v[j];
    }
    __syncthreads();
    const int px = t >> 2, ch = (t & 3) * 8;
    if (u0 + px >= used) return;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = tile[px * kStride + ch + j];
    const size_t o = ((static_cast<size_t>(s) * kSym + k) * used + u0 + px) * kCh + ch;
    bf16x8 b;
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = static_cast<bf16>(v[j]);
    *reinterpret_cast<bf16x8 *>(x + o) = b;
    if constexpr (X32) {
        *reinterpret_cast<float4 *>(x32 + o) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(x32 + o + 4) = make_float4(v[4], v[5], v[6], v[7]);
    }
}

// x32 = float(x) for an input given in bf16 (airan_ce_refiner_Gradients).
__global__ void __launch_bounds__(256) WidenKernel(const bf16 *__restrict__ x, float *__restrict__ x32, size_t n)
{
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        x32[i] = static_cast<float>(x[i]);
}

// Residual scale per slot: sqrt of the mean over Rx of the slot's noise
// variance (fp16 [slots][NR]), or 1 without a noise input.
__global__ void __launch_bounds__(64) NoiseScaleKernel(const __half *__restrict__ noise, uint32_t nr, uint32_t slots,
                                                       float *__restrict__ scale)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    if (noise == nullptr) {
        scale[s] = 1.0f;
        return;
    }
    float acc = 0.0f;
    for (uint32_t r = 0; r < nr; ++r) acc += __half2float(noise[static_cast<size_t>(s) * nr + r]);
    scale[s] = sqrtf(fmaxf(acc / static_cast<float>(nr), 0.0f));
}

// ----------------------------------------------------------------- conv
enum { kEpiRelu = 0, kEpiResid = 1, kEpiMask = 2, kEpiGrid = 3 };

struct ConvArgs {
    const bf16 *in;      // [S][14][W][CIN]
    const bf16 *wt;      // [9][COUT][CIN]
    const float *bias;   // [COUT] (kEpiRelu, kEpiResid, kEpiGrid)
    const bf16 *aux;     // kEpiMask: activation [S][14][W][COUT]
    const float *aux32;  // kEpiResid: X in fp32 [S][14][W][COUT]
    const float *scale;  // kEpiResid, kEpiGrid: residual scale per slot [S]
    bf16 *out;           // kEpiRelu, kEpiMask: [S][14][W][COUT]
    float *out_f32;      // kEpiResid: [S][14][W][COUT]
    __half *g_re;        // kEpiGrid: residual base and output, fp16 [S][COUT / 2][14][pad]
    __half *g_im;
    const uint8_t *keep; // kEpiGrid: slots written (null: all)
    uint32_t pad;        // kEpiGrid: subcarriers per grid row (>= W; zero beyond W)
    uint32_t S, W;
    int dil;             // subcarrier dilation of the taps, 1..HALO
};

constexpr int kMaxDil = 16;

// Position of 16-byte chunk c of pixel p in its LDS row (XOR swizzle).  LDS
// has 64 four-byte banks and serves a 16-byte read 16 lanes at a time, so a
// read of one chunk of 16 consecutive pixels is conflict-free when their
// 16-byte units (pixel p: CH units) are distinct modulo 16; the unit of
// chunk c of pixel p depends on p only through p mod 16, injectively:
// CH = 8: 8 (p & 1) + (c ^ ((p >> 1) & 7)); CH = 4: 4 (p & 3) + (c ^ ((p >> 2) & 3)).
template <int CH>
__device__ inline int ChunkPos(int p, int c)
{
    return CH == 8 ? c ^ ((p >> 1) & 7) : c ^ ((p >> 2) & 3);
}

// One 16-byte-per-lane buffer load of [base, base + bytes) at byte offset
// voff straight into LDS at lds + 16 lane (lds wave-uniform); an offset at or
// beyond `bytes` loads zero (the range check; kOutside with bytes < 2^31
// stays outside for every dword of the 16 bytes, no wrap).
constexpr uint32_t kOutside = 0x80000000u;
__device__ inline void BufferToLds16(const void *base, uint32_t bytes, uint32_t voff, void *lds)
{
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, static_cast<int>(bytes), 0x00020000);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (__attribute__((address_space(3))) void *)lds, 16, voff, 0, 0, 0);
}

// The same load into registers (zero beyond `bytes`).
__device__ inline uint4 BufferLoad16(const void *base, uint32_t bytes, uint32_t voff)
{
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, static_cast<int>(bytes), 0x00020000);
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, 0, 0));
}

// out[p][co] = epilogue(acc), acc = bias[co] + sum_{tap, ci} in[p + off(tap)][ci]
// wt[tap][co][ci] (the bias is the accumulator's initial value; none for kEpiMask):
//   kEpiRelu   bf16(max(acc, 0))
//   kEpiResid  scale[slot] acc + aux32   (fp32)
//   kEpiMask   bf16(aux > 0 ? acc : 0)   (gradient through the ReLU)
//   kEpiGrid   the residual output straight into the channel grid:
//              g[co / 2] = fp16(scale[slot] acc + float(g[co / 2]))
//              (re for even co, im for odd; in place; the same value as
//              kEpiResid with aux32 = float(g) rounded to fp16) for
//              subcarriers 0 .. W - 1, slots with keep[slot] == 0 untouched.
//
// Work item: one TILE-pixel column of half a slot's rows (7 output symbol
// rows; TILE = 64 for 64 output channels, 128 for 32).  The item's 9 input
// rows pass through a ring of 4 LDS rows: step j computes output row h0 + j
// from input rows j .. j + 2 (of h0 - 1 ..) while row j + 3 is on its way,
// so each input row is read from memory once per item.  The row copies are
// buffer loads over the item's slot (descriptor in scalar registers, one
// 32-bit offset per copy); positions outside the grid get an offset beyond
// the buffer's end, which the hardware's range check returns as zero.  A
// step loads row j + 3 into registers before its MFMAs and writes it to LDS
// after them, so the loads' latency hides behind the MFMAs and only the LDS
// writes need waiting for at the next barrier.  The item's first three rows
// are copied global -> LDS directly and waited for before the first step.
//
// MFMA v_mfma_f32_32x32x16_bf16 with the weights as A (32 output channels x
// 16 input channels, in registers for the whole kernel) and the pixels as B
// (16 input channels x 32 pixels, from LDS).  Lane layout: A row = lane % 32,
// its 8 K values 8 (lane / 32) .. + 7; B column = lane % 32, the same K;
// C column (pixel) = lane % 32, row (output channel) of accumulator i =
// 8 (i / 4) + 4 (lane / 32) + i % 4 -- each lane holds 4 consecutive output
// channels of one pixel per group of 4 accumulators.
// HALO: the largest dilation the LDS rows hold (1, 2, 4, 8 or kMaxDil).
template <int CIN, int COUT, int EPI, int HALO>
__global__ void __launch_bounds__(256, 2) ConvKernel(ConvArgs a)
{
    constexpr int CO_BLOCKS = COUT / 32;
    constexpr int PX_WAVES = 4 / CO_BLOCKS;
    constexpr int TILE = 32 * PX_WAVES;
    constexpr int KS = CIN / 16;
    constexpr int CH = CIN / 8;                            // 16-byte chunks per pixel
    constexpr int ROW = ((TILE + 2 * HALO) * CH + 63) / 64 * 64;  // chunks per LDS row (whole waves)
    constexpr int NK = (ROW + 255) / 256;                  // copies per thread per row
    constexpr int HROWS = kSym / 2;                        // output rows per item
    constexpr int RING = 4;                                // LDS input rows
    static_assert(CH == 4 || CH == 8, "32 or 64 input channels");
    __shared__ uint4 ring[RING * ROW];
    __shared__ float4 bias4[COUT / 4];
    // bf16 epilogues: per wave a 32-pixel x 32-channel tile, chunk g (channels
    // 8 g .. + 7) of pixel p at 34 g + p (the 2-chunk pad makes both the
    // pixel-major writes and the chunk-major reads conflict-free).
    constexpr bool kBf16Out = EPI == kEpiRelu || EPI == kEpiMask;
    __shared__ uint4 escr[kBf16Out ? 4 * 4 * 34 : 1];
    const int dil = HALO == 1 ? 1 : a.dil;
    const int tw = TILE + 2 * dil;
    const int W = static_cast<int>(a.W);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wco = wave % CO_BLOCKS, wpx = wave / CO_BLOCKS;
    const int kh8 = lane >> 5;
    const int pl = wpx * 32 + (lane & 31);                 // the lane's pixel in the tile
    const int co0 = wco * 32 + 4 * kh8;                    // + 8 g + e: accumulator 4 g + e
    bf16x8 wf[9][KS];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            wf[tap][ks] = *reinterpret_cast<const bf16x8 *>(
                a.wt + (static_cast<size_t>(tap) * COUT + wco * 32 + (lane & 31)) * CIN + ks * 16 + 8 * kh8);
    if (threadIdx.x < COUT / 4)
        bias4[threadIdx.x] = EPI == kEpiMask ? make_float4(0, 0, 0, 0)
                                             : reinterpret_cast<const float4 *>(a.bias)[threadIdx.x];
    // LDS byte offsets of the lane's B fragments within a ring row.
    uint32_t off[3][KS];
#pragma unroll
    for (int kw = 0; kw < 3; ++kw) {
        const int q = pl + kw * dil;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) off[kw][ks] = 16u * (q * CH + ChunkPos<CH>(q, 2 * ks + kh8));
    }
    const uint32_t wtiles = (a.W + TILE - 1) / TILE;
    const uint32_t items = a.S * 2 * wtiles;
    for (uint32_t t = blockIdx.x; t < items; t += gridDim.x) {
        const uint32_t wt_i = t % wtiles, sh = t / wtiles;
        const int h0 = static_cast<int>(sh % 2) * HROWS;
        const size_t s = sh / 2;
        const int w0 = static_cast<int>(wt_i) * TILE;
        if (EPI == kEpiGrid && a.keep != nullptr && a.keep[s] == 0) continue;
        const uint32_t slot_bytes = kSym * a.W * CIN * 2;
        const bf16 *in_s = a.in + s * kSym * a.W * CIN;
        // Byte offset of copy k's chunk in its input row, or kOutside (outside
        // the grid's subcarriers).
        uint32_t col[NK];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int e = k * 256 + static_cast<int>(threadIdx.x);
            const int px = e / CH;
            const int ww = w0 - dil + px;
            col[k] = px < tw && ww >= 0 && ww < W ? 16u * (static_cast<uint32_t>(ww) * CH + ChunkPos<CH>(px, e % CH))
                                                  : kOutside;
        }
        // Input row h0 - 1 + m into ring row m % 4 (chunk e of the row:
        // pixel e / CH, position e % CH holding chunk ChunkPos(pixel, e % CH)).
        auto voff = [&](int m, int k) {
            const int hh = h0 - 1 + m;
            const uint32_t row_off = hh >= 0 && hh < kSym ? static_cast<uint32_t>(hh) * a.W * CIN * 2 : kOutside;
            return col[k] == kOutside || row_off == kOutside ? kOutside : col[k] + row_off;
        };
        auto copy_row = [&](int m, uint4 *dst) {
#pragma unroll
            for (int k = 0; k < NK; ++k)
                if (k * 256 + wave * 64 < ROW)  // wave-uniform: the row's last waves
                    BufferToLds16(in_s, slot_bytes, voff(m, k), dst + k * 256 + wave * 64);
        };
        uint4 stage[NK];
        auto load_row = [&](int m) {
#pragma unroll
            for (int k = 0; k < NK; ++k)
                if (k * 256 + wave * 64 < ROW) stage[k] = BufferLoad16(in_s, slot_bytes, voff(m, k));
        };
        auto store_row = [&](uint4 *dst) {
#pragma unroll
            for (int k = 0; k < NK; ++k)
                if (k * 256 + wave * 64 < ROW) dst[k * 256 + threadIdx.x] = stage[k];
        };
        __syncthreads();  // the previous item's last step has read its rows
#pragma unroll
        for (int m = 0; m < 3; ++m) copy_row(m, ring + m * ROW);
        // No direct copies are issued inside the row loop, so the compiler
        // must know these are done (it would otherwise wait for them before
        // every operand read).
        __builtin_amdgcn_s_waitcnt(0xF70);
#pragma unroll 1
        for (int j = 0; j < HROWS; ++j) {
            const int h = h0 + j;
            const int w = w0 + pl;
            const size_t px_out = (s * kSym + h) * a.W + w;
            // Rows j .. j + 1 were waited for at earlier steps; row j + 2
            // (this wave's LDS writes of it) now.  The barrier then makes
            // every wave's part visible and frees ring row (j + 3) % 4 (read
            // at step j - 1).  A bare barrier: a fence would also wait for the
            // epilogue's global stores.
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            __builtin_amdgcn_s_barrier();
            if (j + 3 <= HROWS + 1) load_row(j + 3);
            // Epilogue inputs, loaded before this step's copy is issued so
            // that waiting for them does not wait for the copy.
            // bf16 outputs leave transposed: lane -> pixel (lane >> 2) + 16 r of
            // the wave's 32, channels wco 32 + 8 (lane & 3) .. + 7 (16 bytes).
            const int tg = lane & 3;
            uint4 mask[2];
            float4 res32[4];
            if constexpr (EPI == kEpiMask) {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int wt = w0 + wpx * 32 + (lane >> 2) + 16 * r;
                    mask[r] = wt < W ? *reinterpret_cast<const uint4 *>(
                                           a.aux + ((s * kSym + h) * a.W + wt) * COUT + wco * 32 + 8 * tg)
                                     : make_uint4(0, 0, 0, 0);
                }
            }
            if constexpr (EPI == kEpiResid) {
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    res32[g] = w < W ? *reinterpret_cast<const float4 *>(a.aux32 + px_out * COUT + co0 + 8 * g)
                                     : make_float4(0, 0, 0, 0);
            }
            // Grid element of accumulator i = 4 g + e: channel co0 + 8 g + e,
            // i.e. link co0 / 2 + 4 g + e / 2 (re for even e, im for odd).
            const uint32_t gstride = kSym * a.pad;
            const uint32_t gbase = ((static_cast<uint32_t>(s) * (COUT / 2) + (co0 >> 1)) * kSym + h) * a.pad;
            f32x16 acc;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 b = bias4[(co0 >> 2) + 2 * g];
                acc[4 * g] = b.x;
                acc[4 * g + 1] = b.y;
                acc[4 * g + 2] = b.z;
                acc[4 * g + 3] = b.w;
            }
            // 9 KS operand reads, kept kDepth ahead of the MFMA that uses them
            // (an LDS read takes several MFMA durations to return; 2 for the
            // 32-channel layer, whose registers also hold a 5-copy staged row).
            {
                constexpr int kN = 9 * KS, kDepth = COUT == 64 ? 4 : 2;
                auto operand = [&](int n) {
                    const int kh = n / (3 * KS), kw = (n / KS) % 3, ks = n % KS;
                    return *reinterpret_cast<const bf16x8 *>(reinterpret_cast<const char *>(ring + ((j + kh) & 3) * ROW) +
                                                             off[kw][ks]);
                };
                bf16x8 bq[kDepth];
#pragma unroll
                for (int n = 0; n < kDepth; ++n) bq[n] = operand(n);
#pragma unroll
                for (int n = 0; n < kN; ++n) {
                    const bf16x8 b = bq[n % kDepth];
                    if (n + kDepth < kN) bq[n % kDepth] = operand(n + kDepth);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[n / KS][n % KS], b, acc, 0, 0, 0);
                    // Keep the order as written (the scheduler would sink
                    // each read next to its MFMA).
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if (j + 3 <= HROWS + 1) store_row(ring + ((j + 3) & 3) * ROW);
            if constexpr (EPI == kEpiGrid) {
                if (w >= W) continue;
                const float sc = a.scale[s];
                // The residual loads in two groups of 8 (one wait each), after
                // the MFMAs: their operand registers are free by now.
#pragma unroll
                for (int h8 = 0; h8 < 16; h8 += 8) {
                    __half gin[8];
#pragma unroll
                    for (int i = h8; i < h8 + 8; ++i)
                        gin[i - h8] =
                            (i & 1 ? a.g_im : a.g_re)[gbase + (4 * (i >> 2) + ((i & 3) >> 1)) * gstride + w];
#pragma unroll
                    for (int i = h8; i < h8 + 8; ++i) {
                        // Rounded to fp32 first, then to fp16 (as the fp32
                        // route does): the empty asm keeps the compiler from
                        // fusing the two roundings into one mixed-precision fma.
                        float v = fmaf(sc, acc[i], __half2float(gin[i - h8]));
                        asm volatile("" : "+v"(v));
                        (i & 1 ? a.g_im : a.g_re)[gbase + (4 * (i >> 2) + ((i & 3) >> 1)) * gstride + w] =
                            __float2half_rn(v);
                    }
                }
                continue;
            }
            if constexpr (kBf16Out) {
                uint4 *scr = escr + wave * (4 * 34);
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    uint32_t v[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float x = acc[4 * g + e];
                        v[e] = __builtin_bit_cast(uint16_t, static_cast<bf16>(EPI == kEpiRelu ? fmaxf(x, 0.0f) : x));
                    }
                    *reinterpret_cast<uint2 *>(reinterpret_cast<char *>(scr + 34 * g + (lane & 31)) + 8 * kh8) =
                        make_uint2(v[0] | v[1] << 16, v[2] | v[3] << 16);
                }
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int pxw = (lane >> 2) + 16 * r;
                    uint4 v = scr[34 * tg + pxw];
                    if constexpr (EPI == kEpiMask) {
                        // Gradient through the ReLU: zero where the activation is not positive.
                        uint32_t *vw = reinterpret_cast<uint32_t *>(&v);
                        const uint32_t *mw = reinterpret_cast<const uint32_t *>(&mask[r]);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            uint32_t keepw = 0;
#pragma unroll
                            for (int hb = 0; hb < 2; ++hb) {
                                const uint16_t mb = static_cast<uint16_t>(mw[q] >> (16 * hb));
                                if (static_cast<float>(__builtin_bit_cast(bf16, mb)) > 0.0f) keepw |= 0xffffu << (16 * hb);
                            }
                            vw[q] &= keepw;
                        }
                    }
                    const int wt = w0 + wpx * 32 + pxw;
                    if (wt < W)
                        *reinterpret_cast<uint4 *>(a.out + ((s * kSym + h) * a.W + wt) * COUT + wco * 32 + 8 * tg) = v;
                }
                continue;
            }
            if (w >= W) continue;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const size_t o = px_out * COUT + co0 + 8 * g;
                const float sc = a.scale[s];
                *reinterpret_cast<float4 *>(a.out_f32 + o) =
                    make_float4(fmaf(sc, acc[4 * g], res32[g].x), fmaf(sc, acc[4 * g + 1], res32[g].y),
                                fmaf(sc, acc[4 * g + 2], res32[g].z), fmaf(sc, acc[4 * g + 3], res32[g].w));
            }
        }
    }
}

// ---------------------------------------------------------- weight gradient
// Partial sums of one workgroup over its tiles (64 pixels of one symbol
// row): part[tap][co][ci] = sum_p dz[p][co] a[p + off(tap)][ci] and
// part[9 COUT CIN + co] = sum_p dz[p][co].  Thread: co = t % COUT and the
// channel group g = t / COUT of CIN / (256 / COUT) input channels.
template <int CIN, int COUT, int HALO>
__global__ void __launch_bounds__(256) WgradKernel(const bf16 *__restrict__ act, const bf16 *__restrict__ dz,
                                                   float *__restrict__ part, uint32_t S, uint32_t W, int dil_arg)
{
    constexpr int TILE = 64;
    constexpr int G = 256 / COUT;
    constexpr int CG = CIN / G;
    __shared__ float sa[3][TILE + 2 * HALO][CIN];
    const int dil = HALO == 1 ? 1 : dil_arg;
    const int tw = TILE + 2 * dil;
    __shared__ float sd[TILE][COUT];
    const int t = threadIdx.x;
    const int co = t % COUT, g = t / COUT;
    float acc[9][CG];
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[k][c] = 0.0f;
    float db = 0.0f;
    const uint32_t wtiles = (W + TILE - 1) / TILE;
    const uint32_t ntiles = S * kSym * wtiles;
    for (uint32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const uint32_t wt_i = ti % wtiles, sh = ti / wtiles;
        const uint32_t h = sh % kSym;
        const size_t s = sh / kSym;
        const int w0 = static_cast<int>(wt_i) * TILE;
        for (int e = t; e < 3 * tw * CIN; e += 256) {
            const int c = e % CIN, rp = e / CIN, p = rp % tw, r = rp / tw;
            const int hh = static_cast<int>(h) + r - 1, ww = w0 + p - dil;
            float v = 0.0f;
            if (hh >= 0 && hh < kSym && ww >= 0 && ww < static_cast<int>(W))
                v = static_cast<float>(act[((s * kSym + hh) * W + ww) * CIN + c]);
            sa[r][p][c] = v;
        }
        for (int e = t; e < TILE * COUT; e += 256) {
            const int c = e % COUT, p = e / COUT;
            const int ww = w0 + p;
            sd[p][c] = ww < static_cast<int>(W) ? static_cast<float>(dz[((s * kSym + h) * W + ww) * COUT + c]) : 0.0f;
        }
        __syncthreads();
        for (int p = 0; p < TILE; ++p) {
            const float d = sd[p][co];
            db += d;
#pragma unroll
            for (int kh = 0; kh < 3; ++kh)
#pragma unroll
                for (int kw = 0; kw < 3; ++kw) {
                    const float *ar = &sa[kh][p + kw * dil][g * CG];
#pragma unroll
                    for (int c = 0; c < CG; ++c) acc[kh * 3 + kw][c] += d * ar[c];
                }
        }
        __syncthreads();
    }
    float *o = part + static_cast<size_t>(blockIdx.x) * (9 * COUT * CIN + COUT);
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int c = 0; c < CG; ++c) o[(static_cast<size_t>(k) * COUT + co) * CIN + g * CG + c] = acc[k][c];
    if (g == 0) o[9 * COUT * CIN + co] = db;
}

// grad[i] = sum_b part[b][i] in block order.
__global__ void __launch_bounds__(256) ReduceKernel(const float *__restrict__ part, uint32_t nb, uint32_t n,
                                                    float *__restrict__ grad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t b = 0; b < nb; ++b) s += part[static_cast<size_t>(b) * n + i];
    grad[i] = s;
}

// ------------------------------------------------------------------ loss
// dz = bf16(2 w_s (out - target) / N); per-block sums of w_s (out - target)^2,
// w_s the weight of the element's slot (1 without weights; the training
// weights have mean 1 over the batch).
__global__ void __launch_bounds__(256) LossKernel(const float *__restrict__ out, const float *__restrict__ tgt,
                                                  const float *__restrict__ slot_w,
                                                  const float *__restrict__ res_scale, size_t per_slot,
                                                  bf16 *__restrict__ dz, double *__restrict__ part, size_t n)
{
    __shared__ double red[4];
    double acc = 0.0;
    const float scale = static_cast<float>(2.0 / static_cast<double>(n));
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float w = slot_w != nullptr ? slot_w[i / per_slot] : 1.0f;
        const float e = out[i] - tgt[i];
        acc += static_cast<double>(w) * e * e;
        // out = aux + scale_s z: the gradient of the last pre-activation z.
        dz[i] = static_cast<bf16>(w * e * scale * res_scale[i / per_slot]);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(64) LossFinishKernel(const double *__restrict__ part, uint32_t nb, size_t n,
                                                       float *__restrict__ loss)
{
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nb; ++b) s += part[b];
    *loss = static_cast<float>(s / static_cast<double>(n));
}

// ------------------------------------------------------------- optimiser
__global__ void __launch_bounds__(256) AdamKernel(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ g, uint32_t n, float lr, float bc1,
                                                  float bc2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

// bf16 copies of every layer's weights: forward [tap][co][ci] and the
// data-gradient form [8 - tap][ci][co] (flipped taps, transposed channels).
struct PackArgs {
    uint32_t off[kLayers], cin[kLayers], cout[kLayers];
};
__global__ void __launch_bounds__(256) PackKernel(const float *__restrict__ p, bf16 *__restrict__ wf,
                                                  bf16 *__restrict__ wb, PackArgs a)
{
    const uint32_t layer = blockIdx.y;
    const uint32_t cin = a.cin[layer], cout = a.cout[layer];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * cin * cout) return;
    const uint32_t ci = i % cin, co = (i / cin) % cout, tap = i / (cin * cout);
    const float v = p[a.off[layer] + i];
    wf[a.off[layer] + i] = static_cast<bf16>(v);
    wb[a.off[layer] + (static_cast<size_t>(8 - tap) * cin + ci) * cout + co] = static_cast<bf16>(v);
}

// ----------------------------------------------------------------- output
// Network output [slots][14][used][32] -> channel grid fp16
// [slots][NR][L][14][pad] (zero beyond used), slots with keep[s] != 0.
__global__ void __launch_bounds__(256) OutToGridKernel(const float *__restrict__ out, const uint8_t *__restrict__ keep,
                                                       __half *__restrict__ h_re, __half *__restrict__ h_im,
                                                       uint32_t used, uint32_t pad, uint32_t nr, uint32_t nl)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = blockIdx.y, s = blockIdx.z;
    if (u >= pad || (keep != nullptr && keep[s] == 0)) return;
    const float *o = out + ((static_cast<size_t>(s) * kSym + k) * used + u) * kCh;
    for (uint32_t r = 0; r < nr; ++r)
        for (uint32_t l = 0; l < nl; ++l) {
            const uint32_t c = 2 * (r * nl + l);
            const size_t g = (((static_cast<size_t>(s) * nr + r) * nl + l) * kSym + k) * pad + u;
            h_re[g] = __float2half_rn(u < used ? o[c] : 0.0f);
            h_im[g] = __float2half_rn(u < used ? o[c + 1] : 0.0f);
        }
}

void FreeDev(void *p)
{
    if (p != nullptr) (void)hipFree(p);
}

}  // namespace

struct AiranCeRefiner {
    int device = 0;
    int cus = 256;
    PuschMimoConfig cfg{};
    Geom g{};
    uint32_t max_slots = 0;
    uint32_t step = 0;
    float *params = nullptr, *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr;
    bf16 *wf = nullptr, *wb = nullptr;
    bf16 *x = nullptr;
    float *target = nullptr, *out = nullptr;
    bf16 *act[kLayers] = {};  // act[1..5] hidden activations (act[0] unused: the input is x)
    bf16 *dz[2] = {};
    float *part = nullptr;
    size_t part_floats = 0;
    double *loss_part = nullptr;
    float *loss_dev = nullptr;
    float2 *gains = nullptr;
    float *tau = nullptr, *snr = nullptr, *slot_w = nullptr;
    uint32_t max_taps = 32;
    std::vector<float> host_w;  // staged slot weights of the last SynthBatch
    __half *dmrs_re = nullptr, *dmrs_im = nullptr;
    __half *rx_re = nullptr, *rx_im = nullptr;
    __half *hls_re = nullptr, *hls_im = nullptr, *noise = nullptr;
    // Input mode (airan_ce_refiner_SetInput): 0 LS interpolation of h_ls, 1
    // the LMMSE estimate (SynthBatch computes it with `filter` into est_*).
    uint32_t input_mode = 0;
    const LmmseFilter *filter = nullptr;
    void *ce_ws = nullptr;
    size_t ce_ws_bytes = 0;
    __half *est_re = nullptr, *est_im = nullptr;  // [slots][NR][L][14][padded]
    int dil[kLayers] = {1, 1, 1, 1, 1, 1};        // subcarrier dilation per layer (SetDilation)
    float *x32 = nullptr;                         // the input in fp32 (the residual's base)
    float *res_scale = nullptr;                   // residual scale per slot
    uint32_t res_mode = 0;                        // 0: scale 1; 1: the slot's noise standard deviation
};

namespace {

constexpr uint32_t kLossBlocks = 256;

template <int CIN, int COUT, int EPI>
void LaunchConv(const AiranCeRefiner *r, const ConvArgs &a, hipStream_t s)
{
    constexpr int TILE = 32 * (4 / (COUT / 32));
    // Work items: half a slot's symbol rows x one TILE-pixel column.
    const uint32_t items = a.S * 2 * ((a.W + TILE - 1) / TILE);
    const uint32_t grid = std::min<uint32_t>(items, static_cast<uint32_t>(r->cus) * 2);
    // The LDS rows hold the halo of the layer's dilation rounded up to a
    // power of two, so a small dilation keeps the rows short.
    if (a.dil == 1)
        hipLaunchKernelGGL((ConvKernel<CIN, COUT, EPI, 1>), dim3(grid), dim3(256), 0, s, a);
    else if (a.dil <= 2)
        hipLaunchKernelGGL((ConvKernel<CIN, COUT, EPI, 2>), dim3(grid), dim3(256), 0, s, a);
    else if (a.dil <= 4)
        hipLaunchKernelGGL((ConvKernel<CIN, COUT, EPI, 4>), dim3(grid), dim3(256), 0, s, a);
    else if (a.dil <= 8)
        hipLaunchKernelGGL((ConvKernel<CIN, COUT, EPI, 8>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((ConvKernel<CIN, COUT, EPI, kMaxDil>), dim3(grid), dim3(256), 0, s, a);
}

// Forward of the network on x [S][14][W][32] -> out [S][14][W][32] fp32,
// keeping the hidden activations; with grid_re / grid_im the last layer
// instead adds its output to the fp16 channel grid [S][16][14][pad] in place
// (the residual base is the grid itself; slots with keep[s] == 0 untouched).
void Forward(const AiranCeRefiner *r, const bf16 *x, uint32_t S, uint32_t W, hipStream_t s,
             __half *grid_re = nullptr, __half *grid_im = nullptr, const uint8_t *keep = nullptr)
{
    ConvArgs a{};
    a.S = S;
    a.W = W;
    const bf16 *in = x;
    for (int i = 0; i < kLayers; ++i) {
        a.in = in;
        a.dil = r->dil[i];
        a.wt = r->wf + ParamOffset(i);
        a.bias = r->params + ParamOffset(i) + WeightCount(i);
        if (i < kLayers - 1) {
            a.out = r->act[i + 1];
            if (i == 0)
                LaunchConv<32, 64, kEpiRelu>(r, a, s);
            else
                LaunchConv<64, 64, kEpiRelu>(r, a, s);
            in = r->act[i + 1];
        } else if (grid_re != nullptr) {
            a.scale = r->res_scale;
            a.g_re = grid_re;
            a.g_im = grid_im;
            a.keep = keep;
            a.pad = r->g.pad;
            LaunchConv<64, 32, kEpiGrid>(r, a, s);
        } else {
            a.aux32 = r->x32;
            a.scale = r->res_scale;
            a.out_f32 = r->out;
            LaunchConv<64, 32, kEpiResid>(r, a, s);
        }
    }
}

template <int CIN, int COUT>
void LaunchWgrad(const AiranCeRefiner *r, const bf16 *act, const bf16 *dz, uint32_t S, uint32_t W, int layer,
                 hipStream_t s)
{
    const uint32_t ntiles = S * kSym * ((W + 63) / 64);
    const uint32_t nb = std::min<uint32_t>(ntiles, static_cast<uint32_t>(r->cus) * 2);
    const int dil = r->dil[layer];
    if (dil == 1)
        hipLaunchKernelGGL((WgradKernel<CIN, COUT, 1>), dim3(nb), dim3(256), 0, s, act, dz, r->part, S, W, 1);
    else if (dil <= 2)
        hipLaunchKernelGGL((WgradKernel<CIN, COUT, 2>), dim3(nb), dim3(256), 0, s, act, dz, r->part, S, W, dil);
    else if (dil <= 4)
        hipLaunchKernelGGL((WgradKernel<CIN, COUT, 4>), dim3(nb), dim3(256), 0, s, act, dz, r->part, S, W, dil);
    else if (dil <= 8)
        hipLaunchKernelGGL((WgradKernel<CIN, COUT, 8>), dim3(nb), dim3(256), 0, s, act, dz, r->part, S, W, dil);
    else
        hipLaunchKernelGGL((WgradKernel<CIN, COUT, kMaxDil>), dim3(nb), dim3(256), 0, s, act, dz, r->part, S, W,
                           dil);
    const uint32_t n = static_cast<uint32_t>(WeightCount(layer) + kCout[layer]);
    hipLaunchKernelGGL(ReduceKernel, dim3(airan::CeilDiv(n, 256)), dim3(256), 0, s, r->part, nb, n,
                       r->grad + ParamOffset(layer));
}

// Loss and gradients of all parameters (after Forward on the same x).
void Backward(const AiranCeRefiner *r, const bf16 *x, const float *target, const float *slot_w, uint32_t S,
              uint32_t W, float *loss, hipStream_t s)
{
    const size_t n = static_cast<size_t>(S) * kSym * W * kCh;
    hipLaunchKernelGGL(LossKernel, dim3(kLossBlocks), dim3(256), 0, s, r->out, target, slot_w, r->res_scale,
                       static_cast<size_t>(kSym) * W * kCh, r->dz[0], r->loss_part, n);
    hipLaunchKernelGGL(LossFinishKernel, dim3(1), dim3(64), 0, s, r->loss_part, kLossBlocks, n,
                       loss != nullptr ? loss : r->loss_dev);
    int cur = 0;
    for (int i = kLayers - 1; i >= 0; --i) {
        const bf16 *a_in = i == 0 ? x : r->act[i];
        const bf16 *dzi = r->dz[cur];
        if (i == kLayers - 1)
            LaunchWgrad<64, 32>(r, a_in, dzi, S, W, i, s);
        else if (i == 0)
            LaunchWgrad<32, 64>(r, a_in, dzi, S, W, i, s);
        else
            LaunchWgrad<64, 64>(r, a_in, dzi, S, W, i, s);
        if (i == 0) break;
        ConvArgs a{};
        a.S = S;
        a.W = W;
        a.dil = r->dil[i];
        a.in = dzi;
        a.wt = r->wb + ParamOffset(i);
        a.aux = r->act[i];
        a.out = r->dz[cur ^ 1];
        if (i == kLayers - 1)
            LaunchConv<32, 64, kEpiMask>(r, a, s);
        else
            LaunchConv<64, 64, kEpiMask>(r, a, s);
        cur ^= 1;
    }
}

void Pack(const AiranCeRefiner *r, hipStream_t s)
{
    PackArgs a{};
    for (int i = 0; i < kLayers; ++i) {
        a.off[i] = static_cast<uint32_t>(ParamOffset(i));
        a.cin[i] = kCin[i];
        a.cout[i] = kCout[i];
    }
    hipLaunchKernelGGL(PackKernel, dim3(airan::CeilDiv(9 * 64 * 64, 256), kLayers), dim3(256), 0, s, r->params, r->wf,
                       r->wb, a);
}

bool ShapeOk(const Geom &g) { return g.nr == 4 && g.nl == 4 && g.ndmrs >= 1 && g.ndmrs <= 2; }

InputArgs MakeInput(const AiranCeRefiner *r, const __half *hr, const __half *hi, bf16 *x)
{
    const Geom &g = r->g;
    InputArgs a{};
    a.h_re = hr;
    a.h_im = hi;
    a.x = x;
    a.x32 = r->x32;
    a.used = g.used;
    a.nr = g.nr;
    a.nl = g.nl;
    a.ndmrs = g.ndmrs;
    a.pilot_pad = g.pilot_pad;
    for (uint32_t d = 0; d < g.ndmrs; ++d) a.dmrs_sym[d] = g.dmrs_sym[d];
    for (uint32_t l = 0; l < g.nl; ++l) {
        a.comb[l] = g.comb[l];
        a.count[l] = g.occ[l] ? g.n_occ : g.n_dmrs_re;
        a.partner[l] = -1;
        a.sign[l] = 1.0f;
        if (g.occ[l]) {
            for (uint32_t q = 0; q < g.nl; ++q)
                if (q != l && g.comb[q] == g.comb[l]) a.partner[l] = static_cast<int32_t>(q);
            const uint16_t port = r->cfg.dmrs_ports[l];
            a.sign[l] = port == 1000 || port == 1002 ? 1.0f : -1.0f;
        }
    }
    return a;
}

}  // namespace

extern "C" {

int32_t airan_ce_refiner_Create(const PuschMimoConfig *config, uint32_t max_slots, const float *params, uint64_t seed,
                                void *stream, AiranCeRefiner **refiner)
{
    if (config == nullptr || refiner == nullptr || stream == nullptr || max_slots == 0 || max_slots > 1024)
        return AIRAN_INVALID_ARGUMENT;
    *refiner = nullptr;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (!ShapeOk(g)) return AIRAN_UNSUPPORTED_PROFILE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto *r = new AiranCeRefiner{};
    r->cfg = *config;
    r->g = g;
    r->max_slots = max_slots;
    auto fail = [&](int32_t code) {
        airan_ce_refiner_Destroy(r);
        return code;
    };
    hipDeviceProp_t prop;
    if (hipGetDevice(&r->device) != hipSuccess || hipGetDeviceProperties(&prop, r->device) != hipSuccess)
        return fail(AIRAN_LAUNCH_FAILED);
    r->cus = prop.multiProcessorCount;
    const size_t px = static_cast<size_t>(max_slots) * kSym * g.used;
    const size_t nwg = static_cast<size_t>(r->cus) * 2;
    r->part_floats = nwg * (9 * 64 * 64 + 64);
    auto alloc = [](auto **p, size_t bytes) { return hipMalloc(reinterpret_cast<void **>(p), bytes) == hipSuccess; };
    bool ok = alloc(&r->params, kParams * 4) && alloc(&r->adam_m, kParams * 4) && alloc(&r->adam_v, kParams * 4) &&
              alloc(&r->grad, kParams * 4) && alloc(&r->wf, kParams * 2) && alloc(&r->wb, kParams * 2) &&
              alloc(&r->x, px * kCh * 2) && alloc(&r->target, px * kCh * 4) && alloc(&r->out, px * kCh * 4) &&
              alloc(&r->dz[0], px * kHid * 2) && alloc(&r->dz[1], px * kHid * 2) &&
              alloc(&r->part, r->part_floats * 4) && alloc(&r->loss_part, kLossBlocks * 8) &&
              alloc(&r->loss_dev, 4) &&
              alloc(&r->gains, static_cast<size_t>(max_slots) * g.nr * g.nl * r->max_taps * kSym * 8) &&
              alloc(&r->tau, static_cast<size_t>(max_slots) * r->max_taps * 4) && alloc(&r->snr, max_slots * 4) && alloc(&r->slot_w, max_slots * 4) &&
              alloc(&r->dmrs_re, static_cast<size_t>(g.nl) * g.ndmrs * g.ref_pad * 2) &&
              alloc(&r->dmrs_im, static_cast<size_t>(g.nl) * g.ndmrs * g.ref_pad * 2) &&
              alloc(&r->rx_re, static_cast<size_t>(max_slots) * g.nr * kSym * g.pad * 2) &&
              alloc(&r->rx_im, static_cast<size_t>(max_slots) * g.nr * kSym * g.pad * 2) &&
              alloc(&r->hls_re, static_cast<size_t>(max_slots) * g.nr * g.nl * g.ndmrs * g.pilot_pad * 2) &&
              alloc(&r->hls_im, static_cast<size_t>(max_slots) * g.nr * g.nl * g.ndmrs * g.pilot_pad * 2) &&
              alloc(&r->noise, static_cast<size_t>(max_slots) * g.nr * 2);
    for (int i = 1; i < kLayers && ok; ++i) ok = alloc(&r->act[i], px * kHid * 2);
    ok = ok && alloc(&r->x32, px * kCh * 4) && alloc(&r->res_scale, static_cast<size_t>(max_slots) * 4);
    if (!ok) return fail(AIRAN_RESOURCE_TOO_SMALL);
    hipLaunchKernelGGL(NoiseScaleKernel, dim3(airan::CeilDiv(max_slots, 64)), dim3(64), 0, s,
                       static_cast<const __half *>(nullptr), g.nr, max_slots, r->res_scale);
    // Initial weights: given, or He-normal (std sqrt(2 / (9 cin))) with zero
    // biases from a host generator seeded by `seed`.
    std::vector<float> p(kParams, 0.0f);
    if (params != nullptr) {
        std::copy(params, params + kParams, p.begin());
    } else {
        uint64_t z = seed;
        auto next = [&z]() {
            z += 0x9e3779b97f4a7c15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
            return x ^ (x >> 31);
        };
        for (int i = 0; i < kLayers; ++i) {
            const double sd = std::sqrt(2.0 / (9.0 * kCin[i]));
            for (size_t k = 0; k < WeightCount(i); ++k) {
                const double u1 = (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                const double u2 = static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0);
                p[ParamOffset(i) + k] =
                    static_cast<float>(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
            }
        }
        // The last layer starts near zero so the untrained network is close
        // to the identity (the LS interpolation).
        for (size_t k = 0; k < WeightCount(kLayers - 1); ++k) p[ParamOffset(kLayers - 1) + k] *= 0.01f;
    }
    MimoDmrsGenRuntimeV1 *rt = nullptr;
    if ((st = airan_mimo_dmrs_gen_CreateRuntime(&rt)) != AIRAN_OK) return fail(st);
    PuschMimoLayout lay{};
    airan_DerivePuschMimoLayout(config, &lay);
    MimoDmrsGenOpArgsV1 ga{};
    ga.abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    ga.struct_size = sizeof(ga);
    ga.dmrs_re = r->dmrs_re;
    ga.dmrs_im = r->dmrs_im;
    ga.config = config;
    ga.layout = &lay;
    ga.stream = stream;
    st = airan_mimo_dmrs_gen_Enqueue(rt, &ga);
    const bool copy_ok =
        hipMemcpyAsync(r->params, p.data(), kParams * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
        hipMemsetAsync(r->adam_m, 0, kParams * 4, s) == hipSuccess &&
        hipMemsetAsync(r->adam_v, 0, kParams * 4, s) == hipSuccess &&
        hipMemsetAsync(r->rx_re, 0, static_cast<size_t>(max_slots) * g.nr * kSym * g.pad * 2, s) == hipSuccess &&
        hipMemsetAsync(r->rx_im, 0, static_cast<size_t>(max_slots) * g.nr * kSym * g.pad * 2, s) == hipSuccess;
    if (st == AIRAN_OK && copy_ok) Pack(r, s);
    const bool sync_ok = hipStreamSynchronize(s) == hipSuccess;
    airan_mimo_dmrs_gen_DestroyRuntime(rt);
    if (st != AIRAN_OK) return fail(st);
    if (!copy_ok || !sync_ok) return fail(AIRAN_LAUNCH_FAILED);
    *refiner = r;
    return AIRAN_OK;
}

void airan_ce_refiner_Destroy(AiranCeRefiner *r)
{
    if (r == nullptr) return;
    void *bufs[] = {r->params, r->adam_m, r->adam_v, r->grad, r->wf, r->wb, r->x, r->target, r->out, r->dz[0],
                    r->dz[1], r->part, r->loss_part, r->loss_dev, r->gains, r->tau, r->snr, r->slot_w, r->dmrs_re,
                    r->dmrs_im,
                    r->rx_re, r->rx_im, r->hls_re, r->hls_im, r->noise, r->ce_ws, r->est_re, r->est_im,
                    r->x32, r->res_scale};
    for (void *b : bufs) FreeDev(b);
    for (bf16 *a : r->act) FreeDev(a);
    delete r;
}

int32_t airan_ce_refiner_GetParams(AiranCeRefiner *r, float *params)
{
    if (r == nullptr || params == nullptr) return AIRAN_INVALID_ARGUMENT;
    return hipMemcpy(params, r->params, kParams * 4, hipMemcpyDeviceToHost) == hipSuccess ? AIRAN_OK
                                                                                          : AIRAN_LAUNCH_FAILED;
}

int32_t airan_ce_refiner_SetParams(AiranCeRefiner *r, const float *params)
{
    if (r == nullptr || params == nullptr) return AIRAN_INVALID_ARGUMENT;
    if (hipMemcpy(r->params, params, kParams * 4, hipMemcpyHostToDevice) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    Pack(r, nullptr);
    return hipDeviceSynchronize() == hipSuccess ? AIRAN_OK : AIRAN_LAUNCH_FAILED;
}

int32_t airan_ce_refiner_Enqueue(AiranCeRefiner *r, const void *h_ls_re, const void *h_ls_im, uint32_t slots,
                                 const uint8_t *keep, void *h_re, void *h_im, void *stream)
{
    return airan_ce_refiner_EnqueueWithNoise(r, h_ls_re, h_ls_im, nullptr, slots, keep, h_re, h_im, stream);
}

int32_t airan_ce_refiner_EnqueueWithNoise(AiranCeRefiner *r, const void *h_ls_re, const void *h_ls_im,
                                          const void *noise, uint32_t slots, const uint8_t *keep, void *h_re,
                                          void *h_im, void *stream)
{
    if (r == nullptr || h_ls_re == nullptr || h_ls_im == nullptr || h_re == nullptr || h_im == nullptr ||
        stream == nullptr || slots == 0 || slots > r->max_slots || (r->res_mode == 1 && noise == nullptr))
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Geom &g = r->g;
    hipLaunchKernelGGL(NoiseScaleKernel, dim3(airan::CeilDiv(slots, 64)), dim3(64), 0, s,
                       r->res_mode == 1 ? static_cast<const __half *>(noise) : nullptr, g.nr, slots, r->res_scale);
    if (r->input_mode == 1) {
        // The residual base is the estimate grid itself: the last layer
        // writes the refined grid in place, no fp32 copies.
        hipLaunchKernelGGL(EstimateInputKernel<false>, dim3(airan::CeilDiv(g.pad, 64), kSym, slots), dim3(256), 0,
                           s, static_cast<__half *>(h_re), static_cast<__half *>(h_im), r->x, nullptr, g.used, g.pad,
                           true, keep);
        Forward(r, r->x, slots, g.used, s, static_cast<__half *>(h_re), static_cast<__half *>(h_im), keep);
        return airan::LaunchStatus();
    }
    const InputArgs ia =
        MakeInput(r, static_cast<const __half *>(h_ls_re), static_cast<const __half *>(h_ls_im), r->x);
    hipLaunchKernelGGL(InputKernel, dim3(airan::CeilDiv(g.used, 256), slots), dim3(256), 0, s, ia);
    Forward(r, r->x, slots, g.used, s);
    hipLaunchKernelGGL(OutToGridKernel, dim3(airan::CeilDiv(g.pad, 256), kSym, slots), dim3(256), 0, s, r->out, keep,
                       static_cast<__half *>(h_re), static_cast<__half *>(h_im), g.used, g.pad, g.nr, g.nl);
    return airan::LaunchStatus();
}

int32_t airan_ce_refiner_SetInput(AiranCeRefiner *r, uint32_t mode, const struct LmmseFilter *filter)
{
    if (r == nullptr || mode > 1 || (filter != nullptr && !airan::CeFilterMatches(filter, r->g)))
        return AIRAN_INVALID_ARGUMENT;
    if (mode == 1 && filter != nullptr && r->ce_ws == nullptr) {
        const Geom &g = r->g;
        const size_t grid = static_cast<size_t>(r->max_slots) * g.nr * g.nl * kSym * g.pad * sizeof(__half);
        r->ce_ws_bytes = airan::CeWorkspaceBytes(g, r->max_slots);
        if (hipMalloc(&r->ce_ws, r->ce_ws_bytes) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&r->est_re), grid) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&r->est_im), grid) != hipSuccess)
            return AIRAN_RESOURCE_TOO_SMALL;
    }
    r->input_mode = mode;
    r->filter = mode == 1 ? filter : nullptr;
    return AIRAN_OK;
}

int32_t airan_ce_refiner_SetResidualScale(AiranCeRefiner *r, uint32_t mode)
{
    if (r == nullptr || mode > 1) return AIRAN_INVALID_ARGUMENT;
    r->res_mode = mode;
    return AIRAN_OK;
}

int32_t airan_ce_refiner_SetSlotScale(AiranCeRefiner *r, const float *scale, uint32_t slots)
{
    if (r == nullptr || scale == nullptr || slots == 0 || slots > r->max_slots) return AIRAN_INVALID_ARGUMENT;
    return hipMemcpy(r->res_scale, scale, slots * 4, hipMemcpyHostToDevice) == hipSuccess ? AIRAN_OK
                                                                                          : AIRAN_LAUNCH_FAILED;
}

int32_t airan_ce_refiner_SetDilation(AiranCeRefiner *r, const uint32_t *dilation, uint32_t count)
{
    if (r == nullptr || dilation == nullptr || count != kLayers) return AIRAN_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < count; ++i)
        if (dilation[i] < 1 || dilation[i] > static_cast<uint32_t>(kMaxDil)) return AIRAN_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < count; ++i) r->dil[i] = static_cast<int>(dilation[i]);
    return AIRAN_OK;
}

int32_t airan_ce_refiner_GetDilation(const AiranCeRefiner *r, uint32_t *dilation, uint32_t count)
{
    if (r == nullptr || dilation == nullptr || count != kLayers) return AIRAN_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < count; ++i) dilation[i] = static_cast<uint32_t>(r->dil[i]);
    return AIRAN_OK;
}

int32_t airan_ce_refiner_SynthBatch(AiranCeRefiner *r, const float *gains, const float *tap_delay_s, uint32_t taps,
                                    const float *snr_db, uint32_t slots, float scs_hz, uint64_t seed, void *stream)
{
    if (r == nullptr || gains == nullptr || tap_delay_s == nullptr || snr_db == nullptr || stream == nullptr ||
        slots == 0 || slots > r->max_slots || taps == 0 || taps > r->max_taps || !(scs_hz > 0.0f) ||
        (r->input_mode == 1 && r->filter == nullptr))
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Geom &g = r->g;
    const size_t ng = static_cast<size_t>(slots) * g.nr * g.nl * taps * kSym;
    // Training weights: each slot's squared error relative to its noise
    // level (weight proportional to the linear SNR, mean 1), so every SNR of
    // the batch counts alike instead of the noisiest slots dominating.
    double wsum = 0.0;
    for (uint32_t i = 0; i < slots; ++i) wsum += std::pow(10.0, 0.1 * snr_db[i]);
    r->host_w.resize(2 * static_cast<size_t>(slots));
    for (uint32_t i = 0; i < slots; ++i) {
        r->host_w[i] = static_cast<float>(std::pow(10.0, 0.1 * snr_db[i]) * slots / wsum);
        // Residual scale: the noise standard deviation of the slot.
        r->host_w[slots + i] = r->res_mode == 1 ? static_cast<float>(std::pow(10.0, -0.05 * snr_db[i])) : 1.0f;
    }
    if (hipMemcpyAsync(r->gains, gains, ng * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(r->slot_w, r->host_w.data(), slots * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(r->res_scale, r->host_w.data() + slots, slots * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(r->tau, tap_delay_s, static_cast<size_t>(slots) * taps * 4, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(r->snr, snr_db, slots * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return AIRAN_LAUNCH_FAILED;
    SynthArgs a{};
    a.gains = r->gains;
    a.tau = r->tau;
    a.snr_db = r->snr;
    a.dmrs_re = r->dmrs_re;
    a.dmrs_im = r->dmrs_im;
    a.target = r->target;
    a.rx_re = r->rx_re;
    a.rx_im = r->rx_im;
    a.taps = taps;
    a.used = g.used;
    a.pad = g.pad;
    a.nr = g.nr;
    a.nl = g.nl;
    a.ndmrs = g.ndmrs;
    a.ref_pad = g.ref_pad;
    for (uint32_t d = 0; d < g.ndmrs; ++d) a.dmrs_sym[d] = g.dmrs_sym[d];
    for (uint32_t l = 0; l < g.nl; ++l) a.comb[l] = g.comb[l];
    a.scs = scs_hz;
    a.seed = seed;
    hipLaunchKernelGGL(SynthKernel, dim3(airan::CeilDiv(g.used, 256), slots), dim3(256), 0, s, a);
    int32_t st = airan::LaunchDmrsLs(g, r->rx_re, r->rx_im, r->dmrs_re, r->dmrs_im, r->hls_re, r->hls_im, r->noise,
                                     slots, s);
    if (st != AIRAN_OK) return st;
    if (r->input_mode == 1) {
        st = airan::LaunchCeApply(r->filter, r->hls_re, r->hls_im, r->noise, nullptr, r->ce_ws, r->ce_ws_bytes,
                                  r->est_re, r->est_im, g.nl, slots, s);
        if (st != AIRAN_OK) return st;
        hipLaunchKernelGGL(EstimateInputKernel<true>, dim3(airan::CeilDiv(g.used, 64), kSym, slots), dim3(256), 0, s,
                           r->est_re, r->est_im, r->x, r->x32, g.used, g.pad, false, nullptr);
    } else {
        const InputArgs ia = MakeInput(r, r->hls_re, r->hls_im, r->x);
        hipLaunchKernelGGL(InputKernel, dim3(airan::CeilDiv(g.used, 256), slots), dim3(256), 0, s, ia);
    }
    return airan::LaunchStatus();
}

int32_t airan_ce_refiner_Batch(const AiranCeRefiner *r, AiranCeRefinerBatchV1 *b)
{
    if (r == nullptr || b == nullptr || b->abi_version != AIRAN_PUSCH_MIMO_ABI_VERSION ||
        b->struct_size < sizeof(AiranCeRefinerBatchV1))
        return AIRAN_INVALID_ARGUMENT;
    b->x = r->x;
    b->target = r->target;
    b->out = r->out;
    b->h_ls_re = r->hls_re;
    b->h_ls_im = r->hls_im;
    b->noise_ls = r->noise;
    return AIRAN_OK;
}

int32_t airan_ce_refiner_TrainStep(AiranCeRefiner *r, uint32_t slots, float lr, float *loss_out, void *stream)
{
    if (r == nullptr || stream == nullptr || slots == 0 || slots > r->max_slots || !(lr > 0.0f))
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t W = r->g.used;
    Forward(r, r->x, slots, W, s);
    Backward(r, r->x, r->target, r->slot_w, slots, W, loss_out, s);
    ++r->step;
    const float bc1 = 1.0f - std::pow(0.9f, static_cast<float>(r->step));
    const float bc2 = 1.0f - std::pow(0.999f, static_cast<float>(r->step));
    hipLaunchKernelGGL(AdamKernel, dim3(airan::CeilDiv(kParams, 256)), dim3(256), 0, s, r->params, r->adam_m,
                       r->adam_v, r->grad, static_cast<uint32_t>(kParams), lr, bc1, bc2);
    Pack(r, s);
    return airan::LaunchStatus();
}

int32_t airan_ce_refiner_Forward(AiranCeRefiner *r, uint32_t slots, void *stream)
{
    if (r == nullptr || stream == nullptr || slots == 0 || slots > r->max_slots) return AIRAN_INVALID_ARGUMENT;
    Forward(r, r->x, slots, r->g.used, static_cast<hipStream_t>(stream));
    return airan::LaunchStatus();
}

int32_t airan_ce_refiner_Gradients(AiranCeRefiner *r, const void *x, const void *target, uint32_t slots,
                                   uint32_t width, float *grad, float *loss, void *stream)
{
    if (r == nullptr || x == nullptr || target == nullptr || grad == nullptr || loss == nullptr ||
        stream == nullptr || slots == 0 || slots > r->max_slots || width == 0 || width > r->g.used)
        return AIRAN_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bf16 *xb = static_cast<const bf16 *>(x);
    hipLaunchKernelGGL(WidenKernel, dim3(256), dim3(256), 0, s, xb, r->x32,
                       static_cast<size_t>(slots) * kSym * width * kCh);
    Forward(r, xb, slots, width, s);
    Backward(r, xb, static_cast<const float *>(target), nullptr, slots, width, loss, s);
    if (hipMemcpyAsync(grad, r->grad, kParams * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return AIRAN_LAUNCH_FAILED;
    return airan::LaunchStatus();
}

}  // extern "C"
