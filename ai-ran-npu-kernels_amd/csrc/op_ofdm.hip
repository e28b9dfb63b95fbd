// op_ofdm.hip -- OFDM demodulation: CP removal + unitary DFT per symbol.
//
// The reference computes the DFT as two small dense matmuls on its matrix
// unit (ofdm/ofdm_demod_batch/ofdm_demod_kernel.cpp:889, factors from
// ofdm/ofdm_demod_siso/scripts/ofdm_demod_ref.py:100-136).  On gfx950 the
// operation is HBM bound (4 B in, 4 B out per sample against ~5 log2(N)
// flops), so the kernel is a Stockham FFT in fp32 held in LDS: one 16-point
// radix pass per 4 bits of log2(N) plus one 2/4/8-point pass, each thread
// owning 16 points.  The first pass reads the CP-stripped samples straight
// from HBM; the last pass leaves natural order in LDS and the write-out
// permutes into the reference's stage-4 order ([P,64], bin k at
// (k % P) * 64 + k / P) or, fused, straight into the used-subcarrier grid
// that re_demap_batch would produce.  The BF16 / FP16 precisions compute the
// DFT as the reference does, as two dense products on the matrix cores
// (op_ofdm_mfma.hip).
#include <cmath>
#include <mutex>

#include "common.h"
#include "launch.h"
#include "ragged.h"

using airan::Geom;

// No multiply-add contraction: the stage-4 and fused write-outs are separate
// kernel instantiations and must round identically, element for element.
#pragma clang fp contract(off)

namespace airan {

// W_4096^k = exp(-2 pi i k / 4096); smaller sizes index it with stride 4096/N.
__device__ float2 kTwiddle4096[4096];

namespace {
std::mutex g_tw_mutex;
bool g_tw_done[64] = {};
}  // namespace

int InitOfdmTables()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    if (g_tw_done[dev]) return AIRAN_OK;
    static float2 host[4096];
    for (int k = 0; k < 4096; ++k) {
        const double a = -2.0 * M_PI * k / 4096.0;
        host[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(kTwiddle4096), host, sizeof(host)) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    g_tw_done[dev] = true;
    return AIRAN_OK;
}

}  // namespace airan

namespace {

constexpr int kThreads = 256;
constexpr int kMaxN = 4096;
constexpr int kLdsSpan = kMaxN + kMaxN / 16;  // one pad float per 16 (bank spread)

__device__ inline float2 CMul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline float2 CAdd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 CSub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 MulMinusI(float2 a) { return make_float2(a.y, -a.x); }

// Forward 4-point DFT in place.
__device__ inline void Dft4(float2 &x0, float2 &x1, float2 &x2, float2 &x3)
{
    const float2 s0 = CAdd(x0, x2), d0 = CSub(x0, x2);
    const float2 s1 = CAdd(x1, x3), d1 = MulMinusI(CSub(x1, x3));
    x0 = CAdd(s0, s1);
    x2 = CSub(s0, s1);
    x1 = CAdd(d0, d1);
    x3 = CSub(d0, d1);
}

// exp(-2 pi i m / 16), m = 0..9.
__device__ inline float2 W16(int m)
{
    constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, h = 0.70710678118654752f;
    switch (m) {
    case 0: return make_float2(1.0f, 0.0f);
    case 1: return make_float2(c1, -s1);
    case 2: return make_float2(h, -h);
    case 3: return make_float2(s1, -c1);
    case 4: return make_float2(0.0f, -1.0f);
    case 6: return make_float2(-h, -h);
    default: return make_float2(-c1, s1);  // m = 9
    }
}

// Forward R-point DFT of v[0..R-1], natural order out.
template <int R>
__device__ inline void Dft(float2 *v)
{
    if constexpr (R == 2) {
        const float2 t = v[0];
        v[0] = CAdd(t, v[1]);
        v[1] = CSub(t, v[1]);
    } else if constexpr (R == 4) {
        Dft4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // r = b + 2a: DFT4 over a, twiddle W8^(b c), DFT2 over b -> k = c + 4 d.
        float2 u[2][4];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            u[b][0] = v[b];
            u[b][1] = v[b + 2];
            u[b][2] = v[b + 4];
            u[b][3] = v[b + 6];
            Dft4(u[b][0], u[b][1], u[b][2], u[b][3]);
        }
#pragma unroll
        for (int c = 1; c < 4; ++c) u[1][c] = CMul(u[1][c], W16(2 * c));
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            v[c] = CAdd(u[0][c], u[1][c]);
            v[c + 4] = CSub(u[0][c], u[1][c]);
        }
    } else {
        static_assert(R == 16, "radix");
        // r = b + 4a: DFT4 over a, twiddle W16^(b c), DFT4 over b -> k = c + 4 d.
        float2 u[4][4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            u[b][0] = v[b];
            u[b][1] = v[b + 4];
            u[b][2] = v[b + 8];
            u[b][3] = v[b + 12];
            Dft4(u[b][0], u[b][1], u[b][2], u[b][3]);
        }
#pragma unroll
        for (int b = 1; b < 4; ++b)
#pragma unroll
            for (int c = 1; c < 4; ++c) u[b][c] = CMul(u[b][c], W16(b * c));
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            Dft4(u[0][c], u[1][c], u[2][c], u[3][c]);
            v[c] = u[0][c];
            v[c + 4] = u[1][c];
            v[c + 8] = u[2][c];
            v[c + 12] = u[3][c];
        }
    }
}

__device__ inline uint32_t Pad(uint32_t i) { return i + (i >> 4); }

struct OfdmArgs {
    uint32_t total_symbols;  // batch * 14
    uint32_t samples;        // per antenna stream
    uint32_t used, pad;
    float out_scale;         // input scale / sqrt(N)
    uint32_t sym_start[14];
};

// One Stockham pass of radix R over N points held in LDS (natural-order
// in, natural-order out after the last pass).  Each thread of a symbol owns
// 16 / R butterflies.  The last pass applies the output scale before its
// LDS store, so the write-out is a pure fp32 -> fp16 conversion (no multiply
// the compiler could fold into a single-rounding v_fma_mix).
template <int LOGN, int R, bool LAST>
__device__ inline void LdsPass(float *lre, float *lim, uint32_t j, uint32_t ns, bool active, float scale)
{
    constexpr uint32_t N = 1u << LOGN;
    constexpr uint32_t TPS = N / 16;
    constexpr int B = 16 / R;
    float2 v[B][R], tw[B][R];
    if (active) {
        // Twiddle loads first: their latency overlaps the LDS reads and the
        // barrier instead of following them.
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const uint32_t k = (j + t * TPS) & (ns - 1);  // ns is a power of two
            const uint32_t step = (kMaxN / R) / ns;
#pragma unroll
            for (int r = 1; r < R; ++r) tw[t][r] = airan::kTwiddle4096[k * r * step];
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const uint32_t jj = j + t * TPS;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t i = Pad(jj + r * (N / R));
                v[t][r] = make_float2(lre[i], lim[i]);
            }
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const uint32_t jj = j + t * TPS;
            const uint32_t k = jj & (ns - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) v[t][r] = CMul(v[t][r], tw[t][r]);
            Dft<R>(v[t]);
            const uint32_t base = (jj - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t i = Pad(base + r * ns);
                lre[i] = LAST ? v[t][r].x * scale : v[t][r].x;
                lim[i] = LAST ? v[t][r].y * scale : v[t][r].y;
            }
        }
    }
    __syncthreads();
}

// OUT_MODE 0: stage-4 order [14, N] per stream; 1: used-subcarrier grid [14, pad].
// RAGGED (with OUT_MODE 1): stream i is streams[i] (its own IQ and grid
// offsets and used / pad), all of this FFT size.
template <int LOGN, bool IQ_FLOAT, int OUT_MODE, bool RAGGED = false>
__global__ void __launch_bounds__(kThreads) OfdmKernel(const void *__restrict__ iq, __half *__restrict__ out_re,
                                                       __half *__restrict__ out_im, OfdmArgs a,
                                                       const airan::RaggedStream *__restrict__ streams = nullptr)
{
    constexpr uint32_t N = 1u << LOGN;
    constexpr uint32_t TPS = N / 16;
    constexpr uint32_t SPW = kThreads / TPS;  // symbols per workgroup
    __shared__ float lds_re[kLdsSpan], lds_im[kLdsSpan];
    const uint32_t sub = threadIdx.x / TPS, j = threadIdx.x - sub * TPS;
    const uint32_t gsym = blockIdx.x * SPW + sub;
    const bool active = gsym < a.total_symbols;
    const uint32_t stream = active ? gsym / 14 : 0, sym = active ? gsym - stream * 14 : 0;
    float *lre = lds_re + sub * (N + N / 16);
    float *lim = lds_im + sub * (N + N / 16);

    uint64_t iq_base = static_cast<uint64_t>(stream) * a.samples;
    uint64_t out_base = static_cast<uint64_t>(gsym) * a.pad;
    uint32_t used = a.used, pad = a.pad;
    if constexpr (RAGGED) {
        if (active) {
            const airan::RaggedStream st = streams[stream];
            iq_base = st.iq_off;
            used = st.used;
            pad = st.pad;
            out_base = st.out_off + static_cast<uint64_t>(sym) * pad;
        }
    }
    // First radix-16 pass straight from HBM (ns = 1: no twiddles).
    if (active) {
        float2 v[16];
        const size_t s0 = iq_base + a.sym_start[sym];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const size_t n = s0 + j + r * TPS;
            if constexpr (IQ_FLOAT) {
                v[r] = static_cast<const float2 *>(iq)[n];
            } else {
                const uint32_t w = static_cast<const uint32_t *>(iq)[n];
                v[r] = make_float2(static_cast<float>(static_cast<int16_t>(w & 0xffffu)),
                                   static_cast<float>(static_cast<int16_t>(w >> 16)));
            }
        }
        Dft<16>(v);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t i = Pad(j * 16 + r);
            lre[i] = v[r].x;
            lim[i] = v[r].y;
        }
    }
    __syncthreads();
    uint32_t ns = 16;
    constexpr int kPasses16 = LOGN / 4;  // radix-16 passes, the first included
    const float sc = a.out_scale;
#pragma unroll
    for (int p = 1; p < kPasses16; ++p) {
        if (LOGN % 4 == 0 && p == kPasses16 - 1)
            LdsPass<LOGN, 16, true>(lre, lim, j, ns, active, sc);
        else
            LdsPass<LOGN, 16, false>(lre, lim, j, ns, active, sc);
        ns *= 16;
    }
    if constexpr (LOGN % 4 == 1) LdsPass<LOGN, 2, true>(lre, lim, j, ns, active, sc);
    if constexpr (LOGN % 4 == 2) LdsPass<LOGN, 4, true>(lre, lim, j, ns, active, sc);
    if constexpr (LOGN % 4 == 3) LdsPass<LOGN, 8, true>(lre, lim, j, ns, active, sc);
    if (!active) return;

    // Write-out (already scaled): two adjacent outputs per thread per step.
    if constexpr (OUT_MODE == 0) {
        constexpr uint32_t P = N / 64;
        __half2 *ore = reinterpret_cast<__half2 *>(out_re + static_cast<size_t>(gsym) * N);
        __half2 *oim = reinterpret_cast<__half2 *>(out_im + static_cast<size_t>(gsym) * N);
        for (uint32_t q = j; q < N / 2; q += TPS) {
            const uint32_t e = 2 * q;
            const uint32_t k0 = (e & 63) * P + (e >> 6), k1 = ((e + 1) & 63) * P + ((e + 1) >> 6);
            ore[q] = __floats2half2_rn(lre[Pad(k0)], lre[Pad(k1)]);
            oim[q] = __floats2half2_rn(lim[Pad(k0)], lim[Pad(k1)]);
        }
    } else {
        const uint32_t half_used = used / 2;
        __half2 *ore = reinterpret_cast<__half2 *>(out_re + out_base);
        __half2 *oim = reinterpret_cast<__half2 *>(out_im + out_base);
        for (uint32_t q = j; q < pad / 2; q += TPS) {
            const uint32_t u = 2 * q;
            float r0 = 0.0f, i0 = 0.0f, r1 = 0.0f, i1 = 0.0f;
            if (u < used) {
                const uint32_t k = u < half_used ? N - half_used + u : u - half_used;
                r0 = lre[Pad(k)];
                i0 = lim[Pad(k)];
            }
            if (u + 1 < used) {
                const uint32_t k = u + 1 < half_used ? N - half_used + u + 1 : u + 1 - half_used;
                r1 = lre[Pad(k)];
                i1 = lim[Pad(k)];
            }
            ore[q] = __floats2half2_rn(r0, r1);
            oim[q] = __floats2half2_rn(i0, i1);
        }
    }
}

template <bool IQ_FLOAT, int OUT_MODE>
int32_t LaunchLogN(uint32_t logn, dim3 grid, hipStream_t s, const void *iq, __half *ore, __half *oim,
                   const OfdmArgs &a)
{
    switch (logn) {
    case 7: hipLaunchKernelGGL((OfdmKernel<7, IQ_FLOAT, OUT_MODE>), grid, dim3(kThreads), 0, s, iq, ore, oim, a); break;
    case 8: hipLaunchKernelGGL((OfdmKernel<8, IQ_FLOAT, OUT_MODE>), grid, dim3(kThreads), 0, s, iq, ore, oim, a); break;
    case 9: hipLaunchKernelGGL((OfdmKernel<9, IQ_FLOAT, OUT_MODE>), grid, dim3(kThreads), 0, s, iq, ore, oim, a); break;
    case 10: hipLaunchKernelGGL((OfdmKernel<10, IQ_FLOAT, OUT_MODE>), grid, dim3(kThreads), 0, s, iq, ore, oim, a); break;
    case 11: hipLaunchKernelGGL((OfdmKernel<11, IQ_FLOAT, OUT_MODE>), grid, dim3(kThreads), 0, s, iq, ore, oim, a); break;
    case 12: hipLaunchKernelGGL((OfdmKernel<12, IQ_FLOAT, OUT_MODE>), grid, dim3(kThreads), 0, s, iq, ore, oim, a); break;
    default: return AIRAN_UNSUPPORTED_PROFILE;
    }
    return airan::LaunchStatus();
}

int32_t Launch(const OfdmDemodBatchOpArgsV1 *args, int out_mode, void *ore, void *oim)
{
    Geom g;
    airan::DeriveGeom(*args->config, &g);
    if (args->precision != AIRAN_DFT_FP32)
        return airan::LaunchMfmaDft(g, args->precision, args->iq_format, args->batch, args->iq, out_mode,
                                    static_cast<__half *>(ore), static_cast<__half *>(oim),
                                    static_cast<hipStream_t>(args->stream));
    if (airan::InitOfdmTables() != AIRAN_OK) return AIRAN_LAUNCH_FAILED;
    OfdmArgs a{};
    a.total_symbols = args->batch * g.nsym;
    a.samples = g.samples;
    a.used = g.used;
    a.pad = g.pad;
    a.out_scale = static_cast<float>((args->iq_format == AIRAN_IQ_INT16_Q8_8 ? 1.0 / 256.0 : 1.0) /
                                     std::sqrt(static_cast<double>(g.fft)));
    for (uint32_t s = 0; s < g.nsym; ++s) a.sym_start[s] = g.sym_start[s];
    uint32_t logn = 0;
    while ((1u << logn) < g.fft) ++logn;
    const uint32_t spw = kThreads / (g.fft / 16);
    const dim3 grid(airan::CeilDiv(a.total_symbols, spw));
    hipStream_t s = static_cast<hipStream_t>(args->stream);
    auto *re = static_cast<__half *>(ore);
    auto *im = static_cast<__half *>(oim);
    if (args->iq_format == AIRAN_IQ_COMPLEX64)
        return out_mode == 0 ? LaunchLogN<true, 0>(logn, grid, s, args->iq, re, im, a)
                             : LaunchLogN<true, 1>(logn, grid, s, args->iq, re, im, a);
    return out_mode == 0 ? LaunchLogN<false, 0>(logn, grid, s, args->iq, re, im, a)
                         : LaunchLogN<false, 1>(logn, grid, s, args->iq, re, im, a);
}

}  // namespace

namespace airan {

int32_t LaunchOfdmRagged(uint32_t fft, uint32_t iq_format, const RaggedStream *streams, uint32_t num_streams,
                         const void *iq, __half *out_re, __half *out_im, hipStream_t s)
{
    if (num_streams == 0) return AIRAN_OK;
    if (InitOfdmTables() != AIRAN_OK) return AIRAN_LAUNCH_FAILED;
    uint32_t logn = 0;
    while ((1u << logn) < fft) ++logn;
    if ((1u << logn) != fft || logn < 7 || logn > 12) return AIRAN_UNSUPPORTED_PROFILE;
    OfdmArgs a{};
    a.total_symbols = num_streams * kSymbols;
    a.out_scale = static_cast<float>((iq_format == AIRAN_IQ_INT16_Q8_8 ? 1.0 / 256.0 : 1.0) /
                                     std::sqrt(static_cast<double>(fft)));
    // The NR cyclic prefixes of this FFT size (CP 9N/128, plus N/64 on symbol 0).
    const uint32_t cs = fft * 9 / 128, cl = cs + fft / 64;
    uint32_t pos = 0;
    for (uint32_t sym = 0; sym < kSymbols; ++sym) {
        pos += sym == 0 ? cl : cs;
        a.sym_start[sym] = pos;
        pos += fft;
    }
    const dim3 grid(CeilDiv(a.total_symbols, kThreads / (fft / 16)));
    const bool f = iq_format == AIRAN_IQ_COMPLEX64;
#define AIRAN_RAGGED_OFDM(L)                                                                                      \
    case L:                                                                                                      \
        if (f)                                                                                                   \
            hipLaunchKernelGGL((OfdmKernel<L, true, 1, true>), grid, dim3(kThreads), 0, s, iq, out_re, out_im, a, \
                               streams);                                                                         \
        else                                                                                                     \
            hipLaunchKernelGGL((OfdmKernel<L, false, 1, true>), grid, dim3(kThreads), 0, s, iq, out_re, out_im,  \
                               a, streams);                                                                      \
        break;
    switch (logn) {
        AIRAN_RAGGED_OFDM(7)
        AIRAN_RAGGED_OFDM(8)
        AIRAN_RAGGED_OFDM(9)
        AIRAN_RAGGED_OFDM(10)
        AIRAN_RAGGED_OFDM(11)
        AIRAN_RAGGED_OFDM(12)
    default: return AIRAN_UNSUPPORTED_PROFILE;
    }
#undef AIRAN_RAGGED_OFDM
    return LaunchStatus();
}

}  // namespace airan

extern "C" {

// ofdm/ofdm_demod_batch/ofdm_demod_batch.h argument contract: batch of
// antenna streams of 14 symbols each, one slot of samples per stream.
int32_t airan_ofdm_demod_batch_ValidateOpArgs(const OfdmDemodBatchOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(OfdmDemodBatchOpArgsV1)) || a->iq == nullptr || a->config == nullptr ||
        a->stream == nullptr || a->batch == 0 || a->reserved != 0)
        return AIRAN_INVALID_ARGUMENT;
    if (a->iq_format != AIRAN_IQ_INT16_Q8_8 && a->iq_format != AIRAN_IQ_COMPLEX64) return AIRAN_INVALID_ARGUMENT;
    if (a->precision != AIRAN_DFT_FP32 && a->precision != AIRAN_DFT_BF16 && a->precision != AIRAN_DFT_FP16)
        return AIRAN_UNSUPPORTED_PROFILE;
    Geom g;
    const int32_t st = airan::DeriveGeom(*a->config, &g);
    if (st != AIRAN_OK) return st;
    if (a->precision != AIRAN_DFT_FP32 && !airan::MfmaDftSupported(g.fft)) return AIRAN_UNSUPPORTED_PROFILE;
    if (static_cast<uint64_t>(a->batch) * g.nsym > 0xffffffffull / kThreads) return AIRAN_INVALID_ARGUMENT;
    return AIRAN_OK;
}

int32_t airan_ofdm_demod_batch_Enqueue(const OfdmDemodBatchOpArgsV1 *a)
{
    int32_t st = airan_ofdm_demod_batch_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    if (a->out_re == nullptr || a->out_im == nullptr || a->out_re == a->out_im) return AIRAN_INVALID_ARGUMENT;
    return Launch(a, 0, a->out_re, a->out_im);
}

int32_t airan_ofdm_demod_redemap_Enqueue(const OfdmDemodBatchOpArgsV1 *a, void *rx_grid_re, void *rx_grid_im)
{
    int32_t st = airan_ofdm_demod_batch_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    if (rx_grid_re == nullptr || rx_grid_im == nullptr || rx_grid_re == rx_grid_im) return AIRAN_INVALID_ARGUMENT;
    // half2 stores: every row of the padded grid starts 4-byte aligned.
    if (a->config->flags != 0 || (a->config->padded_subcarriers & 1u) != 0) return AIRAN_UNSUPPORTED_PROFILE;
    return Launch(a, 1, rx_grid_re, rx_grid_im);
}

}  // extern "C"
