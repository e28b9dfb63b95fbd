// op_ofdm_mfma.hip -- OFDM demodulation as two dense DFT products on the
// matrix cores, in BF16 or FP16 with FP32 accumulation (the precision
// variants of ofdm_demod_batch; the FP32 path is the FFT of op_ofdm.hip).
//
// The reference factors the N = P * Q point DFT (Q = 64) into two matmuls on
// its matrix unit (ofdm/ofdm_demod_batch/ofdm_demod_kernel.cpp:862-919,
// factors ofdm/ofdm_demod_siso/scripts/ofdm_demod_ref.py:100-136).  With
// A[n1][n2] = x[64 n1 + n2] (the symbol's samples, row-major P x 64),
//   X1 = W_P A                         (P x P times P x 64)
//   X1'[k1][n2] = X1[k1][n2] W_N^(k1 n2)   (twiddle, fp32)
//   Y = X1' W_64                       (P x 64 times 64 x 64)
// gives Y[k1][k2] = X[k1 + P k2], which is the reference's stage-4 order.
// One workgroup per OFDM symbol.  The samples are scaled by a power of two
// per symbol (block scaling: the largest |re|, |im| lands in [128, 256)) and
// rounded to the 16-bit type.  The twiddled stage-1 output is rounded again
// for the second product.  Both products accumulate in fp32
// (v_mfma_f32_32x32x16_{bf16,f16}, four real products per complex product),
// and the output is scaled back and rounded to fp16.  Supported for
// N = 1024, 2048, 4096 (P = 16, 32, 64; P = 16 zero-pads the row block).
#include <cmath>
#include <mutex>

#include "cgemm.h"
#include "common.h"

using airan::Geom;

namespace airan {
extern __device__ float2 kTwiddle4096[4096];  // op_ofdm.hip
}

namespace {

typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

// DFT factor tables in fragment order (cgemm.h, one part, planes re / im):
// kWp[prec][P = 16, 32, 64] rows k1 (padded to 32), K = n1; kWq[prec] rows
// k2, K = n2 (W_64 is symmetric, so this is also its B-operand order).
constexpr uint32_t kTabElems = 2 * 4 * 2 * airan::kFragPlane;  // 64 rows x 64 K x 2 planes
__device__ uint16_t kWp[2][3][kTabElems];
__device__ uint16_t kWq[2][kTabElems];

std::mutex g_mutex;
bool g_done[64] = {};
uint16_t *g_wp[64] = {}, *g_wq[64] = {};  // device addresses of kWp / kWq per device

uint16_t HostBf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

uint16_t Host16(int prec, double v)
{
    const float f = static_cast<float>(v);
    return prec == 0 ? HostBf16(f) : __half_as_ushort(__float2half_rn(f));
}

// exp(-2 pi i m / n) with the quarter-turn values exact.
void Root(uint64_t m, uint32_t n, double *re, double *im)
{
    m %= n;
    if (4 * m % n == 0) {
        const uint64_t q = 4 * m / n;
        *re = q == 0 ? 1.0 : (q == 2 ? -1.0 : 0.0);
        *im = q == 1 ? -1.0 : (q == 3 ? 1.0 : 0.0);
        return;
    }
    const double a = -2.0 * M_PI * static_cast<double>(m) / n;
    *re = std::cos(a);
    *im = std::sin(a);
}

// Fragment-order table of W_n (rows r < rows_valid, K = n, rows padded to a
// multiple of 32): element (r, k) = exp(-2 pi i r k / n).
void BuildTable(int prec, uint32_t n, uint32_t rows_valid, uint16_t *out)
{
    const uint32_t rows = rows_valid < 32 ? 32 : rows_valid, kblocks = n / 16;
    std::memset(out, 0, kTabElems * sizeof(uint16_t));
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t k = 0; k < n; ++k) {
            double re = 0.0, im = 0.0;
            if (r < rows_valid) Root(static_cast<uint64_t>(r) * k, n, &re, &im);
            const size_t i0 = (((static_cast<size_t>(r >> 5) * kblocks + (k >> 4)) * 2 + 0) * 2 + ((k >> 3) & 1u)) * 256 +
                              (r & 31u) * 8 + (k & 7u);
            out[i0] = Host16(prec, re);
            out[i0 + 2 * 256] = Host16(prec, im);  // plane 1
        }
}

template <int PREC>
__device__ inline uint16_t To16(float x)
{
    if constexpr (PREC == AIRAN_DFT_BF16)
        return static_cast<uint16_t>(airan::Bf16Bits(x));
    else
        return __half_as_ushort(__float2half_rn(x));
}

template <int PREC>
__device__ inline airan::f32x16_t Mfma(uint4 a, uint4 b, airan::f32x16_t c)
{
    if constexpr (PREC == AIRAN_DFT_BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(airan::bf16x8_t, a),
                                                       __builtin_bit_cast(airan::bf16x8_t, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b),
                                                      c, 0, 0, 0);
}

struct MfmaDftArgs {
    uint32_t total_symbols, samples, used, pad;
    float out_scale;
    uint32_t sym_start[14];
};

constexpr int kScaleExp = 8;  // block scale: max |re|, |im| in [2^7, 2^8)

// OUT_MODE 0: stage-4 order [14, N] per stream; 1: used-subcarrier grid [14, pad].
template <int PREC, int P, bool IQ_FLOAT, int OUT_MODE>
__global__ void __launch_bounds__(256) OfdmMfmaKernel(const void *__restrict__ iq, __half *__restrict__ out_re,
                                                      __half *__restrict__ out_im, MfmaDftArgs a,
                                                      const uint16_t *__restrict__ wp, const uint16_t *__restrict__ wq)
{
    constexpr uint32_t Q = 64, N = P * Q, MR = P < 32 ? 32 : P;
    constexpr uint32_t XA_LD = P + 8, XT_LD = Q + 8;   // row strides (16-bit elements), rows 16-byte aligned
    constexpr uint32_t PER = N / 256;                  // samples per thread
    constexpr uint32_t MB = MR / 32, NB = Q / 32, BLOCKS = MB * NB, KB1 = P / 16, KB2 = Q / 16;
    static_assert(P >= 16 && P <= 64 && 2 * P * Q <= 2 * Q * XA_LD, "tile shape");
    // xa: the scaled samples transposed, [plane][n2][n1] (stage-1 B operand);
    // reused for the fp16 output [plane][k1 * 64 + k2] after stage 1.
    __shared__ __align__(16) uint16_t xa[2][Q * XA_LD];
    // xt: the twiddled stage-1 output, [plane][k1][n2] (stage-2 A operand).
    __shared__ __align__(16) uint16_t xt[2][MR * XT_LD];
    __shared__ float red[4];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t gsym = blockIdx.x;
    const uint32_t stream = gsym / 14, sym = gsym - stream * 14;
    const size_t s0 = static_cast<size_t>(stream) * a.samples + a.sym_start[sym];

    float2 v[PER];
    float m = 0.0f;
#pragma unroll
    for (uint32_t r = 0; r < PER; ++r) {
        const size_t n = s0 + t + 256 * r;
        if constexpr (IQ_FLOAT) {
            v[r] = static_cast<const float2 *>(iq)[n];
        } else {
            const uint32_t w = static_cast<const uint32_t *>(iq)[n];
            v[r] = make_float2(static_cast<float>(static_cast<int16_t>(w & 0xffffu)),
                               static_cast<float>(static_cast<int16_t>(w >> 16)));
        }
        m = fmaxf(m, fmaxf(fabsf(v[r].x), fabsf(v[r].y)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (lane == 0) red[wave] = m;
    if (P < 32) {
        // Rows P..31 of the stage-2 A operand are zero (the padded row block).
        for (uint32_t i = t; i < (MR - P) * XT_LD; i += 256) {
            xt[0][P * XT_LD + i] = 0;
            xt[1][P * XT_LD + i] = 0;
        }
    }
    __syncthreads();
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    int e = 0;
    (void)frexpf(m, &e);
    const float scale = m > 0.0f ? ldexpf(1.0f, kScaleExp - e) : 1.0f;
#pragma unroll
    for (uint32_t r = 0; r < PER; ++r) {
        const uint32_t n = t + 256 * r, n1 = n >> 6, n2 = n & 63u;
        xa[0][n2 * XA_LD + n1] = To16<PREC>(v[r].x * scale);
        xa[1][n2 * XA_LD + n1] = To16<PREC>(v[r].y * scale);
    }
    __syncthreads();

    const uint32_t mb = wave / NB, nb = wave - mb * NB;
    if (wave < BLOCKS) {
        // Stage 1: X1 block (rows 32 mb.., columns n2 = 32 nb..).
        airan::f32x16_t cr = {}, ci = {};
#pragma unroll
        for (uint32_t kb = 0; kb < KB1; ++kb) {
            const uint4 *w = reinterpret_cast<const uint4 *>(wp + (mb * KB1 + kb) * 2 * airan::kFragPlane) + lane;
            const uint4 ar = w[0], ai = w[airan::kFragPlane / 8];
            const uint32_t n2 = 32 * nb + (lane & 31u), n1 = 16 * kb + 8 * (lane >> 5);
            const uint4 br = *reinterpret_cast<const uint4 *>(&xa[0][n2 * XA_LD + n1]);
            const uint4 bi = *reinterpret_cast<const uint4 *>(&xa[1][n2 * XA_LD + n1]);
            cr = Mfma<PREC>(ar, br, cr);
            cr = Mfma<PREC>(airan::NegBf16x8(ai), bi, cr);
            ci = Mfma<PREC>(ar, bi, ci);
            ci = Mfma<PREC>(ai, br, ci);
        }
        // Twiddle W_N^(k1 n2) in fp32, then round for the second product.
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) {
            const uint32_t k1 = 32 * mb + 8 * (i / 4) + 4 * (lane >> 5) + (i & 3u);
            const uint32_t n2 = 32 * nb + (lane & 31u);
            if (k1 < P) {
                const float2 w = airan::kTwiddle4096[(k1 * n2 * (4096u / N)) & 4095u];
                const float xr = cr[i] * w.x - ci[i] * w.y;
                const float xi = cr[i] * w.y + ci[i] * w.x;
                xt[0][k1 * XT_LD + n2] = To16<PREC>(xr);
                xt[1][k1 * XT_LD + n2] = To16<PREC>(xi);
            }
        }
    }
    __syncthreads();
    const float osc = a.out_scale / scale;  // exact: scale is a power of two
    if (wave < BLOCKS) {
        // Stage 2: Y block (rows k1 = 32 mb.., columns k2 = 32 nb..).
        airan::f32x16_t cr = {}, ci = {};
#pragma unroll
        for (uint32_t kb = 0; kb < KB2; ++kb) {
            const uint32_t k1 = 32 * mb + (lane & 31u), n2 = 16 * kb + 8 * (lane >> 5);
            const uint4 ar = *reinterpret_cast<const uint4 *>(&xt[0][k1 * XT_LD + n2]);
            const uint4 ai = *reinterpret_cast<const uint4 *>(&xt[1][k1 * XT_LD + n2]);
            const uint4 *w = reinterpret_cast<const uint4 *>(wq + (nb * KB2 + kb) * 2 * airan::kFragPlane) + lane;
            const uint4 br = w[0], bi = w[airan::kFragPlane / 8];
            cr = Mfma<PREC>(ar, br, cr);
            cr = Mfma<PREC>(airan::NegBf16x8(ai), bi, cr);
            ci = Mfma<PREC>(ar, bi, ci);
            ci = Mfma<PREC>(ai, br, ci);
        }
        __half *ys_re = reinterpret_cast<__half *>(xa[0]), *ys_im = reinterpret_cast<__half *>(xa[1]);
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) {
            const uint32_t k1 = 32 * mb + 8 * (i / 4) + 4 * (lane >> 5) + (i & 3u);
            const uint32_t k2 = 32 * nb + (lane & 31u);
            if (k1 < P) {
                ys_re[k1 * Q + k2] = __float2half_rn(cr[i] * osc);
                ys_im[k1 * Q + k2] = __float2half_rn(ci[i] * osc);
            }
        }
    }
    __syncthreads();
    const __half *ys_re = reinterpret_cast<const __half *>(xa[0]), *ys_im = reinterpret_cast<const __half *>(xa[1]);
    if constexpr (OUT_MODE == 0) {
        // Stage-4 order: bin k at (k % P) * 64 + k / P = k1 * 64 + k2.
        __half2 *ore = reinterpret_cast<__half2 *>(out_re + static_cast<size_t>(gsym) * N);
        __half2 *oim = reinterpret_cast<__half2 *>(out_im + static_cast<size_t>(gsym) * N);
        const __half2 *sr = reinterpret_cast<const __half2 *>(ys_re), *si = reinterpret_cast<const __half2 *>(ys_im);
        for (uint32_t q = t; q < N / 2; q += 256) {
            ore[q] = sr[q];
            oim[q] = si[q];
        }
    } else {
        const uint32_t half_used = a.used / 2;
        __half2 *ore = reinterpret_cast<__half2 *>(out_re + static_cast<size_t>(gsym) * a.pad);
        __half2 *oim = reinterpret_cast<__half2 *>(out_im + static_cast<size_t>(gsym) * a.pad);
        const __half zero = __ushort_as_half(0);
        for (uint32_t q = t; q < a.pad / 2; q += 256) {
            const uint32_t u = 2 * q;
            __half r0 = zero, i0 = zero, r1 = zero, i1 = zero;
            if (u < a.used) {
                const uint32_t k = u < half_used ? N - half_used + u : u - half_used;
                const uint32_t s = (k % P) * Q + k / P;
                r0 = ys_re[s];
                i0 = ys_im[s];
            }
            if (u + 1 < a.used) {
                const uint32_t k = u + 1 < half_used ? N - half_used + u + 1 : u + 1 - half_used;
                const uint32_t s = (k % P) * Q + k / P;
                r1 = ys_re[s];
                i1 = ys_im[s];
            }
            ore[q] = __halves2half2(r0, r1);
            oim[q] = __halves2half2(i0, i1);
        }
    }
}

template <int PREC, int P>
int32_t LaunchP(bool iq_float, int out_mode, uint32_t grid, hipStream_t s, const void *iq, __half *re, __half *im,
                const MfmaDftArgs &a)
{
    const int pi = P == 16 ? 0 : (P == 32 ? 1 : 2);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    uint16_t *wp = nullptr, *wq = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        wp = g_wp[dev];
        wq = g_wq[dev];
    }
    if (wp == nullptr || wq == nullptr) return AIRAN_LAUNCH_FAILED;
    const int pr = PREC == AIRAN_DFT_BF16 ? 0 : 1;
    wp += (static_cast<size_t>(pr) * 3 + pi) * kTabElems;
    wq += static_cast<size_t>(pr) * kTabElems;
    if (iq_float) {
        if (out_mode == 0)
            hipLaunchKernelGGL((OfdmMfmaKernel<PREC, P, true, 0>), dim3(grid), dim3(256), 0, s, iq, re, im, a, wp, wq);
        else
            hipLaunchKernelGGL((OfdmMfmaKernel<PREC, P, true, 1>), dim3(grid), dim3(256), 0, s, iq, re, im, a, wp, wq);
    } else {
        if (out_mode == 0)
            hipLaunchKernelGGL((OfdmMfmaKernel<PREC, P, false, 0>), dim3(grid), dim3(256), 0, s, iq, re, im, a, wp, wq);
        else
            hipLaunchKernelGGL((OfdmMfmaKernel<PREC, P, false, 1>), dim3(grid), dim3(256), 0, s, iq, re, im, a, wp, wq);
    }
    return airan::LaunchStatus();
}

template <int PREC>
int32_t LaunchPrec(uint32_t fft, bool iq_float, int out_mode, uint32_t grid, hipStream_t s, const void *iq,
                   __half *re, __half *im, const MfmaDftArgs &a)
{
    switch (fft) {
    case 1024: return LaunchP<PREC, 16>(iq_float, out_mode, grid, s, iq, re, im, a);
    case 2048: return LaunchP<PREC, 32>(iq_float, out_mode, grid, s, iq, re, im, a);
    case 4096: return LaunchP<PREC, 64>(iq_float, out_mode, grid, s, iq, re, im, a);
    default: return AIRAN_UNSUPPORTED_PROFILE;
    }
}

}  // namespace

namespace airan {

bool MfmaDftSupported(uint32_t fft) { return fft == 1024 || fft == 2048 || fft == 4096; }

int32_t InitMfmaDftTables()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_done[dev]) return AIRAN_OK;
    static uint16_t wp[2][3][kTabElems], wq[2][kTabElems];
    for (int prec = 0; prec < 2; ++prec) {
        for (int pi = 0; pi < 3; ++pi) BuildTable(prec, 16u << pi, 16u << pi, wp[prec][pi]);
        BuildTable(prec, 64, 64, wq[prec]);
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(kWp), wp, sizeof(wp)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(kWq), wq, sizeof(wq)) != hipSuccess ||
        hipGetSymbolAddress(reinterpret_cast<void **>(&g_wp[dev]), HIP_SYMBOL(kWp)) != hipSuccess ||
        hipGetSymbolAddress(reinterpret_cast<void **>(&g_wq[dev]), HIP_SYMBOL(kWq)) != hipSuccess)
        return AIRAN_LAUNCH_FAILED;
    g_done[dev] = true;
    return AIRAN_OK;
}

int32_t LaunchMfmaDft(const Geom &g, uint32_t precision, uint32_t iq_format, uint32_t batch, const void *iq,
                      int out_mode, __half *out_re, __half *out_im, hipStream_t s)
{
    if (!MfmaDftSupported(g.fft)) return AIRAN_UNSUPPORTED_PROFILE;
    if (InitOfdmTables() != AIRAN_OK || InitMfmaDftTables() != AIRAN_OK) return AIRAN_LAUNCH_FAILED;
    MfmaDftArgs a{};
    a.total_symbols = batch * g.nsym;
    a.samples = g.samples;
    a.used = g.used;
    a.pad = g.pad;
    a.out_scale = static_cast<float>((iq_format == AIRAN_IQ_INT16_Q8_8 ? 1.0 / 256.0 : 1.0) /
                                     std::sqrt(static_cast<double>(g.fft)));
    for (uint32_t i = 0; i < g.nsym; ++i) a.sym_start[i] = g.sym_start[i];
    const bool f = iq_format == AIRAN_IQ_COMPLEX64;
    if (precision == AIRAN_DFT_BF16)
        return LaunchPrec<AIRAN_DFT_BF16>(g.fft, f, out_mode, a.total_symbols, s, iq, out_re, out_im, a);
    return LaunchPrec<AIRAN_DFT_FP16>(g.fft, f, out_mode, a.total_symbols, s, iq, out_re, out_im, a);
}

}  // namespace airan
