// op_ldpc.hip -- BG1 layered normalized min-sum LDPC decoder for gfx950.
//
// Schedule and arithmetic follow fec/ldpc_decode/ldpc_decode_kernel.cpp:
// rows in order (:533-713), per row gather -> subtract old message ->
// min1/min2/sign product -> alpha = 0.75 (:248-300) -> posterior update and
// scatter; final clamp +-5120 and bit = (lam < 0) (:719-780).
//
// Mapping: one workgroup per code block; the posterior [68][Z] lives in LDS
// as fp16 (52 KB at Z = 384); thread i owns check i of every row, so within a row the
// Z threads touch a permutation of each column (no LDS write conflicts) and a
// barrier separates rows.  A check's messages are stored compressed
// (min1, min2, index of min1, sign bits): 8 bytes per (row, check) in
// registers when the active rows fit REG_ROWS, otherwise in a global workspace.
#include "common.h"
#include "airan_ldpc_bg1.h"
#include "crc.h"
#include "launch.h"
#include "ratematch.h"

// Posterior of the code block owned by the workgroup ([68][Z] x 16 bit).
extern __shared__ uint16_t airan_ldpc_lam[];

namespace {

struct Bg1Dev {
    uint16_t row_start[AIRAN_BG1_ROWS + 1];
    uint8_t col[AIRAN_BG1_EDGES];
    uint16_t shift[AIRAN_BG1_EDGES];
};

__constant__ Bg1Dev kBg1;

bool g_table_loaded[64] = {};

int32_t LoadTable()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    if (g_table_loaded[dev]) return AIRAN_OK;
    Bg1Dev t{};
    for (int i = 0; i <= AIRAN_BG1_ROWS; ++i) t.row_start[i] = AIRAN_BG1_ROW_START[i];
    for (int e = 0; e < AIRAN_BG1_EDGES; ++e) {
        t.col[e] = AIRAN_BG1_COL[e];
        t.shift[e] = AIRAN_BG1_SHIFT[e];
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(kBg1), &t, sizeof(t)) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    g_table_loaded[dev] = true;
    return AIRAN_OK;
}

constexpr int kMaxDeg = AIRAN_BG1_MAX_DEG;

__device__ inline float H16(float v) { return __half2float(__float2half_rn(v)); }

// Message of edge k reconstructed from the compressed state
// (w0 = fp16 min1 | fp16 min2 << 16, w1 = sign bits | index of min1 << 24):
// alpha * sign * magnitude rounded to fp16, then (MODE_REF) stored as int8.
template <int MODE>
__device__ inline float EdgeMsg(uint32_t w0, uint32_t w1, int k)
{
    const uint32_t idx = w1 >> 24;
    const uint32_t signs = w1 & 0x7ffffu;
    const uint16_t mh = static_cast<uint16_t>(static_cast<uint32_t>(k) == idx ? (w0 >> 16) : (w0 & 0xffffu));
    const float mag = __half2float(__ushort_as_half(mh));
    const uint32_t neg = (__popc(signs) ^ (signs >> k)) & 1u;
    const float nm = H16((neg ? -0.75f : 0.75f) * mag);
    if (MODE == AIRAN_LDPC_MODE_REF) {
        float q = rintf(nm);
        return q > 127.0f ? 127.0f : (q < -127.0f ? -127.0f : q);
    }
    return nm;
}

// Check i of row r: gather, subtract old message, min1/min2/signs, scatter
// the updated posterior (fp16 LDS).  Updates the compressed state in place.
template <int MODE>
__device__ inline void ProcessRow(uint16_t *lam, int r, uint32_t i, uint32_t Z, uint32_t &w0,
                                  uint32_t &w1)
{
    const int e0 = kBg1.row_start[r];
    const int deg = kBg1.row_start[r + 1] - e0;
    uint32_t addr[kMaxDeg];
    float ext[kMaxDeg];
    float m1 = INFINITY, m2 = INFINITY;
    uint32_t idx = 0, signs = 0;
#pragma unroll
    for (int k = 0; k < kMaxDeg; ++k) {
        if (k < deg) {
            const uint32_t z = static_cast<uint32_t>(kBg1.shift[e0 + k]) % Z;
            uint32_t p = i + z;
            p = p >= Z ? p - Z : p;
            addr[k] = static_cast<uint32_t>(kBg1.col[e0 + k]) * Z + p;
            const float v = __half2float(__ushort_as_half(lam[addr[k]]));
            ext[k] = H16(v - EdgeMsg<MODE>(w0, w1, k));
            const float a = fabsf(ext[k]);
            signs |= static_cast<uint32_t>(!(ext[k] >= 0.0f)) << k;
            if (a < m1) {
                m2 = m1;
                m1 = a;
                idx = static_cast<uint32_t>(k);
            } else if (a < m2) {
                m2 = a;
            }
        }
    }
    w0 = static_cast<uint32_t>(__half_as_ushort(__float2half_rn(m1))) |
         (static_cast<uint32_t>(__half_as_ushort(__float2half_rn(m2))) << 16);
    w1 = signs | (idx << 24);
#pragma unroll
    for (int k = 0; k < kMaxDeg; ++k) {
        if (k < deg) {
            const uint32_t neg = (__popc(signs) ^ (signs >> k)) & 1u;
            const float mag = (static_cast<uint32_t>(k) == idx) ? m2 : m1;
            const float nm = H16((neg ? -0.75f : 0.75f) * mag);
            lam[addr[k]] = __half_as_ushort(__float2half_rn(ext[k] + nm));
        }
    }
}

template <int MODE>
__device__ inline bool HardBit(const uint16_t *lam, uint32_t a)
{
    return __half2float(__ushort_as_half(lam[a])) < 0.0f;
}

// True when check i of every active row is satisfied by the hard decisions.
template <int MODE>
__device__ inline bool ChecksOk(const uint16_t *lam, int rows, uint32_t i, uint32_t Z)
{
    bool ok = true;
    for (int r = 0; r < rows; ++r) {
        const int e0 = kBg1.row_start[r], e1 = kBg1.row_start[r + 1];
        uint32_t par = 0;
        for (int e = e0; e < e1; ++e) {
            uint32_t p = i + static_cast<uint32_t>(kBg1.shift[e]) % Z;
            p = p >= Z ? p - Z : p;
            par ^= HardBit<MODE>(lam, static_cast<uint32_t>(kBg1.col[e]) * Z + p) ? 1u : 0u;
        }
        ok &= par == 0;
    }
    return ok;
}

struct LdpcDev {
    uint32_t z, k, iters, rows, early_stop;
};

template <int MODE, int REG_ROWS>
__global__ void __launch_bounds__(384) LdpcKernel(const int16_t *__restrict__ lam_in,
                                                  int8_t *__restrict__ bits,
                                                  int16_t *__restrict__ lam_out,
                                                  uint8_t *__restrict__ iters_out,
                                                  uint2 *__restrict__ msg_ws, LdpcDev d)
{
    uint16_t *lam = airan_ldpc_lam;
    const uint32_t Z = d.z, N = 68u * Z;
    const uint32_t cb = blockIdx.x;
    const uint32_t i = threadIdx.x;
    const int16_t *src = lam_in + static_cast<size_t>(cb) * N;
    for (uint32_t t = i; t < N; t += blockDim.x) {
        const float scale = MODE == AIRAN_LDPC_MODE_REF ? 1.0f : 1.0f / 256.0f;
        lam[t] = __half_as_ushort(__float2half_rn(static_cast<float>(src[t]) * scale));
    }
    uint32_t rw0[REG_ROWS > 0 ? REG_ROWS : 1], rw1[REG_ROWS > 0 ? REG_ROWS : 1];
#pragma unroll
    for (int r = 0; r < (REG_ROWS > 0 ? REG_ROWS : 1); ++r) {
        rw0[r] = 0;
        rw1[r] = 0;
    }
    uint2 *ws = msg_ws + static_cast<size_t>(cb) * d.rows * Z;
    if (REG_ROWS == 0 && i < Z)
        for (uint32_t r = 0; r < d.rows; ++r) ws[r * Z + i] = make_uint2(0u, 0u);
    __syncthreads();
    const int rows = static_cast<int>(d.rows);
    uint32_t it = 0;
    for (; it < d.iters; ++it) {
        if (REG_ROWS > 0) {
#pragma unroll
            for (int r = 0; r < REG_ROWS; ++r) {
                if (r < rows) {
                    if (i < Z) ProcessRow<MODE>(lam, r, i, Z, rw0[r], rw1[r]);
                    __syncthreads();
                }
            }
        } else {
            for (int r = 0; r < rows; ++r) {
                if (i < Z) {
                    uint2 s = ws[r * Z + i];
                    ProcessRow<MODE>(lam, r, i, Z, s.x, s.y);
                    ws[r * Z + i] = s;
                }
                __syncthreads();
            }
        }
        if (d.early_stop) {
            const bool ok = i < Z ? ChecksOk<MODE>(lam, rows, i, Z) : true;
            if (__syncthreads_and(ok ? 1 : 0)) {
                ++it;
                break;
            }
        }
    }
    if (iters_out != nullptr && i == 0) iters_out[cb] = static_cast<uint8_t>(it);
    int8_t *bo = bits + static_cast<size_t>(cb) * d.k;
    for (uint32_t t = i; t < d.k; t += blockDim.x) bo[t] = HardBit<MODE>(lam, t) ? 1 : 0;
    if (lam_out != nullptr) {
        int16_t *lo = lam_out + static_cast<size_t>(cb) * N;
        const float clip = MODE == AIRAN_LDPC_MODE_REF ? 5120.0f : 20.0f;
        const float out_scale = MODE == AIRAN_LDPC_MODE_REF ? 1.0f : 256.0f;
        for (uint32_t t = i; t < N; t += blockDim.x) {
            float v = __half2float(__ushort_as_half(lam[t]));
            v = v > clip ? clip : (v < -clip ? -clip : v);
            lo[t] = static_cast<int16_t>(rintf(v * out_scale));
        }
    }
}

constexpr int kRegRows = 8;

// ------------------------------------------------------------------ fast path
// FP16 messages and at most kFastRows active rows (every code rate above
// about 0.6), the same arithmetic as LdpcKernel<MODE_FP16> with:
//  * row degrees fixed at compile time and per-lifting-size edge words
//    (col * Z | (shift mod Z) << 16) in constant memory, so an edge costs a
//    scalar load and three VALU ops of addressing;
//  * fp16 VALU arithmetic on the posterior: a single fp16 add/sub/mul of
//    fp16 operands rounds exactly like the fp32 operation rounded to fp16
//    (fp32 has 24 >= 2*11 + 2 significand bits, so rounding twice is
//    innocuous), and the compressed state keeps the scaled magnitudes
//    fp16(0.75 * min), which are the message magnitudes themselves;
//  * LDS holding only the (22 + rows) columns the active rows touch, so
//    several code blocks share a CU; columns beyond never change.
// With FUSED the decoder input is de-rate-matched on the fly from the
// bit-plane codeword LLRs (the standalone rate-dematch arithmetic).
constexpr int kNumLift = 5;  // Z = 24 << zi
constexpr int kFastRows = 8;
constexpr int kFastDeg[kFastRows] = {19, 19, 19, 19, 3, 8, 9, 7};

constexpr int FastRowStart(int r)
{
    int s = 0;
    for (int i = 0; i < r; ++i) s += kFastDeg[i];
    return s;
}

constexpr int kFastEdges = FastRowStart(kFastRows);

struct Bg1LiftDev {
    uint32_t w[kNumLift][kFastEdges];
};

__constant__ Bg1LiftDev kLift;

bool g_lift_loaded[64] = {};

int32_t LoadLiftTable()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    if (g_lift_loaded[dev]) return AIRAN_OK;
    for (int r = 0; r < kFastRows; ++r)
        if (AIRAN_BG1_ROW_START[r] != FastRowStart(r) ||
            AIRAN_BG1_ROW_START[r + 1] - AIRAN_BG1_ROW_START[r] != kFastDeg[r])
            return AIRAN_LAUNCH_FAILED;
    Bg1LiftDev t{};
    for (int zi = 0; zi < kNumLift; ++zi) {
        const uint32_t z = 24u << zi;
        for (int e = 0; e < kFastEdges; ++e)
            t.w[zi][e] = static_cast<uint32_t>(AIRAN_BG1_COL[e]) * z |
                         (static_cast<uint32_t>(AIRAN_BG1_SHIFT[e]) % z) << 16;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(kLift), &t, sizeof(t)) != hipSuccess) return AIRAN_LAUNCH_FAILED;
    g_lift_loaded[dev] = true;
    return AIRAN_OK;
}

using h16 = _Float16;

__device__ __forceinline__ h16 HBits(uint32_t b) { return __builtin_bit_cast(h16, static_cast<uint16_t>(b)); }
__device__ __forceinline__ uint32_t BitsH(h16 h) { return __builtin_bit_cast(uint16_t, h); }

// Negative and non-zero (fp16 bits), as `value < 0` on the decoded posterior.
__device__ __forceinline__ uint32_t NegBit(uint32_t h) { return (h & 0xffffu) - 0x8001u <= 0x7bffu ? 1u : 0u; }

__device__ __forceinline__ uint32_t LiftAddr(uint32_t w, uint32_t i, uint32_t Z)
{
    uint32_t p = i + (w >> 16);
    p = min(p, p - Z);
    return (w & 0xffffu) + p;
}

// Check i of row R: gather, extrinsic, min1/min2/signs, scatter.  State
// sa = fp16(0.75 min1) | fp16(0.75 min2) << 16, sb = message sign bits
// (bit k: sign of the message on edge k) | index of min1 << 24.
// Branch-free on fp16 bit patterns: magnitudes (|x| bits, +inf = 0x7c00,
// NaN above) order like unsigned integers, so min1 and its first index are
// the minimum of the keys (|x| << 5 | k) and min2 is the median of
// (min1 key, key, min2 key) accumulated over the edges.  The recorded sign
// is !(x >= 0), i.e. bits above 0x7c00 (the posterior is never -0: it starts
// from integer LLRs and IEEE sums give -0 only from -0).  The scatter takes
// the new message's sign from the extrinsic's sign bit, which differs from
// the recorded one only for a NaN extrinsic, whose posterior is NaN anyway.
template <int R>
__device__ __forceinline__ void FastRow(h16 *lam, const uint32_t *lw, uint32_t i, uint32_t Z, uint32_t &sa,
                                        uint32_t &sb)
{
    constexpr int DEG = kFastDeg[R], E0 = FastRowStart(R);
    const uint32_t o1 = sa & 0xffffu, o2 = sa >> 16;
    const uint32_t omsg = sb & 0x7ffffu, oidx = sb >> 24;
    uint32_t w[DEG];
#pragma unroll
    for (int k = 0; k < DEG; ++k) w[k] = lw[E0 + k];
    uint32_t addr[DEG];
    h16 ext[DEG];
    uint32_t m1key = 0x7c00u << 5, m2key = 0x7c00u << 5, sg = 0;
#pragma unroll
    for (int k = 0; k < DEG; ++k) {
        addr[k] = LiftAddr(w[k], i, Z);
        const uint32_t mag = static_cast<uint32_t>(k) == oidx ? o2 : o1;
        const uint32_t sbit = k <= 15 ? omsg << (15 - k) : omsg >> (k - 15);
        const h16 x = lam[addr[k]] - HBits(mag | (sbit & 0x8000u));
        ext[k] = x;
        const uint32_t xb = BitsH(x);
        const uint32_t key = ((xb & 0x7fffu) << 5) | static_cast<uint32_t>(k);
        m2key = min(m2key, max(m1key, key));
        m1key = min(m1key, key);
        sg |= ((0x7c00u - xb) >> 31) << k;
    }
    const uint32_t idx = m1key & 31u;
    const uint32_t s1 = BitsH(static_cast<h16>(0.75f) * HBits(m1key >> 5));
    const uint32_t s2 = BitsH(static_cast<h16>(0.75f) * HBits(m2key >> 5));
    const uint32_t par = __popc(sg) & 1u;
    const uint32_t s1p = s1 | (par << 15), s2p = s2 | (par << 15);
#pragma unroll
    for (int k = 0; k < DEG; ++k) {
        const uint32_t msel = static_cast<uint32_t>(k) == idx ? s2p : s1p;
        lam[addr[k]] = ext[k] + HBits(msel ^ (BitsH(ext[k]) & 0x8000u));
    }
    sa = s1 | (s2 << 16);
    sb = (sg ^ (par ? 0x7ffffu : 0u)) | (idx << 24);
}

template <int R>
__device__ __forceinline__ void FastIteration(h16 *lam, const uint32_t *lw, uint32_t i, uint32_t Z, int rows,
                                              uint32_t *sa, uint32_t *sb)
{
    if constexpr (R < kFastRows) {
        if (R < rows) {
            if (i < Z) FastRow<R>(lam, lw, i, Z, sa[R], sb[R]);
            __syncthreads();
            FastIteration<R + 1>(lam, lw, i, Z, rows, sa, sb);
        }
    }
}

// OR over the active rows of check i's parity (hard decisions).
template <int R>
__device__ __forceinline__ uint32_t FastSyndrome(const uint16_t *lam, const uint32_t *lw, uint32_t i, uint32_t Z,
                                                 int rows)
{
    if constexpr (R < kFastRows) {
        if (R < rows) {
            constexpr int DEG = kFastDeg[R], E0 = FastRowStart(R);
            uint32_t par = 0;
#pragma unroll
            for (int k = 0; k < DEG; ++k) par ^= NegBit(lam[LiftAddr(lw[E0 + k], i, Z)]);
            // One row's loads at a time (keeps the register count at the
            // occupancy target instead of every row's loads in flight).
            asm volatile("" : "+v"(par) : : "memory");
            return par | FastSyndrome<R + 1>(lam, lw, i, Z, rows);
        }
    }
    return 0;
}

struct LdpcFastDev {
    uint32_t z, zi, k, iters, rows, early_stop;
    airan::RateDev rate;  // FUSED: codeword geometry
};

// One workgroup per code block.  src: int16 decoder LLRs [cb][68 Z]
// (!FUSED) or bit-plane codeword LLRs with per-block descriptors (FUSED).
// Six waves per SIMD (80 VGPRs): four Z = 384 blocks per CU.
// With crc.cb_ok set, the epilogue also checks the block's CRC (crc.h).
template <bool FUSED>
__global__ void __launch_bounds__(384, 6) LdpcFastKernel(const int16_t *__restrict__ src,
                                                      const RateMatchDescriptor *__restrict__ desc,
                                                      int8_t *__restrict__ bits, int16_t *__restrict__ lam_out,
                                                      uint8_t *__restrict__ iters_out, LdpcFastDev d,
                                                      airan::CrcFuse crc)
{
    uint16_t *lam16 = airan_ldpc_lam;
    h16 *lam = reinterpret_cast<h16 *>(airan_ldpc_lam);
    const uint32_t Z = d.z, N = 68u * Z, i = threadIdx.x, cb = blockIdx.x;
    const uint32_t nl = (22u + d.rows) * Z;  // posterior entries held in LDS
    const int16_t *blk = FUSED ? src : src + static_cast<size_t>(cb) * N;
    airan::RateCb rc{};
    if (FUSED) rc = airan::RateCbSetup(desc[cb], d.rate);
    auto input = [&](uint32_t t) -> h16 {
        const int16_t v = FUSED ? airan::RateDematchValue(t, rc, d.rate, src) : blk[t];
        return static_cast<h16>(static_cast<float>(v) * (1.0f / 256.0f));
    };
    if (FUSED && airan::RateSingle(rc, d.rate)) {
        // One LLR (or a constant) per position; a thread holds at most
        // kMaxLoad positions (22 + rows <= 30 columns of Z, blockDim >= Z or
        // 64 for Z < 64), so every load of the block is issued before the
        // first is waited on.
        constexpr int kMaxLoad = 30;
        const int16_t *base = src + static_cast<size_t>(rc.slot0) * d.rate.qm * d.rate.cw_stride + rc.rem0;
        int32_t v[kMaxLoad];
#pragma unroll
        for (int u = 0; u < kMaxLoad; ++u) {
            const uint32_t t = i + static_cast<uint32_t>(u) * blockDim.x;
            const int32_t idx = t < nl ? airan::RateSingleIndex(t, rc, d.rate) : -1;
            v[u] = idx == -2 ? 5120 : 0;
            if (idx >= 0) v[u] = base[idx] * 8;
        }
#pragma unroll
        for (int u = 0; u < kMaxLoad; ++u) {
            const uint32_t t = i + static_cast<uint32_t>(u) * blockDim.x;
            const int32_t x = v[u] > 5120 ? 5120 : (v[u] < -5120 ? -5120 : v[u]);
            if (t < nl) lam[t] = static_cast<h16>(static_cast<float>(x) * (1.0f / 256.0f));
        }
    } else if (FUSED) {
        for (uint32_t t = i; t < nl; t += blockDim.x) lam[t] = input(t);
    } else {
        // nl is a multiple of 4 (Z is a multiple of 24) and blk 8-byte aligned.
        const uint2 *s4 = reinterpret_cast<const uint2 *>(blk);
        uint2 *l4 = reinterpret_cast<uint2 *>(lam16);
        for (uint32_t t = i; t < nl / 4u; t += blockDim.x) {
            const uint2 v = s4[t];
            auto cv = [](uint32_t x) {
                return BitsH(static_cast<h16>(static_cast<float>(static_cast<int16_t>(x)) * (1.0f / 256.0f)));
            };
            l4[t] = make_uint2(cv(v.x & 0xffffu) | cv(v.x >> 16) << 16, cv(v.y & 0xffffu) | cv(v.y >> 16) << 16);
        }
    }
    uint32_t sa[kFastRows], sb[kFastRows];
#pragma unroll
    for (int r = 0; r < kFastRows; ++r) sa[r] = sb[r] = 0;
    __syncthreads();
    const uint32_t *lw = kLift.w[d.zi];
    const int rows = static_cast<int>(d.rows);
    uint32_t it = 0;
    for (; it < d.iters; ++it) {
        FastIteration<0>(lam, lw, i, Z, rows, sa, sb);
        if (d.early_stop) {
            const uint32_t bad = i < Z ? FastSyndrome<0>(lam16, lw, i, Z, rows) : 0u;
            if (__syncthreads_and(bad == 0 ? 1 : 0)) {
                ++it;
                break;
            }
        }
    }
    if (iters_out != nullptr && i == 0) iters_out[cb] = static_cast<uint8_t>(it);
    int8_t *bo = bits + static_cast<size_t>(cb) * d.k;
    const bool check = crc.cb_ok != nullptr;
    uint32_t rb = 0, ra = 0;
    if ((d.k & 3u) == 0) {
        const uint2 *l4 = reinterpret_cast<const uint2 *>(lam16);
        // With kprime and L multiples of 4, the power-table entries of bits
        // 4t .. 4t+3 are four consecutive words (in reverse order): one
        // 16-byte load per table instead of a branch and a load per bit.
        const bool vec = check && (crc.kprime & 3u) == 0 && (crc.L & 3u) == 0;
        for (uint32_t t = i; t < d.k / 4u; t += blockDim.x) {
            const uint2 v = l4[t];
            const uint32_t b4 = NegBit(v.x) | NegBit(v.x >> 16) << 8 | NegBit(v.y) << 16 | NegBit(v.y >> 16) << 24;
            reinterpret_cast<uint32_t *>(bo)[t] = b4;
            if (vec) {
                const uint32_t j = 4 * t;
                auto sel = [b4](uint4 p) {
                    return ((b4 & 1u) ? p.w : 0u) ^ ((b4 & 0x100u) ? p.z : 0u) ^ ((b4 & 0x10000u) ? p.y : 0u) ^
                           ((b4 & 0x1000000u) ? p.x : 0u);
                };
                if (crc.crc_cb && j < crc.kprime)
                    rb ^= sel(*reinterpret_cast<const uint4 *>(crc.pow_b + (crc.kprime - 4 - j)));
                if (j < crc.L) ra ^= sel(*reinterpret_cast<const uint4 *>(crc.pow_tb + (crc.L - 4 - j)));
            } else if (check) {
                for (uint32_t q = 0; q < 4; ++q)
                    if ((b4 >> (8 * q)) & 1u) airan::CrcAddBit(crc, 4 * t + q, rb, ra);
            }
        }
    } else {
        for (uint32_t t = i; t < d.k; t += blockDim.x) {
            const uint32_t b = NegBit(lam16[t]);
            bo[t] = static_cast<int8_t>(b);
            if (check && b) airan::CrcAddBit(crc, t, rb, ra);
        }
    }
    if (check) {
        __shared__ uint32_t red[2][8];
        for (int off = 32; off > 0; off >>= 1) {
            rb ^= __shfl_xor(rb, off);
            ra ^= __shfl_xor(ra, off);
        }
        if ((i & 63u) == 0) {
            red[0][i >> 6] = rb;
            red[1][i >> 6] = ra;
        }
        __syncthreads();
        if (i == 0) {
            uint32_t xb = 0, xa = 0;
            for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
                xb ^= red[0][w];
                xa ^= red[1][w];
            }
            airan::CrcFinishBlock(crc, cb, xb, xa);
        }
    }
    if (lam_out != nullptr) {
        int16_t *lo = lam_out + static_cast<size_t>(cb) * N;
        for (uint32_t t = i; t < N; t += blockDim.x) {
            float v = static_cast<float>(t < nl ? lam[t] : input(t));
            v = v > 20.0f ? 20.0f : (v < -20.0f ? -20.0f : v);
            lo[t] = static_cast<int16_t>(rintf(v * 256.0f));
        }
    }
}

bool FastEligible(uint32_t mode, uint32_t z, uint32_t rows)
{
    return mode == AIRAN_LDPC_MODE_FP16 && rows <= static_cast<uint32_t>(kFastRows) && z >= 24 && z <= 384;
}

uint32_t LiftIndex(uint32_t z)
{
    uint32_t zi = 0;
    while ((24u << zi) < z) ++zi;
    return zi;
}

}  // namespace

extern "C" {

int32_t airan_ldpc_decode_ValidateOpArgs(const LdpcDecodeOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(LdpcDecodeOpArgsV1)) || a->lam_in == nullptr ||
        a->bits == nullptr || a->stream == nullptr || a->num_cb == 0)
        return AIRAN_INVALID_ARGUMENT;
    if (a->lam_in == a->bits || a->lam_in == a->lam_out || a->bits == a->lam_out)
        return AIRAN_INVALID_ARGUMENT;
    bool z_ok = false;
    for (uint32_t z = 3; z <= 384; z *= 2) z_ok |= a->z == z;
    if (!z_ok || a->z < 24) return AIRAN_UNSUPPORTED_PROFILE;
    if (a->k == 0 || a->k > 22 * a->z || a->max_iterations == 0 || a->max_iterations > 100 ||
        a->active_rows < 4 || a->active_rows > AIRAN_BG1_ROWS ||
        (a->mode != AIRAN_LDPC_MODE_REF && a->mode != AIRAN_LDPC_MODE_FP16) || a->num_cb > 1u << 20)
        return AIRAN_INVALID_ARGUMENT;
    return AIRAN_OK;
}

size_t airan_ldpc_decode_WorkspaceBytes(uint32_t num_cb, uint32_t z, uint32_t active_rows)
{
    if (active_rows <= static_cast<uint32_t>(kRegRows)) return 0;
    return static_cast<size_t>(num_cb) * active_rows * z * sizeof(uint2);
}

int32_t airan_ldpc_decode_Enqueue(const LdpcDecodeOpArgsV1 *a, void *workspace, size_t workspace_bytes)
{
    int32_t st = airan_ldpc_decode_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    if ((st = LoadTable()) != AIRAN_OK) return st;
    const size_t need = airan_ldpc_decode_WorkspaceBytes(a->num_cb, a->z, a->active_rows);
    if (need > 0 && (workspace == nullptr || workspace_bytes < need)) return AIRAN_RESOURCE_TOO_SMALL;
    LdpcDev d{a->z, a->k, a->max_iterations, a->active_rows, a->early_stop ? 1u : 0u};
    const uint32_t threads = airan::AlignUp(a->z, 64);
    const size_t lds = static_cast<size_t>(68) * a->z * sizeof(uint16_t);
    hipStream_t s = static_cast<hipStream_t>(a->stream);
    const auto *in = static_cast<const int16_t *>(a->lam_in);
    auto *bits = static_cast<int8_t *>(a->bits);
    auto *lo = static_cast<int16_t *>(a->lam_out);
    auto *io = static_cast<uint8_t *>(a->iterations_out);
    auto *ws = static_cast<uint2 *>(workspace);
    const bool aligned = (reinterpret_cast<uintptr_t>(in) & 7u) == 0 && (reinterpret_cast<uintptr_t>(bits) & 3u) == 0;
    if (aligned && FastEligible(a->mode, a->z, a->active_rows)) {
        if ((st = LoadLiftTable()) != AIRAN_OK) return st;
        LdpcFastDev fd{a->z, LiftIndex(a->z), a->k, a->max_iterations, a->active_rows, a->early_stop ? 1u : 0u, {}};
        const size_t fast_lds = static_cast<size_t>(22 + a->active_rows) * a->z * sizeof(uint16_t);
        hipLaunchKernelGGL(LdpcFastKernel<false>, dim3(a->num_cb), dim3(threads), fast_lds, s, in, nullptr, bits, lo,
                           io, fd, airan::CrcFuse{});
        return airan::LaunchStatus();
    }
    const bool reg = a->active_rows <= static_cast<uint32_t>(kRegRows);
    if (a->mode == AIRAN_LDPC_MODE_FP16) {
        if (reg)
            hipLaunchKernelGGL((LdpcKernel<AIRAN_LDPC_MODE_FP16, kRegRows>), dim3(a->num_cb), dim3(threads), lds, s, in, bits, lo, io, ws, d);
        else
            hipLaunchKernelGGL((LdpcKernel<AIRAN_LDPC_MODE_FP16, 0>), dim3(a->num_cb), dim3(threads), lds, s, in, bits, lo, io, ws, d);
    } else {
        if (reg)
            hipLaunchKernelGGL((LdpcKernel<AIRAN_LDPC_MODE_REF, kRegRows>), dim3(a->num_cb), dim3(threads), lds, s, in, bits, lo, io, ws, d);
        else
            hipLaunchKernelGGL((LdpcKernel<AIRAN_LDPC_MODE_REF, 0>), dim3(a->num_cb), dim3(threads), lds, s, in, bits, lo, io, ws, d);
    }
    return airan::LaunchStatus();
}

}  // extern "C"

namespace airan {

int32_t LaunchLdpcFromCodeword(const Geom &g, const AiranTbPlanV1 &p, const RateMatchDescriptor *desc,
                               const int16_t *cw, uint32_t num_cb, uint32_t rows, uint32_t iters, bool early_stop,
                               int8_t *bits, uint8_t *iters_out, const CrcFuse *crc, hipStream_t s)
{
    if (num_cb == 0 || !FastEligible(AIRAN_LDPC_MODE_FP16, p.Z, rows) || rows < 4 || iters == 0 || iters > 100 ||
        (p.Z % 24) != 0 || p.K == 0 || p.K > 22 * p.Z || p.Ncb <= p.filler || (reinterpret_cast<uintptr_t>(bits) & 3u) ||
        p.E_low >= (1u << 22) || p.E_high >= (1u << 22))
        return AIRAN_UNSUPPORTED_PROFILE;
    int32_t st = LoadLiftTable();
    if (st != AIRAN_OK) return st;
    LdpcFastDev fd{p.Z, LiftIndex(p.Z), p.K, iters, rows, early_stop ? 1u : 0u,
                   {g.qm, g.cw_symbols, g.cw_stride, p.Z, p.Kprime, p.filler, p.Ncb, num_cb}};
    const size_t lds = static_cast<size_t>(22 + rows) * p.Z * sizeof(uint16_t);
    hipLaunchKernelGGL(LdpcFastKernel<true>, dim3(num_cb), dim3(AlignUp(p.Z, 64)), lds, s, cw, desc, bits,
                       static_cast<int16_t *>(nullptr), iters_out, fd, crc != nullptr ? *crc : CrcFuse{});
    return LaunchStatus();
}

}  // namespace airan

extern "C" {

// Rows 0..3 always; extension row r >= 4 only when its parity column 22 + r
// holds a transmitted circular-buffer position.
uint32_t airan_ldpc_ActiveRows(const AiranTbPlanV1 *p)
{
    if (p == nullptr || p->Z == 0) return AIRAN_BG1_ROWS;
    const uint32_t nf = p->Ncb - p->filler;
    const uint32_t e_max = p->E_high > p->E_low ? p->E_high : p->E_low;
    if (p->k0 != 0 || e_max >= nf) return AIRAN_BG1_ROWS;
    const uint32_t fstart = p->Kprime - 2 * p->Z;
    const uint32_t last = e_max - 1;
    const uint32_t pos = last < fstart ? last : last + p->filler;
    const uint32_t col = (pos + 2 * p->Z) / p->Z;
    int rows = static_cast<int>(col) - 22 + 1;
    rows = rows < 4 ? 4 : (rows > AIRAN_BG1_ROWS ? AIRAN_BG1_ROWS : rows);
    return static_cast<uint32_t>(rows);
}

}  // extern "C"
