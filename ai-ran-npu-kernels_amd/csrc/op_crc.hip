// op_crc.hip -- code-block CRC24B and transport-block CRC24A/CRC16 checks
// (TS 38.212 5.1) on decoded bits, plus payload packing.
//
// The CRC remainder is linear in the message bits: for an n-bit message
// (data followed by its parity) the remainder is XOR over set bits j of
// x^(L + n - 1 - j) mod g, a GF(2) matrix times the bit vector.  Thread t
// of the checking kernel owns the 24 positions [24 t, 24 t + 24) of every
// code block and holds that slice of the matrix as masks m[k] (bit 23 - jj of
// m[k] = bit k of position 24 t + jj's term): its share of remainder bit k is
// the parity of (its 24 bits & m[k]); the shares are XOR-reduced over the
// workgroup.  The TB remainder of code block r is its local remainder times
// x^(bits after r) mod g, accumulated with atomicXor (order independent) per
// TB.
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"
#include "crc.h"
#include "launch.h"
#include "ragged.h"

namespace {

constexpr uint32_t kPolyA = 0x864CFBu;   // CRC24A
constexpr uint32_t kPolyB = 0x800063u;   // CRC24B
constexpr uint32_t kPoly16 = 0x1021u;    // CRC16
constexpr uint32_t kMaxPow = 8448 + 64;  // exponents 0..kMaxPow-1 above the CRC length

struct CrcTables {
    uint32_t *pow_a = nullptr;    // x^(24 + e) mod g_A
    uint32_t *pow_b = nullptr;    // x^(24 + e) mod g_B
    uint32_t *pow_16 = nullptr;   // x^(16 + e) mod g_16
};

std::mutex g_crc_mutex;
CrcTables g_tables[64];

uint32_t HostPow(uint32_t poly, uint32_t len, uint32_t e)
{
    // x^(len + e) mod g by repeated shift.
    uint32_t mask = (1u << len) - 1u;
    uint32_t r = 1;  // x^0
    for (uint32_t i = 0; i < len + e; ++i) {
        uint32_t top = (r >> (len - 1)) & 1u;
        r = (r << 1) & mask;
        if (top) r ^= poly;
    }
    return r;
}

int32_t EnsureTables(CrcTables **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    std::lock_guard<std::mutex> lock(g_crc_mutex);
    CrcTables &t = g_tables[dev];
    if (t.pow_a == nullptr) {
        std::vector<uint32_t> a(kMaxPow), b(kMaxPow), c(kMaxPow);
        // iterative: x^(len+e) = x * x^(len+e-1)
        auto fill = [](std::vector<uint32_t> &v, uint32_t poly, uint32_t len) {
            uint32_t mask = (1u << len) - 1u;
            uint32_t r = HostPow(poly, len, 0);
            for (uint32_t e = 0; e < kMaxPow; ++e) {
                v[e] = r;
                uint32_t top = (r >> (len - 1)) & 1u;
                r = (r << 1) & mask;
                if (top) r ^= poly;
            }
        };
        fill(a, kPolyA, 24);
        fill(b, kPolyB, 24);
        fill(c, kPoly16, 16);
        if (hipMalloc(&t.pow_a, kMaxPow * 4) != hipSuccess || hipMalloc(&t.pow_b, kMaxPow * 4) != hipSuccess ||
            hipMalloc(&t.pow_16, kMaxPow * 4) != hipSuccess)
            return AIRAN_LAUNCH_FAILED;
        if (hipMemcpy(t.pow_a, a.data(), kMaxPow * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t.pow_b, b.data(), kMaxPow * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t.pow_16, c.data(), kMaxPow * 4, hipMemcpyHostToDevice) != hipSuccess)
            return AIRAN_LAUNCH_FAILED;
    }
    *out = &t;
    return AIRAN_OK;
}

struct CrcDev {
    uint32_t C, K, kprime, crc_cb, crc_tb, tbs;
    uint32_t poly_tb, len_tb;
};

__global__ void __launch_bounds__(256) ZeroKernel(uint32_t *p, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

constexpr uint32_t kCrcW = 24;           // bits per thread and block
constexpr uint32_t kCrcThreads = 384;     // 24 * 384 >= 8448, the largest K
constexpr uint32_t kCrcMaxLen = 24;
static_assert(airan::kCrcTableBytes >= 2 * kCrcMaxLen * kCrcThreads * sizeof(uint32_t), "mask table size");

// The four 0/1 bytes of w as 4 bits, the first byte highest.
__device__ inline uint32_t Pack4(uint32_t w)
{
    return (((w & 0x01010101u) * 0x08040201u) >> 24) & 0xfu;
}

// Masks of one CRC for positions [w0, w0 + 24) of an n-bit message;
// pw[e] = x^(len + e) mod g.  Positions >= n get no bits.
__device__ inline void CrcMasks(const uint32_t *pw, uint32_t poly, uint32_t len, uint32_t n, uint32_t w0,
                                uint32_t (&m)[kCrcMaxLen])
{
#pragma unroll
    for (uint32_t k = 0; k < kCrcMaxLen; ++k) m[k] = 0;
    if (w0 >= n) return;
    const uint32_t last = min(n - w0, kCrcW) - 1;  // last position that counts
    uint32_t p = pw[n - 1 - (w0 + last)];          // its term; earlier ones are times x
    const uint32_t top = 1u << (len - 1), mask = (top << 1) - 1u;
#pragma unroll
    for (int jj = kCrcW - 1; jj >= 0; --jj) {
        if (static_cast<uint32_t>(jj) > last) continue;
#pragma unroll
        for (uint32_t k = 0; k < kCrcMaxLen; ++k) m[k] |= ((p >> k) & 1u) << (kCrcW - 1 - jj);
        p = ((p << 1) & mask) ^ ((p & top) ? poly : 0u);
    }
}

// Remainder share of the 24 bits v (first position highest) under masks m.
__device__ inline uint32_t CrcApply(uint32_t v, const uint32_t (&m)[kCrcMaxLen])
{
    uint32_t r = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCrcMaxLen; ++k) r |= (__popc(v & m[k]) & 1u) << k;
    return r;
}

// The masks of both CRCs for every thread slot: masks[k][t] (CRC24B rows
// k < 24, then the TB CRC's), built once per plan.
__global__ void __launch_bounds__(kCrcThreads) CrcMaskKernel(airan::CrcJob c, uint32_t *__restrict__ masks)
{
    const uint32_t t = threadIdx.x, w0 = kCrcW * t;
    uint32_t mb[kCrcMaxLen], ma[kCrcMaxLen];
    CrcMasks(c.pow_b, airan::kCrcPolyB, 24, c.crc_cb ? c.kprime : 0u, w0, mb);
    CrcMasks(c.pow_tb, c.poly_tb, c.len_tb, c.L, w0, ma);
#pragma unroll
    for (uint32_t k = 0; k < kCrcMaxLen; ++k) {
        masks[k * kCrcThreads + t] = mb[k];
        masks[(kCrcMaxLen + k) * kCrcThreads + t] = ma[k];
    }
}

constexpr uint32_t kCrcBatch = 4;  // blocks a workgroup checks together

// bits: int8 0/1 per bit, [cb][K].  Persistent: workgroup b checks blocks
// kCrcBatch at a time (b kCrcBatch, ..., then G kCrcBatch further on) with
// kCrcThreads threads, thread t owning positions [24 t, 24 t + 24) and
// holding their masks (CrcMaskKernel) in registers; the batch's loads are in
// flight together and threads 0 .. kCrcBatch - 1 finish its blocks.
__global__ void __launch_bounds__(kCrcThreads) CbCrcKernel(const int8_t *__restrict__ bits, airan::CrcJob c,
                                                          uint32_t K, uint32_t num_cb,
                                                          const uint32_t *__restrict__ masks)
{
    const uint32_t t = threadIdx.x, w0 = kCrcW * t;
    uint32_t mb[kCrcMaxLen], ma[kCrcMaxLen];
#pragma unroll
    for (uint32_t k = 0; k < kCrcMaxLen; ++k) {
        mb[k] = masks[k * kCrcThreads + t];
        ma[k] = masks[(kCrcMaxLen + k) * kCrcThreads + t];
    }
    // Three 8-byte loads when the block base and w0 are 8-aligned and the
    // slice lies inside the block.
    const bool vec = (K % 8u) == 0 && (reinterpret_cast<uintptr_t>(bits) & 7u) == 0 && w0 + kCrcW <= K;
    const bool any = w0 < c.kprime;
    __shared__ uint32_t red[kCrcBatch][2][kCrcThreads / 64];
    for (uint32_t base = blockIdx.x * kCrcBatch; base < num_cb; base += gridDim.x * kCrcBatch) {
        uint32_t v[kCrcBatch];
        if (vec) {
            // Unconditional loads (clamped block, own slice or the block's
            // first), masked at use, so the batch is in flight together.
            uint2 q[kCrcBatch][3];
#pragma unroll
            for (uint32_t u = 0; u < kCrcBatch; ++u) {
                const uint32_t cbu = min(base + u, num_cb - 1);
                const uint2 *p = reinterpret_cast<const uint2 *>(bits + static_cast<size_t>(cbu) * K + (any ? w0 : 0u));
                q[u][0] = p[0];
                q[u][1] = p[1];
                q[u][2] = p[2];
            }
#pragma unroll
            for (uint32_t u = 0; u < kCrcBatch; ++u) {
                v[u] = Pack4(q[u][0].x) << 20 | Pack4(q[u][0].y) << 16 | Pack4(q[u][1].x) << 12 |
                       Pack4(q[u][1].y) << 8 | Pack4(q[u][2].x) << 4 | Pack4(q[u][2].y);
                if (!any || base + u >= num_cb) v[u] = 0;
            }
        } else {
#pragma unroll
            for (uint32_t u = 0; u < kCrcBatch; ++u) {
                v[u] = 0;
                if (!any || base + u >= num_cb) continue;
                const int8_t *b = bits + static_cast<size_t>(base + u) * K + w0;
                for (uint32_t j = 0; j < kCrcW && w0 + j < K; ++j)
                    v[u] |= static_cast<uint32_t>(b[j] & 1) << (kCrcW - 1 - j);
            }
        }
        uint32_t rb[kCrcBatch], ra[kCrcBatch];
#pragma unroll
        for (uint32_t u = 0; u < kCrcBatch; ++u) {
            rb[u] = CrcApply(v[u], mb);
            ra[u] = CrcApply(v[u], ma);
        }
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (uint32_t u = 0; u < kCrcBatch; ++u) {
                rb[u] ^= __shfl_xor(rb[u], off);
                ra[u] ^= __shfl_xor(ra[u], off);
            }
        }
        if ((t & 63u) == 0) {
#pragma unroll
            for (uint32_t u = 0; u < kCrcBatch; ++u) {
                red[u][0][t >> 6] = rb[u];
                red[u][1][t >> 6] = ra[u];
            }
        }
        __syncthreads();
        if (t < kCrcBatch && base + t < num_cb) {
            uint32_t xb = 0, xa = 0;
            for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
                xb ^= red[t][0][w];
                xa ^= red[t][1][w];
            }
            airan::CrcFinishBlock(c, base + t, xb, xa);
        }
        __syncthreads();
    }
}

// The same products on the matrix cores when K is a multiple of 32: the
// remainders of all blocks are R = M B over the integers, then mod 2, with
// M [64][K] the 0/1 term bits (rows 0..23 CRC24B, rows 32.. the TB CRC) and
// B [K][blocks] the 0/1 decoded bytes as they lie in memory (block-major).
// One v_mfma_i32_32x32x16_i8 step takes 16 positions; positions are taken
// in pairs of steps, lane l (half h = l / 32) supplying bytes
// [32 p + 16 h, 32 p + 16 h + 16) of its row (A) or block (B): the first 8
// for step 2p, the next 8 for step 2p+1 (any order of the positions serves,
// A and B agreeing).  The fragment table holds M in that order,
// frag[(tile P + p) 64 + l], so a wave loads 1 KB contiguous per tile and
// pair.
typedef int i32x16_t __attribute__((ext_vector_type(16)));
constexpr uint32_t kCrcCols = 32;   // blocks per workgroup (MFMA columns)
constexpr uint32_t kCrcWaves = 4;   // waves splitting the positions
constexpr uint32_t kCrcUnroll = 6;  // position pairs of a wave in flight together

__global__ void __launch_bounds__(256) CrcFragKernel(airan::CrcJob c, uint32_t P, uint4 *__restrict__ frag)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * P * 64) return;
    const uint32_t l = e % 64, p = (e / 64) % P, tile = e / (64 * P);
    const uint32_t row = l % 32, pos0 = 32 * p + 16 * (l / 32);
    const uint32_t *pw = tile ? c.pow_tb : c.pow_b;
    const uint32_t n = tile ? c.L : (c.crc_cb ? c.kprime : 0u);
    const uint32_t len = tile ? c.len_tb : 24u;
    uint32_t w[4] = {0, 0, 0, 0};
    if (row < len) {
        for (uint32_t j = 0; j < 16; ++j) {
            const uint32_t pos = pos0 + j;
            if (pos < n) w[j / 4] |= ((pw[n - 1 - pos] >> row) & 1u) << (8 * (j % 4));
        }
    }
    frag[e] = make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ inline long Lo8(uint4 v) { return static_cast<long>(static_cast<uint64_t>(v.y) << 32 | v.x); }
__device__ inline long Hi8(uint4 v) { return static_cast<long>(static_cast<uint64_t>(v.w) << 32 | v.z); }

// bits: int8 0/1, [cb][K], K a multiple of 32 and bits 16-byte aligned.
// Workgroup b checks blocks 32 b .. 32 b + 31; wave w takes a contiguous
// quarter of the position pairs (consecutive loads of a lane share cache
// lines); the parities of the waves' partial sums are XOR-combined.
__global__ void __launch_bounds__(kCrcWaves * 64) CrcMfmaKernel(const int8_t *__restrict__ bits, airan::CrcJob c,
                                                               uint32_t K, uint32_t num_cb,
                                                               const uint4 *__restrict__ frag)
{
    const uint32_t l = threadIdx.x & 63u, w = threadIdx.x >> 6, h = l >> 5, col = l & 31u;
    const uint32_t P = K / 32;
    const uint32_t cb = blockIdx.x * kCrcCols + col;
    const bool live = cb < num_cb;
    const uint4 *src = reinterpret_cast<const uint4 *>(bits + static_cast<size_t>(live ? cb : 0u) * K) + h;
    i32x16_t acc_b = {}, acc_a = {};
    const uint32_t span = (P + kCrcWaves - 1) / kCrcWaves, pe = min(P, (w + 1) * span);
    for (uint32_t p0 = w * span; p0 < pe; p0 += kCrcUnroll) {
        // Every load is unconditional (clamped to a valid pair, masked at
        // use): a guarded load would make the compiler wait for it before
        // the next one.
        uint4 b[kCrcUnroll], a0[kCrcUnroll], a1[kCrcUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kCrcUnroll; ++u) {
            const uint32_t p = min(p0 + u, pe - 1);
            b[u] = src[2 * p];
            a0[u] = frag[p * 64 + l];
            a1[u] = frag[(P + p) * 64 + l];
        }
#pragma unroll
        for (uint32_t u = 0; u < kCrcUnroll; ++u) {
            if (!live || p0 + u >= pe) b[u] = make_uint4(0, 0, 0, 0);
            acc_b = __builtin_amdgcn_mfma_i32_32x32x16_i8(Lo8(a0[u]), Lo8(b[u]), acc_b, 0, 0, 0);
            acc_b = __builtin_amdgcn_mfma_i32_32x32x16_i8(Hi8(a0[u]), Hi8(b[u]), acc_b, 0, 0, 0);
            acc_a = __builtin_amdgcn_mfma_i32_32x32x16_i8(Lo8(a1[u]), Lo8(b[u]), acc_a, 0, 0, 0);
            acc_a = __builtin_amdgcn_mfma_i32_32x32x16_i8(Hi8(a1[u]), Hi8(b[u]), acc_a, 0, 0, 0);
        }
    }
    // Accumulator r of lane l is row 8 (r / 4) + 4 h + r % 4 of column col.
    uint32_t pb = 0, pa = 0;
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) {
        const uint32_t row = 8 * (r / 4) + 4 * h + r % 4;
        pb |= (static_cast<uint32_t>(acc_b[r]) & 1u) << row;
        pa |= (static_cast<uint32_t>(acc_a[r]) & 1u) << row;
    }
    pb |= __shfl_xor(pb, 32);
    pa |= __shfl_xor(pa, 32);
    __shared__ uint32_t red[kCrcWaves][2][kCrcCols];
    if (h == 0) {
        red[w][0][col] = pb;
        red[w][1][col] = pa;
    }
    __syncthreads();
    if (threadIdx.x < kCrcCols && live) {
        uint32_t xb = 0, xa = 0;
        for (uint32_t v = 0; v < kCrcWaves; ++v) {
            xb ^= red[v][0][col];
            xa ^= red[v][1][col];
        }
        airan::CrcFinishBlock(c, cb, xb, xa);
    }
}

// tb_ok of TB blockIdx.x (one wave): the XOR of its C blocks' shifted
// remainder parts is zero.
__global__ void __launch_bounds__(64) TbFinishKernel(const uint32_t *__restrict__ tb_part,
                                                     uint8_t *__restrict__ tb_ok, uint32_t C)
{
    const uint32_t tb = blockIdx.x, l = threadIdx.x;
    const uint32_t *p = tb_part + static_cast<size_t>(tb) * C;
    uint32_t x = 0;
    for (uint32_t r = l; r < C; r += 64) x ^= p[r];
    for (int off = 32; off > 0; off >>= 1) x ^= __shfl_xor(x, off);
    if (l == 0) tb_ok[tb] = x == 0;
}
// Payload byte b of TB tb: bits 8b..8b+7 (MSB first) of the TB bit stream.
__global__ void __launch_bounds__(256) PackKernel(const int8_t *__restrict__ bits,
                                                  uint8_t *__restrict__ out, CrcDev d)
{
    const uint32_t nbytes = (d.tbs + 7) / 8;
    const uint32_t byte = blockIdx.x * blockDim.x + threadIdx.x;
    if (byte >= nbytes) return;
    const uint32_t tb = blockIdx.y;
    const uint32_t L = d.kprime - d.crc_cb;
    uint32_t v = 0;
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t t = byte * 8 + k;
        uint32_t bit = 0;
        if (t < d.tbs) {
            const uint32_t r = t / L, j = t - r * L;
            bit = bits[(static_cast<size_t>(tb) * d.C + r) * d.K + j] & 1;
        }
        v |= bit << (7 - k);
    }
    out[static_cast<size_t>(tb) * nbytes + byte] = static_cast<uint8_t>(v);
}


// Ragged batch: one workgroup per code block of any grant.  The block's
// CRC24B remainder over its K' bits and the remainder of its part of the
// TB (its first K' - crc_cb bits) as the last bits of a message, shifted by
// x^(bits after it in its TB) mod g_TB (the host's per-block `xpow`).
__global__ void __launch_bounds__(256) RaggedCrcKernel(const int8_t *__restrict__ bits,
                                                       const airan::RaggedCb *__restrict__ cbs,
                                                       const uint32_t *__restrict__ pow_b,
                                                       const uint32_t *__restrict__ pow_a,
                                                       const uint32_t *__restrict__ pow_16,
                                                       uint8_t *__restrict__ cb_ok, uint32_t *__restrict__ tb_part)
{
    const airan::RaggedCb c = cbs[blockIdx.x];
    const int8_t *b = bits + c.bits_off;
    const uint32_t n = c.kprime, L = c.kprime - c.crc_cb;
    const uint32_t *pw = c.len_tb == 24 ? pow_a : pow_16;
    uint32_t rb = 0, ra = 0;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        if (b[j] != 0) {
            if (c.crc_cb) rb ^= pow_b[n - 1 - j];
            if (j < L) ra ^= pw[L - 1 - j];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        rb ^= __shfl_xor(rb, off);
        ra ^= __shfl_xor(ra, off);
    }
    __shared__ uint32_t part[2][4];
    if ((threadIdx.x & 63u) == 0) {
        part[0][threadIdx.x / 64] = rb;
        part[1][threadIdx.x / 64] = ra;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rb = part[0][0] ^ part[0][1] ^ part[0][2] ^ part[0][3];
        ra = part[1][0] ^ part[1][1] ^ part[1][2] ^ part[1][3];
        cb_ok[blockIdx.x] = c.crc_cb ? static_cast<uint8_t>(rb == 0) : static_cast<uint8_t>(1);
        tb_part[blockIdx.x] = airan::CrcMulMod(ra, c.xpow, c.poly_tb, c.len_tb);
    }
}

// tb_ok[t] = 1 when the parts of TB t's code blocks (first[t] .. first[t] +
// count[t] - 1) XOR to zero.
__global__ void __launch_bounds__(64) RaggedCrcFinishKernel(const uint32_t *__restrict__ tb_part,
                                                            const uint32_t *__restrict__ first,
                                                            const uint32_t *__restrict__ count,
                                                            uint8_t *__restrict__ tb_ok, uint32_t num_tbs)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tbs) return;
    uint32_t x = 0;
    for (uint32_t r = 0; r < count[t]; ++r) x ^= tb_part[first[t] + r];
    tb_ok[t] = x == 0 ? 1 : 0;
}

}  // namespace

namespace airan {

uint32_t CrcXpowHost(const AiranTbPlanV1 &p, uint32_t r)
{
    const uint32_t poly = p.crc_tb == 24 ? kPolyA : kPoly16, len = p.crc_tb;
    const uint32_t mask = (1u << len) - 1u;
    uint64_t e = static_cast<uint64_t>(p.C - 1 - r) * (p.Kprime - p.crc_cb);
    // x^e mod g by square-and-multiply on the host.
    auto mul = [&](uint32_t a, uint32_t b) {
        uint32_t res = 0;
        for (int i = static_cast<int>(len) - 1; i >= 0; --i) {
            const uint32_t top = (res >> (len - 1)) & 1u;
            res = (res << 1) & mask;
            if (top) res ^= poly;
            if ((b >> i) & 1u) res ^= a;
        }
        return res;
    };
    uint32_t x = 1u, base = 2u;
    while (e != 0) {
        if (e & 1u) x = mul(x, base);
        base = mul(base, base);
        e >>= 1;
    }
    return x;
}

uint32_t CrcTbPoly(uint32_t crc_tb) { return crc_tb == 24 ? kPolyA : kPoly16; }

int32_t LaunchCrcRagged(const RaggedCb *cbs, uint32_t num_cb, const int8_t *bits, uint8_t *cb_ok, uint32_t *tb_part,
                        const uint32_t *tb_first, const uint32_t *tb_count, uint8_t *tb_ok, uint32_t num_tbs,
                        hipStream_t s)
{
    CrcTables *t = nullptr;
    const int32_t st = EnsureTables(&t);
    if (st != AIRAN_OK) return st;
    if (num_cb == 0 || num_tbs == 0) return AIRAN_INVALID_ARGUMENT;
    hipLaunchKernelGGL(RaggedCrcKernel, dim3(num_cb), dim3(256), 0, s, bits, cbs, t->pow_b, t->pow_a, t->pow_16,
                       cb_ok, tb_part);
    hipLaunchKernelGGL(RaggedCrcFinishKernel, dim3(CeilDiv(num_tbs, 64)), dim3(64), 0, s, tb_part, tb_first,
                       tb_count, tb_ok, num_tbs);
    return LaunchStatus();
}

int32_t CrcJobSetup(const AiranTbPlanV1 &p, uint8_t *cb_ok, uint32_t *tb_part, CrcJob *out)
{
    if (p.C == 0 || p.C > kCrcMaxCb || p.Kprime > p.K || p.K > 8448 || p.Kprime < p.crc_cb ||
        (p.crc_tb != 24 && p.crc_tb != 16) ||
        (p.crc_cb != 0 && p.crc_cb != 24) ||
        static_cast<uint64_t>(p.C) * (p.Kprime - p.crc_cb) != static_cast<uint64_t>(p.tbs) + p.crc_tb)
        return AIRAN_INVALID_ARGUMENT;
    CrcTables *t = nullptr;
    const int32_t st = EnsureTables(&t);
    if (st != AIRAN_OK) return st;
    *out = CrcJob{t->pow_b, p.crc_tb == 24 ? t->pow_a : t->pow_16, cb_ok, tb_part, nullptr, p.C, p.Kprime, p.crc_cb,
                   p.Kprime - p.crc_cb, p.crc_tb == 24 ? kPolyA : kPoly16, p.crc_tb};
    return AIRAN_OK;
}

void LaunchCrcZero(uint32_t *words, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(ZeroKernel, dim3(CeilDiv(n, 256)), dim3(256), 0, s, words, n);
}

// The matrix-core path needs whole position pairs and aligned 16-byte loads.
bool CrcMfmaEligible(const void *bits, uint32_t K, const void *tables)
{
    return K % 32 == 0 && ((reinterpret_cast<uintptr_t>(bits) | reinterpret_cast<uintptr_t>(tables)) & 15u) == 0;
}

void LaunchCrcTables(const CrcJob &c, uint32_t K, void *tables, hipStream_t s)
{
    if (K % 32 == 0) {
        const uint32_t P = K / 32;
        static_assert(kCrcTableBytes >= 2 * (8448 / 32) * 64 * sizeof(uint4), "fragment table size");
        hipLaunchKernelGGL(CrcFragKernel, dim3(CeilDiv(2 * P * 64, 256)), dim3(256), 0, s, c, P,
                           static_cast<uint4 *>(tables));
    }
    // The masks of the VALU path follow the fragment table (used when the
    // bits are not 16-byte aligned).
    hipLaunchKernelGGL(CrcMaskKernel, dim3(1), dim3(kCrcThreads), 0, s, c,
                       reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(tables) + kCrcFragBytes));
}

void LaunchCbCrc(const int8_t *bits, const CrcJob &c, uint32_t K, uint32_t num_cb, const void *tables,
                 hipStream_t s)
{
    if (CrcMfmaEligible(bits, K, tables)) {
        hipLaunchKernelGGL(CrcMfmaKernel, dim3(CeilDiv(num_cb, kCrcCols)), dim3(kCrcWaves * 64), 0, s, bits, c, K,
                           num_cb, static_cast<const uint4 *>(tables));
        return;
    }
    const uint32_t *masks = reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(tables) + kCrcFragBytes);
    // One resident wave of workgroups (the masks are built once per workgroup).
    uint32_t cap = 1024;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(CbCrcKernel),
                                                     static_cast<int>(kCrcThreads), 0) == hipSuccess &&
        cus > 0 && per_cu > 0)
        cap = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
    const uint32_t grid = std::max(1u, std::min(CeilDiv(num_cb, kCrcBatch), cap));
    hipLaunchKernelGGL(CbCrcKernel, dim3(grid), dim3(kCrcThreads), 0, s, bits, c, K, num_cb, masks);
}

void LaunchCrcFinish(const uint32_t *tb_part, uint8_t *tb_ok, uint32_t num_tbs, uint32_t C, hipStream_t s)
{
    hipLaunchKernelGGL(TbFinishKernel, dim3(num_tbs), dim3(64), 0, s, tb_part, tb_ok, C);
}

}  // namespace airan

extern "C" size_t airan_crc_check_WorkspaceBytes(uint32_t num_tbs)
{
    // Per-block TB remainder parts, then the check masks.
    return airan::CrcTableOffset(num_tbs) + airan::kCrcTableBytes;
}

extern "C" int32_t airan_crc_check_Enqueue(const AiranTbPlanV1 *p, uint32_t num_tbs, const void *bits,
                                           void *cb_ok, void *tb_ok, void *tb_bytes, void *workspace,
                                           size_t workspace_bytes, void *stream)
{
    if (p == nullptr || bits == nullptr || cb_ok == nullptr || tb_ok == nullptr || stream == nullptr ||
        num_tbs == 0 || p->C == 0 || p->C > airan::kCrcMaxCb || p->Kprime > p->K || p->K > 8448 ||
        p->Kprime < p->crc_cb || (p->crc_tb != 24 && p->crc_tb != 16) || (p->crc_cb != 0 && p->crc_cb != 24) ||
        static_cast<uint64_t>(p->C) * (p->Kprime - p->crc_cb) != static_cast<uint64_t>(p->tbs) + p->crc_tb)
        return AIRAN_INVALID_ARGUMENT;
    if (workspace == nullptr || workspace_bytes < airan_crc_check_WorkspaceBytes(num_tbs))
        return AIRAN_RESOURCE_TOO_SMALL;
    CrcTables *t = nullptr;
    int32_t st = EnsureTables(&t);
    if (st != AIRAN_OK) return st;
    CrcDev d{p->C, p->K, p->Kprime, p->crc_cb, p->crc_tb, p->tbs,
             p->crc_tb == 24 ? kPolyA : kPoly16, p->crc_tb};
    airan::CrcJob c{t->pow_b, p->crc_tb == 24 ? t->pow_a : t->pow_16, static_cast<uint8_t *>(cb_ok),
                     static_cast<uint32_t *>(workspace), nullptr, p->C, p->Kprime, p->crc_cb,
                     p->Kprime - p->crc_cb, d.poly_tb, d.len_tb};
    hipStream_t s = static_cast<hipStream_t>(stream);
    void *tables = static_cast<uint8_t *>(workspace) + airan::CrcTableOffset(num_tbs);
    airan::LaunchCrcTables(c, p->K, tables, s);
    airan::LaunchCbCrc(static_cast<const int8_t *>(bits), c, p->K, num_tbs * p->C, tables, s);
    airan::LaunchCrcFinish(static_cast<const uint32_t *>(workspace), static_cast<uint8_t *>(tb_ok), num_tbs, p->C, s);
    if (tb_bytes != nullptr) {
        dim3 grid(airan::CeilDiv((p->tbs + 7) / 8, 256), num_tbs);
        hipLaunchKernelGGL(PackKernel, grid, dim3(256), 0, s, static_cast<const int8_t *>(bits),
                           static_cast<uint8_t *>(tb_bytes), d);
    }
    return airan::LaunchStatus();
}
