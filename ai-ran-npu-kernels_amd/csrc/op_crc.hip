// op_crc.hip -- code-block CRC24B and transport-block CRC24A/CRC16 checks
// (TS 38.212 5.1) on decoded bits, plus payload packing.
//
// The CRC remainder is linear in the message bits: for an n-bit message
// (data followed by its parity) the remainder is XOR over set bits j of
// x^(L + n - 1 - j) mod g.  Each workgroup owns one code block, each thread a
// strided subset of its bits; the partial remainders are XOR-reduced.  The TB
// remainder of code block r is its local remainder times x^(bits after r)
// mod g, accumulated with atomicXor (order independent) per TB.
#include <mutex>
#include <vector>

#include "common.h"
#include "crc.h"
#include "launch.h"

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

__global__ void __launch_bounds__(256) CbCrcKernel(const int8_t *__restrict__ bits, airan::CrcFuse c, uint32_t K)
{
    const uint32_t cbg = blockIdx.x;            // global CB index over all TBs
    const int8_t *b = bits + static_cast<size_t>(cbg) * K;
    uint32_t rb = 0, ra = 0;
    for (uint32_t j = threadIdx.x; j < c.kprime; j += blockDim.x)
        if (b[j] & 1) airan::CrcAddBit(c, j, rb, ra);
    // Wavefront then block XOR reduction.
    for (int off = 32; off > 0; off >>= 1) {
        rb ^= __shfl_xor(rb, off);
        ra ^= __shfl_xor(ra, off);
    }
    __shared__ uint32_t sb[8], sa[8];
    const uint32_t w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        sb[w] = rb;
        sa[w] = ra;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t xb = 0, xa = 0;
        for (uint32_t k = 0; k < blockDim.x / 64; ++k) {
            xb ^= sb[k];
            xa ^= sa[k];
        }
        airan::CrcFinishBlock(c, cbg, xb, xa);
    }
}

__global__ void __launch_bounds__(256) TbFinishKernel(const uint32_t *__restrict__ tb_acc,
                                                      uint8_t *__restrict__ tb_ok, uint32_t n)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) tb_ok[t] = tb_acc[t] == 0;
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

int32_t CrcFuseSetup(const AiranTbPlanV1 &p, uint8_t *cb_ok, uint32_t *tb_acc, CrcFuse *out)
{
    if (p.C == 0 || p.Kprime > p.K || p.K > 8448 || p.Kprime < p.crc_cb || (p.crc_tb != 24 && p.crc_tb != 16) ||
        (p.crc_cb != 0 && p.crc_cb != 24) ||
        static_cast<uint64_t>(p.C) * (p.Kprime - p.crc_cb) != static_cast<uint64_t>(p.tbs) + p.crc_tb)
        return AIRAN_INVALID_ARGUMENT;
    CrcTables *t = nullptr;
    const int32_t st = EnsureTables(&t);
    if (st != AIRAN_OK) return st;
    *out = CrcFuse{t->pow_b, p.crc_tb == 24 ? t->pow_a : t->pow_16, cb_ok, tb_acc, nullptr, p.C, p.Kprime, p.crc_cb,
                   p.Kprime - p.crc_cb, p.crc_tb == 24 ? kPolyA : kPoly16, p.crc_tb};
    return AIRAN_OK;
}

void LaunchCrcZero(uint32_t *tb_acc, uint32_t num_tbs, hipStream_t s)
{
    hipLaunchKernelGGL(ZeroKernel, dim3(CeilDiv(num_tbs, 256)), dim3(256), 0, s, tb_acc, num_tbs);
}

void LaunchCrcFinish(const uint32_t *tb_acc, uint8_t *tb_ok, uint32_t num_tbs, hipStream_t s)
{
    hipLaunchKernelGGL(TbFinishKernel, dim3(CeilDiv(num_tbs, 256)), dim3(256), 0, s, tb_acc, tb_ok, num_tbs);
}

}  // namespace airan

extern "C" size_t airan_crc_check_WorkspaceBytes(uint32_t num_tbs)
{
    return static_cast<size_t>(num_tbs) * sizeof(uint32_t);
}

extern "C" int32_t airan_crc_check_Enqueue(const AiranTbPlanV1 *p, uint32_t num_tbs, const void *bits,
                                           void *cb_ok, void *tb_ok, void *tb_bytes, void *workspace,
                                           size_t workspace_bytes, void *stream)
{
    if (p == nullptr || bits == nullptr || cb_ok == nullptr || tb_ok == nullptr || stream == nullptr ||
        num_tbs == 0 || p->C == 0 || p->Kprime > p->K || p->K > 8448 || p->Kprime < p->crc_cb ||
        (p->crc_tb != 24 && p->crc_tb != 16) || (p->crc_cb != 0 && p->crc_cb != 24) ||
        static_cast<uint64_t>(p->C) * (p->Kprime - p->crc_cb) != static_cast<uint64_t>(p->tbs) + p->crc_tb)
        return AIRAN_INVALID_ARGUMENT;
    if (workspace == nullptr || workspace_bytes < airan_crc_check_WorkspaceBytes(num_tbs))
        return AIRAN_RESOURCE_TOO_SMALL;
    CrcTables *t = nullptr;
    int32_t st = EnsureTables(&t);
    if (st != AIRAN_OK) return st;
    CrcDev d{p->C, p->K, p->Kprime, p->crc_cb, p->crc_tb, p->tbs,
             p->crc_tb == 24 ? kPolyA : kPoly16, p->crc_tb};
    airan::CrcFuse c{t->pow_b, p->crc_tb == 24 ? t->pow_a : t->pow_16, static_cast<uint8_t *>(cb_ok),
                     static_cast<uint32_t *>(workspace), nullptr, p->C, p->Kprime, p->crc_cb,
                     p->Kprime - p->crc_cb, d.poly_tb, d.len_tb};
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(ZeroKernel, dim3(airan::CeilDiv(num_tbs, 256)), dim3(256), 0, s, static_cast<uint32_t *>(workspace), num_tbs);
    hipLaunchKernelGGL(CbCrcKernel, dim3(num_tbs * p->C), dim3(256), 0, s, static_cast<const int8_t *>(bits), c, p->K);
    hipLaunchKernelGGL(TbFinishKernel, dim3(airan::CeilDiv(num_tbs, 256)), dim3(256), 0, s, static_cast<uint32_t *>(workspace),
                       static_cast<uint8_t *>(tb_ok), num_tbs);
    if (tb_bytes != nullptr) {
        dim3 grid(airan::CeilDiv((p->tbs + 7) / 8, 256), num_tbs);
        hipLaunchKernelGGL(PackKernel, grid, dim3(256), 0, s, static_cast<const int8_t *>(bits),
                           static_cast<uint8_t *>(tb_bytes), d);
    }
    return airan::LaunchStatus();
}
