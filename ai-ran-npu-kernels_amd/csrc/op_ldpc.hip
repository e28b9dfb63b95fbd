// op_ldpc.hip -- BG1 layered normalized min-sum LDPC decoder for gfx950.
//
// Schedule and arithmetic follow fec/ldpc_decode/ldpc_decode_kernel.cpp:
// rows in order (:533-713), per row gather -> subtract old message ->
// min1/min2/sign product -> alpha = 0.75 (:248-300) -> posterior update and
// scatter; final clamp +-5120, int16 conversion with CAST_RINT and bit =
// (int16 value < 0) (:719-780), i.e. a REF-mode posterior in [-0.5, 0)
// decodes to 0.
//
// Mapping: one workgroup per code block; the posterior [68][Z] lives in LDS
// as fp16 (52 KB at Z = 384); thread i owns check i of every row, so within a row the
// Z threads touch a permutation of each column (no LDS write conflicts) and a
// barrier separates rows.  A check's messages are stored compressed
// (min1, min2, index of min1, sign bits): 8 bytes per (row, check) in
// registers when the active rows fit REG_ROWS, otherwise in a global workspace.
#include <algorithm>
#include <mutex>

#include "common.h"
#include "airan_ldpc_bg1.h"
#include "launch.h"
#include "ragged.h"
#include "ratematch.h"

// Posterior of the code block owned by the workgroup ([68][Z] x 16 bit).
extern __shared__ uint16_t airan_ldpc_lam[];

// Message state of airan_ldpc_decode_EnqueueNoAlloc: a fixed device array
// instead of a caller workspace, for callers that cannot allocate.  Each GPU
// the library runs on holds its own copy until the process ends.  32 MiB
// covers the reference's 23-slot transport block (143 code blocks at all 46
// rows, Z = 384: 20.2 MB).
constexpr size_t kNoAllocScratchBytes = size_t{32} << 20;
static __device__ __align__(256) uint2 scratch[kNoAllocScratchBytes / sizeof(uint2)];

namespace {

struct Bg1Dev {
    uint16_t row_start[AIRAN_BG1_ROWS + 1];
    uint8_t col[AIRAN_BG1_EDGES];
    uint16_t shift[AIRAN_BG1_EDGES];
};

__constant__ Bg1Dev kBg1;

// Per-device "constant tables loaded" flags of kBg1 and kLift, guarded by
// one mutex (callers on several host threads may load them at once).
std::mutex g_table_mutex;
bool g_table_loaded[64] = {};

int32_t LoadTable()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    std::lock_guard<std::mutex> lock(g_table_mutex);
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

// Device address of `scratch` per device (hipGetSymbolAddress with that
// device current, at its first airan_InitDevice); the array's name is a host
// stand-in and must not be passed to a kernel.
void *g_scratch_ptr[64] = {};

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

// MODE_REF: the reference's rint(clamp(v, +-5120)) < 0, which for an fp16
// posterior is v < -0.5 (a NaN gives 0).  MODE_FP16 (no reference
// counterpart): v < 0.
template <int MODE>
__device__ inline bool HardBit(const uint16_t *lam, uint32_t a)
{
    const float v = __half2float(__ushort_as_half(lam[a]));
    return MODE == AIRAN_LDPC_MODE_REF ? v < -0.5f : v < 0.0f;
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

// RAGGED: block b decodes code block list[b] of a ragged batch, with that
// block's Z, K, rows and buffer offsets (no lam_out).
template <int MODE, int REG_ROWS, bool RAGGED = false>
__global__ void __launch_bounds__(384) LdpcKernel(const int16_t *__restrict__ lam_in,
                                                  int8_t *__restrict__ bits,
                                                  int16_t *__restrict__ lam_out,
                                                  uint8_t *__restrict__ iters_out,
                                                  uint2 *__restrict__ msg_ws, LdpcDev d,
                                                  const airan::RaggedCb *__restrict__ cbs = nullptr,
                                                  const uint32_t *__restrict__ list = nullptr)
{
    uint16_t *lam = airan_ldpc_lam;
    uint32_t cb = blockIdx.x;
    size_t lam_off, bits_off, ws_off;
    if constexpr (RAGGED) {
        cb = list[blockIdx.x];
        const airan::RaggedCb &c = cbs[cb];
        d.z = c.z;
        d.k = c.k;
        d.rows = c.rows;
        lam_off = c.lam_off;
        bits_off = c.bits_off;
        ws_off = c.ws_off;
    } else {
        lam_off = static_cast<size_t>(cb) * 68u * d.z;
        bits_off = static_cast<size_t>(cb) * d.k;
        ws_off = static_cast<size_t>(cb) * d.rows * d.z;
    }
    const uint32_t Z = d.z, N = 68u * Z;
    const uint32_t i = threadIdx.x;
    const int16_t *src = lam_in + lam_off;
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
    uint2 *ws = msg_ws + ws_off;
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
    int8_t *bo = bits + bits_off;
    for (uint32_t t = i; t < d.k; t += blockDim.x) bo[t] = HardBit<MODE>(lam, t) ? 1 : 0;
    if (!RAGGED && lam_out != nullptr) {
        int16_t *lo = lam_out + lam_off;
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
    std::lock_guard<std::mutex> lock(g_table_mutex);
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

// Two code blocks per workgroup: block "A" in the low and block "B" in the
// high fp16 half of every 32-bit posterior word, so one LDS access, one set
// of edge addresses and one packed fp16 add (v_pk_add_f16) serve both.
// Every message of the thread's check lives in a register (one packed pair
// per edge), so the extrinsic is lam - msg with no reconstruction of the old
// message from a compressed state.  Check i of row R, per half:
//   x_k = lam_k - msg_k; min1/min2 of the magnitudes |x_k| (fp16
//   magnitudes order like unsigned integers), sign bits !(x_k >= 0) (bits
//   above 0x7c00: negative, or a NaN), parity p;
//   msg_k = (|x_k| == |min1| ? 0.75 min2 : 0.75 min1) with sign
//   p ^ sign(x_k), and lam_k = x_k + msg_k.
// This is the compressed-state decoder's arithmetic exactly: an edge whose
// magnitude equals min1 without being min1's index implies min2 == min1, so
// selecting by magnitude instead of index changes no value; and a message's
// sign differs from the recorded sign bit only for a NaN extrinsic, whose
// posterior is NaN and stays NaN (every later extrinsic on it is NaN), so the
// stored message sign is never observed.
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2_t H2(uint32_t x) { return __builtin_bit_cast(h2_t, x); }
__device__ __forceinline__ uint32_t WH(h2_t x) { return __builtin_bit_cast(uint32_t, x); }

// Packed unsigned 16-bit operations on both halves (VOP3P).
__device__ __forceinline__ uint32_t PkMinU16(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ uint32_t PkMaxU16(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_max_u16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// max(a - b, 0) per half.
__device__ __forceinline__ uint32_t PkSubSatU16(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_sub_u16 %0, %1, %2 clamp" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// (a - b) mod 2^16 per half.
__device__ __forceinline__ uint32_t PkSubU16(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_sub_u16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// (a * b + c) mod 2^16 per half.
__device__ __forceinline__ uint32_t PkMadU16(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Signed 16-bit max / min per half.
__device__ __forceinline__ uint32_t PkMaxI16(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_max_i16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ uint32_t PkMinI16(uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_pk_min_i16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
typedef short s16x2 __attribute__((ext_vector_type(2)));

// Per half: 1 if the fp16 bits exceed 0x7c00 (negative, or a NaN), else 0.
__device__ __forceinline__ uint32_t AboveInf2(uint32_t x) { return PkMinU16(PkSubSatU16(x, 0x7c007c00u), 0x00010001u); }

// Per half: NegBit (negative, non-zero, not NaN) as 0 / 1: x - 0x8001 < 0x7c00.
__device__ __forceinline__ uint32_t NegBit2(uint32_t x)
{
    return PkMinU16(PkSubSatU16(0x7c007c00u, PkSubU16(x, 0x80018001u)), 0x00010001u);
}

// Byte offset in LDS of edge e's posterior word for check i: column base +
// 4 ((i + shift) mod Z), i4 = 4 i, i4m = 4 i - 4 Z (mod 2^32).  With ZT != 0,
// Z = ZT and the edge's column and shift are compile-time constants (callers
// unroll over e): the lanes past the wrap (4 i >= 4 (Z - s)) take i4m, the
// others i4 (a compare and a select), and the constant column base + 4 s
// goes to the LDS instruction's offset field, which the hardware adds to the
// address register modulo 2^32 (so i4m + offset lands on 4 (i + s - Z));
// the posterior must start at LDS address 0 for the offset to hold all of
// it.  Otherwise the edge word comes from the per-lifting-size table.
template <uint32_t ZT>
__device__ __forceinline__ uint32_t EdgeByte(const uint32_t *lw, int e, uint32_t i, uint32_t i4, uint32_t i4m,
                                             uint32_t Z)
{
    if constexpr (ZT != 0) {
        const uint32_t s4 = 4u * (static_cast<uint32_t>(AIRAN_BG1_SHIFT[e]) % ZT);
        const uint32_t base = i4 < 4u * ZT - s4 ? i4 : i4m;
        return base + (4u * static_cast<uint32_t>(AIRAN_BG1_COL[e]) * ZT + s4);
    } else {
        return 4u * LiftAddr(lw[e], i, Z);
    }
}

__device__ __forceinline__ uint32_t &LdsWord(uint32_t *lam2, uint32_t byte)
{
    return *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(lam2) + byte);
}

// Check i of row R for both blocks.  min1 / min2 are tracked on the fp16
// magnitudes (with duplicates: two edges at the minimum give min2 == min1),
// starting from two virtual edges of magnitude 0x7c00, which is what the
// keys (|x| << 5 | k) from (0x7c00 << 5) give for the magnitudes.  A row
// with an infinite or NaN extrinsic sets *inf_flag.
template <int R, uint32_t ZT>
__device__ __forceinline__ void PairRow(uint32_t *lam2, uint32_t zi, uint32_t i, uint32_t Z, uint32_t *msg,
                                        uint32_t *inf_flag)
{
    constexpr int DEG = kFastDeg[R], E0 = FastRowStart(R);
    const uint32_t *lw = kLift.w[zi];
    // Opaque per row: with compile-time edges the addresses depend on i
    // alone and would otherwise be hoisted out of the iteration loop, every
    // row's at once, in registers.
    uint32_t i4 = 4u * i, i4m = 4u * i - 4u * Z;
    asm volatile("" : "+v"(i4), "+v"(i4m));
    uint32_t m1 = 0x7c007c00u, m2 = 0x7c007c00u;
    // XOR of the extrinsics (bits 15 / 31: the sign parity of each half) and
    // the largest magnitude per half (above 0x7c00 a NaN, 0x7c00 an infinity).
    uint32_t sx = 0, mx = 0;
    // The row's extrinsics and addresses are not held: the message pass
    // reads each posterior again (no other thread writes it within the row)
    // and recomputes its address, and the loads go kChunk at a time, so that
    // the kernel fits 128 VGPRs without spills and a CU holds two workgroups
    // (12 waves) whatever SIMDs their waves land on.
    constexpr int kChunk = 6;
#pragma unroll
    for (int k = 0; k < DEG; ++k) {
        if (k % kChunk == 0 && k > 0) __builtin_amdgcn_sched_barrier(0);
        const uint32_t x = WH(H2(LdsWord(lam2, EdgeByte<ZT>(lw, E0 + k, i, i4, i4m, Z))) - H2(msg[E0 + k]));
        const uint32_t mg = x & 0x7fff7fffu;
        m2 = PkMinU16(m2, PkMaxU16(m1, mg));
        m1 = PkMinU16(m1, mg);
        sx ^= x;
        mx = PkMaxU16(mx, mg);
    }
    const h2_t q = {static_cast<_Float16>(0.75f), static_cast<_Float16>(0.75f)};
    // The parity counts the edges whose bits exceed 0x7c00 (negative, or a
    // NaN of either sign).  Without a NaN in the row that is the sign bit,
    // so the parity is sx's; a row with an infinite or NaN extrinsic (rare,
    // and sticky once a posterior is NaN) counts the extrinsics again and
    // sets the flag.
    uint32_t par = sx & 0x80008000u;
    if (PkSubSatU16(mx, 0x7bff7bffu) != 0) {
        *inf_flag = 1u;
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < DEG; ++k)
            cnt += AboveInf2(WH(H2(LdsWord(lam2, EdgeByte<ZT>(lw, E0 + k, i, i4, i4m, Z))) - H2(msg[E0 + k])));
        par = (cnt & 1u) << 15 | ((cnt >> 16) & 1u) << 31;
    }
    const uint32_t s1p = WH(q * H2(m1)) | par, s2p = WH(q * H2(m2)) | par;
    const uint32_t d12 = PkSubU16(s2p, s1p);
    // Message pass, kUpd edges at a time: their posteriors are all read
    // before any is written, so the reads of a group are in flight together
    // (within a row every posterior word belongs to one edge of one thread,
    // so no read of the row aliases another edge's write).
    constexpr int kUpd = 6;
#pragma unroll
    for (int k0 = 0; k0 < DEG; k0 += kUpd) {
        constexpr int kMax = kUpd;
        const int n = DEG - k0 < kMax ? DEG - k0 : kMax;
        if (k0 > 0) __builtin_amdgcn_sched_barrier(0);
        uint32_t a[kMax], v[kMax];
#pragma unroll
        for (int k = 0; k < kMax; ++k)
            if (k < n) {
                a[k] = EdgeByte<ZT>(lw, E0 + k0 + k, i, i4, i4m, Z);
                v[k] = LdsWord(lam2, a[k]);
            }
#pragma unroll
        for (int k = 0; k < kMax; ++k)
            if (k < n) {
                const uint32_t x = WH(H2(v[k]) - H2(msg[E0 + k0 + k]));
                // 1 per half where |x_k| == min1, then s1p + that * (s2p - s1p).
                const uint32_t e = PkSubSatU16(0x00010001u, (x & 0x7fff7fffu) ^ m1);
                const uint32_t m = PkMadU16(e, d12, s1p) ^ (x & 0x80008000u);
                msg[E0 + k0 + k] = m;
                LdsWord(lam2, a[k]) = WH(H2(x) + H2(m));
            }
    }
}

template <int R, int ROWS, uint32_t ZT>
__device__ __forceinline__ void PairIteration(uint32_t *lam2, uint32_t zi, uint32_t i, uint32_t Z, uint32_t *msg,
                                              uint32_t *inf_flag)
{
    if constexpr (R < ROWS) {
        if (i < Z) PairRow<R, ZT>(lam2, zi, i, Z, msg, inf_flag);
        __syncthreads();
        PairIteration<R + 1, ROWS, ZT>(lam2, zi, i, Z, msg, inf_flag);
    }
}

// Bit 0: A's check i fails a row's parity; bit 16: B's.  Rows are taken
// one at a time within the 128-VGPR budget of up to 6 active rows, three at
// a time above (their loads in flight together).  A hard decision is
// NegBit (negative, non-zero, not NaN).  No posterior is ever -0 (the
// start values are +0 or non-zero, and round-to-nearest sums and
// differences give -0 only from -0 operands), so without a NaN it is the
// sign bit and a row's parity is the XOR of its posteriors' words (!EXACT);
// EXACT evaluates NegBit per posterior.
template <int R, int ROWS, uint32_t ZT, bool EXACT>
__device__ __forceinline__ uint32_t PairSyndrome(uint32_t *lam2, const uint32_t *lw, uint32_t i, uint32_t Z)
{
    constexpr int G = ROWS <= 6 ? 1 : 3;
    if constexpr (R < ROWS) {
        constexpr int RE = R + G < ROWS ? R + G : ROWS;
        constexpr int E0 = FastRowStart(R), N = FastRowStart(RE) - E0;
        uint32_t i4 = 4u * i, i4m = 4u * i - 4u * Z;
        asm volatile("" : "+v"(i4), "+v"(i4m));
        uint32_t v[N];
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = LdsWord(lam2, EdgeByte<ZT>(lw, E0 + k, i, i4, i4m, Z));
        __builtin_amdgcn_sched_barrier(0);
        uint32_t par = 0;
#pragma unroll
        for (int r = R; r < RE; ++r) {
            uint32_t pr = 0;
#pragma unroll
            for (int k = FastRowStart(r) - E0; k < FastRowStart(r + 1) - E0; ++k) pr ^= EXACT ? NegBit2(v[k]) : v[k];
            par |= EXACT ? pr : (pr >> 15) & 0x00010001u;
        }
        asm volatile("" : "+v"(par) : : "memory");
        return par | PairSyndrome<RE, ROWS, ZT, EXACT>(lam2, lw, i, Z);
    }
    return 0;
}

// Shared words of the pair decoder after its posterior: taken_cb[2],
// syn_bad[8], syn_exact[8], inf_flag[2].
constexpr uint32_t kPairTailWords = 20;

// The pair decoder's LDS: posterior then kPairTailWords.  With Z = ZT known
// at compile time a static array, whose address is a compile-time constant
// (0: the kernel has no other shared variable), so that every edge's column
// base and shift fold into the LDS instructions' offsets; otherwise the
// dynamic allocation.
template <int ROWS, uint32_t ZT>
__device__ __forceinline__ uint32_t *PairLds()
{
    if constexpr (ZT != 0) {
        __shared__ uint32_t buf[(22 + ROWS) * ZT + kPairTailWords];
        return buf;
    } else {
        return reinterpret_cast<uint32_t *>(airan_ldpc_lam);
    }
}

struct LdpcFastDev {
    uint32_t z, zi, k, iters, rows, early_stop, num_cb;
    airan::RateDev rate;  // FUSED: codeword geometry
};

// Decoder input LLR of position t of a block, as the fp16 posterior start
// value (LLR / 256).
__device__ inline uint32_t InputH(int32_t v)
{
    return BitsH(static_cast<h16>(static_cast<float>(v) * (1.0f / 256.0f)));
}

// Loads block cb's posterior start values into half h of lam2 (nl
// positions; cb >= num_cb: a zero block).  FUSED: de-rate-matched from the
// bit-plane codeword LLRs with descriptor desc[cb].
template <bool FUSED>
__device__ inline void PairLoad(uint16_t *lam16, uint32_t h, uint32_t nl, const int16_t *__restrict__ src,
                                const RateMatchDescriptor *__restrict__ desc, const LdpcFastDev &d, uint32_t cb)
{
    const uint32_t i = threadIdx.x;
    if (cb >= d.num_cb) {
        for (uint32_t t = i; t < nl; t += blockDim.x) lam16[2 * t + h] = 0;
        return;
    }
    if constexpr (FUSED) {
        const airan::RateCb rc = airan::RateCbSetup(desc[cb], d.rate);
        if (airan::RateSingle(rc, d.rate)) {
            // Every transmitted bit kk < E lands on its own position (the
            // inverse of RateSingleIndex): q = (q0 + kk) mod (Ncb - filler),
            // position 2Z + (q < fstart ? q : q + filler).  So the block's
            // codeword LLRs -- Qm planes of E/Qm consecutive symbols -- are
            // scattered into LDS after the other positions are set (zero, or
            // +5120 for filler bits).  The symbols of a plane are read in
            // aligned 16-byte chunks of 8 (cw_stride is a multiple of 8) where
            // a chunk lies inside the plane's [rem0, rem0 + eq), and one by one
            // at the plane's two ends (at most 7 each).  A chunk whose 8
            // positions are consecutive (no wrap of q, no filler boundary, all
            // below nl: every chunk but about one per block) is converted with
            // packed int16 clamps and an exact int16 -> fp16 conversion:
            // InputH(clamp(8 x, +-5120)) = fp16(clamp(x, +-640)) / 32, exact.
            const uint32_t two_z = 2u * d.z, fstart = d.rate.kprime - two_z;
            const uint32_t nf = d.rate.ncb - d.rate.filler;
            const int16_t *slot = src + static_cast<size_t>(rc.slot0) * d.rate.qm * d.rate.cw_stride;
            const uint32_t cfirst = (rc.rem0 + 7u) / 8u, clast = (rc.rem0 + rc.eq) / 8u;
            const uint32_t ni = clast > cfirst ? clast - cfirst : 0u;  // whole chunks per plane
            const uint32_t nh = ni ? 8u * cfirst - rc.rem0 : rc.eq;     // symbols before them
            const uint32_t nt = ni ? rc.rem0 + rc.eq - 8u * clast : 0u; // and after them
            const uint32_t total = ni * d.rate.qm;
            const float rcp_ni = 1.0f / static_cast<float>(ni ? ni : 1u);
            // Symbol j (< eq) of plane b, raw codeword LLR v: its position and value.
            auto put = [&](uint32_t b, uint32_t j, int32_t v) {
                uint32_t q = rc.q0 + b * rc.eq + j;
                q = q >= nf ? q - nf : q;
                const uint32_t t = two_z + (q < fstart ? q : q + d.rate.filler);
                if (t >= nl) return;
                v *= 8;
                v = v > 5120 ? 5120 : (v < -5120 ? -5120 : v);
                lam16[2 * t + h] = static_cast<uint16_t>(InputH(v));
            };
            auto chunk_of = [&](uint32_t e, uint32_t &b, uint32_t &cc) {
                b = static_cast<uint32_t>(static_cast<float>(e) * rcp_ni);
                b = b * ni > e ? b - 1u : b;
                b = (b + 1u) * ni <= e ? b + 1u : b;
                cc = cfirst + (e - b * ni);
            };
            constexpr uint32_t kBatch = 2;  // chunk loads of a thread in flight together
            uint4 raw[kBatch];
            auto fetch = [&](uint32_t e0) {
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) {
                    const uint32_t e = e0 + u * blockDim.x;
                    if (e < total) {
                        uint32_t b, cc;
                        chunk_of(e, b, cc);
                        raw[u] = *reinterpret_cast<const uint4 *>(slot + static_cast<size_t>(b) * d.rate.cw_stride +
                                                                  8u * cc);
                    }
                }
            };
            // The first batch is in flight while the other positions are set.
            uint32_t e0 = i;
            fetch(e0);
            const uint32_t fill = InputH(5120);
            for (uint32_t t = i; t < nl; t += blockDim.x)
                lam16[2 * t + h] = static_cast<uint16_t>(
                    t >= two_z + fstart && t < two_z + fstart + d.rate.filler ? fill : 0u);
            __syncthreads();
            // The plane ends, one symbol per thread (at most 14 Qm of them).
            const uint32_t ne = nh + nt;
            for (uint32_t f = i; f < ne * d.rate.qm; f += blockDim.x) {
                const uint32_t b = f / ne, r = f - b * ne;
                const uint32_t j = r < nh ? r : 8u * clast - rc.rem0 + (r - nh);
                put(b, j, slot[static_cast<size_t>(b) * d.rate.cw_stride + rc.rem0 + j]);
            }
            const h2_t scale = {static_cast<_Float16>(1.0f / 32.0f), static_cast<_Float16>(1.0f / 32.0f)};
            while (e0 < total) {
#pragma unroll
                for (uint32_t u = 0; u < kBatch; ++u) {
                    const uint32_t e = e0 + u * blockDim.x;
                    if (e >= total) continue;
                    uint32_t b, cc;
                    chunk_of(e, b, cc);
                    const uint32_t j0 = 8u * cc - rc.rem0;
                    uint32_t q = rc.q0 + b * rc.eq + j0;
                    q = q >= nf ? q - nf : q;
                    const uint32_t t0 = two_z + (q < fstart ? q : q + d.rate.filler);
                    const uint32_t w[4] = {raw[u].x, raw[u].y, raw[u].z, raw[u].w};
                    if (q + 8u <= nf && (q >= fstart || q + 8u <= fstart) && t0 + 8u <= nl) {
#pragma unroll
                        for (uint32_t k = 0; k < 4; ++k) {
                            const uint32_t c = PkMinI16(PkMaxI16(w[k], 0xfd80fd80u), 0x02800280u);
                            const h2_t v = __builtin_convertvector(__builtin_bit_cast(s16x2, c), h2_t) * scale;
                            const uint32_t vb = WH(v);
                            lam16[2 * (t0 + 2u * k) + h] = static_cast<uint16_t>(vb);
                            lam16[2 * (t0 + 2u * k + 1u) + h] = static_cast<uint16_t>(vb >> 16);
                        }
                    } else {
#pragma unroll
                        for (uint32_t x = 0; x < 8; ++x)
                            put(b, j0 + x, static_cast<int16_t>(w[x >> 1] >> (16u * (x & 1u))));
                    }
                }
                e0 += kBatch * blockDim.x;
                if (e0 < total) fetch(e0);
            }
        } else {
            for (uint32_t t = i; t < nl; t += blockDim.x)
                lam16[2 * t + h] = static_cast<uint16_t>(InputH(airan::RateDematchValue(t, rc, d.rate, src)));
        }
    } else {
        // nl is a multiple of 4 (Z is a multiple of 24) and the block 8-byte aligned.
        const uint2 *s4 = reinterpret_cast<const uint2 *>(src + static_cast<size_t>(cb) * 68u * d.z);
        for (uint32_t t = i; t < nl / 4u; t += blockDim.x) {
            const uint2 v = s4[t];
            lam16[8 * t + h] = static_cast<uint16_t>(InputH(static_cast<int16_t>(v.x & 0xffffu)));
            lam16[8 * t + 2 + h] = static_cast<uint16_t>(InputH(static_cast<int16_t>(v.x >> 16)));
            lam16[8 * t + 4 + h] = static_cast<uint16_t>(InputH(static_cast<int16_t>(v.y & 0xffffu)));
            lam16[8 * t + 6 + h] = static_cast<uint16_t>(InputH(static_cast<int16_t>(v.y >> 16)));
        }
    }
}

// Writes block cb (half h of lam2): iteration count, hard bits and, when
// lam_out is set, the int16 posterior (positions beyond the LDS part keep
// their input).  Called by the whole workgroup; ends with a barrier, after
// which half h may be overwritten.
template <bool FUSED>
__device__ inline void PairEmit(const uint32_t *lam2, uint32_t h, uint32_t cb, uint32_t it, uint32_t nl,
                                const int16_t *__restrict__ src, const RateMatchDescriptor *__restrict__ desc,
                                int8_t *__restrict__ bits, int16_t *__restrict__ lam_out,
                                uint8_t *__restrict__ iters_out, const LdpcFastDev &d)
{
    const uint32_t i = threadIdx.x;
    const uint16_t *lam16 = reinterpret_cast<const uint16_t *>(lam2);
    if (iters_out != nullptr && i == 0) iters_out[cb] = static_cast<uint8_t>(it);
    int8_t *bo = bits + static_cast<size_t>(cb) * d.k;
    const uint32_t sh = 16u * h;
    if ((d.k & 3u) == 0) {
        const uint4 *l4 = reinterpret_cast<const uint4 *>(lam2);
        for (uint32_t t = i; t < d.k / 4u; t += blockDim.x) {
            const uint4 v = l4[t];
            reinterpret_cast<uint32_t *>(bo)[t] =
                NegBit(v.x >> sh) | NegBit(v.y >> sh) << 8 | NegBit(v.z >> sh) << 16 | NegBit(v.w >> sh) << 24;
        }
    } else {
        for (uint32_t t = i; t < d.k; t += blockDim.x) bo[t] = static_cast<int8_t>(NegBit(lam16[2 * t + h]));
    }
    if (lam_out != nullptr) {
        const uint32_t N = 68u * d.z;
        const airan::RateCb rc = FUSED ? airan::RateCbSetup(desc[cb], d.rate) : airan::RateCb{};
        int16_t *lo = lam_out + static_cast<size_t>(cb) * N;
        for (uint32_t t = i; t < N; t += blockDim.x) {
            float v;
            if (t < nl) {
                v = static_cast<float>(HBits(lam16[2 * t + h]));
            } else {
                const int16_t x = FUSED ? airan::RateDematchValue(t, rc, d.rate, src)
                                        : src[static_cast<size_t>(cb) * N + t];
                v = static_cast<float>(HBits(InputH(x)));
            }
            v = v > 20.0f ? 20.0f : (v < -20.0f ? -20.0f : v);
            lo[t] = static_cast<int16_t>(rintf(v * 256.0f));
        }
    }
    __syncthreads();
}

// Persistent decoder: the G = gridDim.x workgroups decode the blocks two at
// a time, one per half of the posterior words.  When a half's block is done
// (syndrome clear under early stop, or the iteration limit) its outputs are
// written and the next block is loaded into that half (its messages reset)
// while the other half runs on; a half past the last block holds a zero
// block and is never written.  The workgroups take blocks 0, 1, ... from
// the counter queue[0] (two words, zero at launch and left zero) as halves
// free up -- workgroups that start late hold no block meanwhile -- or, with
// queue null, workgroup b walks its static list
// b, b + G, b + 2 G, ...  Blocks are independent of the pairing
// (the halves never interact), so every block's outputs equal a decode of
// that block alone.  src: int16 decoder LLRs [cb][68 Z] (!FUSED) or
// bit-plane codeword LLRs with per-block descriptors (FUSED).  LDS holds the
// (22 + rows) Z posterior pairs the active rows touch.
//
// The refill code (emit, take, load) exists once, in a rolled loop over the
// two halves that also does the first loads.  Up to 6 active rows the kernel
// is held to 4 waves per SIMD (128 VGPRs): a 6-wave workgroup's waves land
// 2, 2, 1, 1 on the SIMDs, and only at 4 waves per SIMD does a second
// workgroup always fit beside it (at 3 the hardware kept one per CU).
template <bool FUSED, int ROWS, uint32_t ZT>
__global__ void __launch_bounds__(384)
    __attribute__((amdgpu_waves_per_eu(ROWS <= 6 ? 4 : 1, ROWS <= 6 ? 4 : 8)))
    LdpcPairKernel(const int16_t *__restrict__ src, const RateMatchDescriptor *__restrict__ desc,
                   int8_t *__restrict__ bits, int16_t *__restrict__ lam_out, uint8_t *__restrict__ iters_out,
                   LdpcFastDev d, uint32_t *__restrict__ queue)
{
    uint32_t *lam2 = PairLds<ROWS, ZT>();
    const uint32_t Z = d.z, i = threadIdx.x;
    const uint32_t nl = (22u + ROWS) * Z;
    // The few shared words live after the posterior (kPairTailWords).  taken_cb: one slot per half (a wave may
    // still be reading half 0's take when thread 0 takes for half 1:
    // PairLoad need not end with a barrier).  syn_bad / syn_exact: per wave,
    // bit h set when half h fails a check.  inf_flag[it0 & 1] (it0 before
    // the pass's increment): some row of the pass saw an infinite or NaN
    // extrinsic, or (harmless: it only selects the exact syndrome) an
    // earlier pass that used the same word did.
    uint32_t *taken_cb = lam2 + nl, *syn_bad = taken_cb + 2, *syn_exact = syn_bad + 8, *inf_flag = syn_exact + 8;
    if (i < 2) inf_flag[i] = 0;  // before the first barrier
    constexpr int kEdges = FastRowStart(ROWS);
    uint32_t msg[kEdges];
#pragma unroll
    for (int e = 0; e < kEdges; ++e) msg[e] = 0;
    uint32_t taken = 0;
    // Per half: its block (>= num_cb: none) and the iterations run on it.
    uint32_t cur0 = ~0u, cur1 = ~0u, it0 = 0, it1 = 0;
    uint32_t refill = 3u;  // bit h: half h is written (if it holds a block) and refilled
    const uint32_t *lw = kLift.w[d.zi];
    for (;;) {
        if (refill != 0) {
#pragma nounroll
            for (uint32_t h = 0; h < 2; ++h) {
                if (((refill >> h) & 1u) == 0) continue;
                const uint32_t cb = h ? cur1 : cur0;
                // The next block is taken before this one's outputs are
                // written, so the queue atomic overlaps the emission, whose
                // closing barrier also publishes the taken index.
                if (queue != nullptr && i == 0) taken_cb[h] = atomicAdd(queue, 1u);
                if (cb < d.num_cb)
                    PairEmit<FUSED>(lam2, h, cb, h ? it1 : it0, nl, src, desc, bits, lam_out, iters_out, d);
                else
                    __syncthreads();
                const uint32_t next =
                    queue == nullptr ? blockIdx.x + taken * gridDim.x : __builtin_amdgcn_readfirstlane(taken_cb[h]);
                ++taken;
                PairLoad<FUSED>(reinterpret_cast<uint16_t *>(lam2), h, nl, src, desc, d, next);
                if (h) {
                    cur1 = next;
                    it1 = 0;
                } else {
                    cur0 = next;
                    it0 = 0;
                }
                const uint32_t keep = h ? 0x0000ffffu : 0xffff0000u;
#pragma unroll
                for (int e = 0; e < kEdges; ++e) msg[e] &= keep;
            }
            __syncthreads();
            if (cur0 >= d.num_cb && cur1 >= d.num_cb) break;
        }
        // Uniform (it0 is the same in every lane): a scalar, so that the
        // flag's address takes no vector register across the pass.
        const uint32_t slot = __builtin_amdgcn_readfirstlane(it0) & 1u;
        PairIteration<0, ROWS, ZT>(lam2, d.zi, i, Z, msg, inf_flag + slot);
        ++it0;
        ++it1;
        uint32_t ok = 0;  // bit h: half h's syndrome is clear
        if (d.early_stop) {
            // Both halves' flags in one barrier: each wave's ballots into
            // syn_bad[wave] (rewritten only after further barriers).  The
            // parity is the sign-bit XOR unless some thread saw an infinite
            // or NaN extrinsic in the iteration: a NaN posterior is created
            // only from an infinite operand, which the row reading it would
            // have seen, and every posterior the syndrome reads was read and
            // rewritten in the iteration.  Then the syndrome is evaluated
            // again with NegBit (rare: saturated blocks only).
            //
            // inf_flag[slot] was set (if at all) before the last row's
            // barrier and is read below; the other word is cleared here for
            // a later pass (this pass writes and reads only inf_flag[slot]):
            // its last reads were in an earlier pass, before this pass's
            // first row barrier, and its next writes follow this pass's
            // syndrome barrier.
            const uint32_t bad = i < Z ? PairSyndrome<0, ROWS, ZT, false>(lam2, lw, i, Z) : 0u;
            const uint32_t wb = (__ballot((bad & 1u) != 0) != 0 ? 1u : 0u) | (__ballot((bad >> 16) != 0) != 0 ? 2u : 0u);
            if ((i & 63u) == 0) syn_bad[i >> 6] = wb;
            if (i == 0) inf_flag[slot ^ 1u] = 0;
            const bool exact = inf_flag[slot] != 0;
            __syncthreads();
            uint32_t any = 0;
            for (uint32_t w = 0; w < blockDim.x / 64u; ++w) any |= syn_bad[w];
            if (exact) {
                const uint32_t bx = i < Z ? PairSyndrome<0, ROWS, ZT, true>(lam2, lw, i, Z) : 0u;
                const uint32_t wx =
                    (__ballot((bx & 1u) != 0) != 0 ? 1u : 0u) | (__ballot((bx >> 16) != 0) != 0 ? 2u : 0u);
                if ((i & 63u) == 0) syn_exact[i >> 6] = wx;
                __syncthreads();
                any = 0;
                for (uint32_t w = 0; w < blockDim.x / 64u; ++w) any |= syn_exact[w];
            }
            ok = ~any & 3u;
        }
        refill = 0;
        if (cur0 < d.num_cb && ((ok & 1u) || it0 >= d.iters)) refill |= 1u;
        if (cur1 < d.num_cb && ((ok & 2u) || it1 >= d.iters)) refill |= 2u;
    }
    // The last workgroup to finish leaves the queue zeroed for the next
    // launch: every take (queue[0]) precedes its workgroup's count in queue[1].
    if (queue != nullptr && i == 0) {
        __threadfence();
        if (atomicAdd(queue + 1, 1u) == gridDim.x - 1) {
            atomicExch(queue, 0u);
            atomicExch(queue + 1, 0u);
        }
    }
}

// Workgroups resident on the device at once for a pair-decoder launch shape
// (0 when the runtime cannot tell).
uint32_t PairCapacity(const void *kernel, uint32_t threads, size_t lds)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, static_cast<int>(threads), lds) != hipSuccess)
        return 0;
    return static_cast<uint32_t>(std::max(cus, 0)) * static_cast<uint32_t>(std::max(per_cu, 0));
}

// Launches LdpcPairKernel<FUSED, rows, ZT> (rows 4 .. kFastRows); lds: the
// posterior's bytes (the shared words after it are added here; at Z = 384
// the kernel's LDS is static and none is allocated at launch).
template <bool FUSED>
void LaunchPair(uint32_t rows, uint32_t threads, size_t lds, hipStream_t s, const int16_t *src,
                const RateMatchDescriptor *desc, int8_t *bits, int16_t *lam_out, uint8_t *iters_out,
                const LdpcFastDev &d, uint32_t *queue)
{
    using Kern = void (*)(const int16_t *, const RateMatchDescriptor *, int8_t *, int16_t *, uint8_t *, LdpcFastDev,
                          uint32_t *);
    // Z = 384 (every PUSCH profile here): edges folded at compile time.
    const bool z384 = d.z == 384;
    Kern k = rows <= 4 ? (z384 ? LdpcPairKernel<FUSED, 4, 384> : LdpcPairKernel<FUSED, 4, 0>)
           : rows == 5 ? (z384 ? LdpcPairKernel<FUSED, 5, 384> : LdpcPairKernel<FUSED, 5, 0>)
           : rows == 6 ? (z384 ? LdpcPairKernel<FUSED, 6, 384> : LdpcPairKernel<FUSED, 6, 0>)
           : rows == 7 ? (z384 ? LdpcPairKernel<FUSED, 7, 384> : LdpcPairKernel<FUSED, 7, 0>)
                       : (z384 ? LdpcPairKernel<FUSED, 8, 384> : LdpcPairKernel<FUSED, 8, 0>);
    lds = z384 ? 0 : lds + kPairTailWords * sizeof(uint32_t);
    // One resident wave of workgroups, each walking its list of blocks.
    uint32_t grid = (d.num_cb + 1u) / 2u;
    const uint32_t cap = PairCapacity(reinterpret_cast<const void *>(k), threads, lds);
    if (cap > 0) grid = std::min(grid, cap);
    hipLaunchKernelGGL(k, dim3(grid), dim3(threads), lds, s, src, desc, bits, lam_out, iters_out, d, queue);
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

size_t airan_ldpc_decode_NoAllocMaxWorkspaceBytes(void) { return kNoAllocScratchBytes; }

int32_t airan_ldpc_decode_EnqueueNoAlloc(const LdpcDecodeOpArgsV1 *a)
{
    int32_t st = airan_ldpc_decode_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    const size_t need = airan_ldpc_decode_WorkspaceBytes(a->num_cb, a->z, a->active_rows);
    if (need > kNoAllocScratchBytes) return AIRAN_RESOURCE_TOO_SMALL;
    if (need == 0) return airan_ldpc_decode_Enqueue(a, nullptr, 0);
    if ((st = airan_InitDevice()) != AIRAN_OK) return st;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    void *ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_table_mutex);
        ws = g_scratch_ptr[dev];
    }
    if (ws == nullptr) return AIRAN_LAUNCH_FAILED;
    return airan_ldpc_decode_Enqueue(a, ws, need);
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
        const uint32_t rows = a->active_rows;  // validated >= 4
        LdpcFastDev fd{a->z, LiftIndex(a->z), a->k, a->max_iterations, rows, a->early_stop ? 1u : 0u, a->num_cb, {}};
        const size_t fast_lds = static_cast<size_t>(22 + rows) * a->z * sizeof(uint32_t);
        LaunchPair<false>(rows, threads, fast_lds, s, in, nullptr, bits, lo, io, fd, nullptr);
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

int32_t LaunchLdpcRagged(const RaggedCb *cbs, const uint32_t *reg_list, uint32_t n_reg, const uint32_t *ws_list,
                         uint32_t n_ws, uint32_t iters, bool early_stop, const int16_t *lam, int8_t *bits,
                         uint8_t *iters_out, uint2 *ws, hipStream_t s)
{
    int32_t st = LoadTable();
    if (st != AIRAN_OK) return st;
    const LdpcDev d{0, 0, iters, 0, early_stop ? 1u : 0u};
    const size_t lds = static_cast<size_t>(68) * 384 * sizeof(uint16_t);
    if (n_reg > 0)
        hipLaunchKernelGGL((LdpcKernel<AIRAN_LDPC_MODE_FP16, kRegRows, true>), dim3(n_reg), dim3(384), lds, s, lam,
                           bits, nullptr, iters_out, nullptr, d, cbs, reg_list);
    if (n_ws > 0)
        hipLaunchKernelGGL((LdpcKernel<AIRAN_LDPC_MODE_FP16, 0, true>), dim3(n_ws), dim3(384), lds, s, lam, bits,
                           nullptr, iters_out, ws, d, cbs, ws_list);
    return LaunchStatus();
}

int32_t InitLdpcScratch()
{
    // The decoder's constant tables too, so that a caller who ran
    // airan_InitDevice launches nothing but the decoder afterwards.
    int32_t st = LoadTable();
    if (st == AIRAN_OK) st = LoadLiftTable();
    if (st != AIRAN_OK) return st;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return AIRAN_LAUNCH_FAILED;
    std::lock_guard<std::mutex> lock(g_table_mutex);
    if (g_scratch_ptr[dev] != nullptr) return AIRAN_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return AIRAN_LAUNCH_FAILED;
    g_scratch_ptr[dev] = p;
    return AIRAN_OK;
}

int32_t LaunchLdpcFromCodeword(const Geom &g, const AiranTbPlanV1 &p, const RateMatchDescriptor *desc,
                               const int16_t *cw, uint32_t num_cb, uint32_t rows, uint32_t iters, bool early_stop,
                               int8_t *bits, uint8_t *iters_out, uint32_t *queue,
                               hipStream_t s)
{
    if (num_cb == 0 || !FastEligible(AIRAN_LDPC_MODE_FP16, p.Z, rows) || rows < 4 || iters == 0 || iters > 100 ||
        (p.Z % 24) != 0 || p.K == 0 || p.K > 22 * p.Z || p.Ncb <= p.filler || (reinterpret_cast<uintptr_t>(bits) & 3u) ||
        p.E_low >= (1u << 22) || p.E_high >= (1u << 22))
        return AIRAN_UNSUPPORTED_PROFILE;
    int32_t st = LoadLiftTable();
    if (st != AIRAN_OK) return st;
    LdpcFastDev fd{p.Z, LiftIndex(p.Z), p.K, iters, rows, early_stop ? 1u : 0u, num_cb,
                   {g.qm, g.cw_symbols, g.cw_stride, p.Z, p.Kprime, p.filler, p.Ncb, num_cb}};
    const size_t lds = static_cast<size_t>(22 + rows) * p.Z * sizeof(uint32_t);
    LaunchPair<true>(rows, AlignUp(p.Z, 64), lds, s, cw, desc, bits, nullptr, iters_out, fd, queue);
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
