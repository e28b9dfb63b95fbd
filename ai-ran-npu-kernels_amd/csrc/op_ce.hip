// op_ce.hip -- LMMSE channel estimation from DMRS least-squares observations.
//
// Estimator (channel_est/channel_est_lmmse_mimo/scripts/channel_est_lmmse_ref.py:
// 50-117): exponential-PDP frequency prior R_f(dk) = 1 / (1 + j 2 pi dk scs
// tau), top-`rank` eigenpairs of the observation covariance, the
// Sionna-style rescale of ref.py:109-116 and per-subcarrier time Wiener
// weights with R_t = J0(2 pi f_d dt) (ref.py:75-88).
//
// Observation models.  A layer alone on its comb is observed at comb + 2p
// (one "comb model").  For two layers sharing a comb with FD-OCC2 the
// reference averages each pair (observation_model_ref.py:41-79), which
// assumes the channel flat over the pair and leaks the partner layer's
// channel into the estimate (about -25 dB at 300 ns delay spread).  Here the
// pair is estimated jointly instead, without that assumption: with
// z_m = y_m r_m* on comb + 2m, z_{2p} = h0 + h1 and z_{2p+1} = h0 - h1
// (FD-OCC signs +,+ and +,-), recovered exactly from the two layers' h_ls
// (z_{2p} = h_ls0[p] + h_ls1[p], z_{2p+1} = h_ls0[p] - h_ls1[p]).  For
// independent layers of equal power g+ = h0 + h1 and g- = h0 - h1 are
// uncorrelated, each observed on a comb of spacing 4 (offsets comb and
// comb + 2); each gets its own comb-model filter and h0,1 = (g+ +- g-) / 2.
//
// The eigen-factors of every observation model are built once per grant
// shape in FP64 on the device (covariances by a kernel, rocSOLVER zheevd,
// rocBLAS zgemm); the noise-dependent parts are applied per slot, with the
// noise estimated from the energy outside the kept eigenspace.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "cgemm.h"
#include "launch.h"

using airan::Geom;
using airan::f32x16_t;

namespace {

constexpr uint32_t kWeightModelMagic = 0x43455731u;  // "CEW1"
constexpr int kMaxRank = 512;
constexpr uint32_t kMaxModels = 4;
constexpr uint32_t kMaxObs = 2048;         // observations per model (used / 2 at 273 PRB: 1638)
constexpr double kMinNoiseDims = 32.0;     // residual dimensions for the subspace noise estimate
// The projection's observation axis is cut into this many fixed ranges of
// whole 16-observation blocks (partial products summed in a fixed order),
// independent of the batch, so results do not depend on how many slots share
// a launch.
constexpr uint32_t kProjSplits = 4;
// Projection GEMM: a workgroup computes 64 columns (slot, receive antenna,
// DMRS symbol) x kProjKTile eigenvectors of one model and split.
constexpr uint32_t kProjKTile = 128;
// Interpolation GEMM: a workgroup computes kInterpTile columns (slot,
// receive antenna, DMRS symbol) x kInterpTile subcarriers.
constexpr uint32_t kInterpTile = 128;

// One observation model: n observations at offset + step * p.
struct CeModelFactors {
    uint32_t n = 0;      // observations
    uint32_t step = 0;   // subcarrier spacing of the observations (2 or 4)
    uint32_t offset = 0;
    uint32_t in_a = 0;   // z = h_ls[in_a] + sign_b * h_ls[in_b]
    uint32_t in_b = 0;
    float sign_b = 0.0f;
    uint32_t rank = 0;   // stored eigenpairs (multiple of 16; columns >= kept are zero)
    uint32_t kept = 0;   // eigenpairs actually kept
    float2 *bt = nullptr;  // [n][rank]    bt[p][k] = conj(U[p][k])
    float *a2 = nullptr;   // [rank][used] |(R_hp U)[sc][k]|^2, subcarrier contiguous
    uint16_t *a16 = nullptr;  // R_hp U, 2-part fragment planes: rows sc < AlignUp(used, kInterpTile), K = rank
    uint16_t *w16 = nullptr;  // conj(U) transposed, 3-part fragment planes: rows k < AlignUp(rank, kProjKTile),
                              // K = AlignUp(n, 16) observations (zero beyond n)
    float *lam = nullptr;  // [rank] eigenvalues (1 for padded columns)
};

}  // namespace

struct LmmseFilter {
    int device;
    Geom g;
    LmmseFilterParamsV1 params;
    uint32_t nm;                       // observation models
    CeModelFactors model[kMaxModels];
    int32_t layer_model[4][2];         // models of layer l (second = -1 for a comb model)
    float layer_sign[4];               // h_l = (g_a + sign g_b) / 2 for an FD-OCC pair
    float rt[14][4];                   // J0(2 pi f_d (s - d_j) T) per symbol and DMRS symbol
    float rtt[4][4];                   // J0 between DMRS symbols
};

namespace {

__device__ inline double2 RfDev(double dsc, double scs_tau)
{
    // 1 / (1 + j a) = (1 - j a) / (1 + a^2)
    const double a = 2.0 * M_PI * dsc * scs_tau;
    const double den = 1.0 / (1.0 + a * a);
    return make_double2(den, -a * den);
}

struct ObsGeom {
    uint32_t n, step, offset, used;
    double scs_tau;
};

// R_pp, column-major n x n (element (p, q) at p + q n).
__global__ void __launch_bounds__(256) BuildRppKernel(double2 *__restrict__ rpp, ObsGeom o)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = blockIdx.y;
    if (p >= o.n) return;
    const double dk = (static_cast<double>(p) - static_cast<double>(q)) * o.step;
    rpp[p + static_cast<size_t>(q) * o.n] = RfDev(dk, o.scs_tau);
}

// R_hp, column-major used x n.
__global__ void __launch_bounds__(256) BuildRhpKernel(double2 *__restrict__ rhp, ObsGeom o)
{
    const uint32_t sc = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t p = blockIdx.y;
    if (sc >= o.used) return;
    const double dk = static_cast<double>(sc) - static_cast<double>(o.offset + o.step * p);
    rhp[sc + static_cast<size_t>(p) * o.used] = RfDev(dk, o.scs_tau);
}

// Keep the top `rank` eigenpairs (rocSOLVER returns ascending order) in
// descending order: bt[p][k] = conj(V[p][n-1-k]), at[k][sc] = (R_hp V)[sc][n-1-k].
__global__ void __launch_bounds__(256) PackBtKernel(const double2 *__restrict__ v, float2 *__restrict__ bt,
                                                    uint32_t n, uint32_t rank, uint32_t kept)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t p = blockIdx.y;
    if (k >= rank) return;
    float2 o = make_float2(0.0f, 0.0f);
    if (k < kept) {
        const double2 e = v[p + static_cast<size_t>(n - 1 - k) * n];
        o = make_float2(static_cast<float>(e.x), static_cast<float>(-e.y));
    }
    bt[static_cast<size_t>(p) * rank + k] = o;
}

__global__ void __launch_bounds__(256) PackAtKernel(const double2 *__restrict__ av, float2 *__restrict__ at,
                                                    uint32_t used, uint32_t rank, uint32_t kept)
{
    const uint32_t sc = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = blockIdx.y;
    if (sc >= used) return;
    float2 o = make_float2(0.0f, 0.0f);
    if (k < kept) {
        // av = R_hp * V_top (column-major used x kept, columns ascending in lambda)
        const double2 e = av[sc + static_cast<size_t>(kept - 1 - k) * used];
        o = make_float2(static_cast<float>(e.x), static_cast<float>(e.y));
    }
    at[static_cast<size_t>(k) * used + sc] = o;
}

// a16 = at transposed into bf16 fragment planes (rows sc < sc_pad, zero
// beyond used; K = rank) and a2 = |at|^2, one thread per subcarrier and 8
// eigenvectors.
__global__ void __launch_bounds__(256) PackA16Kernel(const float2 *__restrict__ at, uint16_t *__restrict__ a16,
                                                     float *__restrict__ a2, uint32_t used, uint32_t sc_pad,
                                                     uint32_t rank)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= sc_pad * (rank / 8u)) return;
    const uint32_t sc = e % sc_pad, k0 = (e / sc_pad) * 8u;
    float2 v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        v[j] = sc < used ? at[static_cast<size_t>(k0 + j) * used + sc] : make_float2(0.0f, 0.0f);
        if (sc < used) a2[static_cast<size_t>(k0 + j) * used + sc] = v[j].x * v[j].x + v[j].y * v[j].y;
    }
    airan::StoreFrag8<2>(a16, sc, k0, rank / 16u, v);
}

// w16 = bt (= conj(U), [n][rank]) transposed into 3-part fragment planes:
// rows k < k_pad (zero beyond rank), K = AlignUp(n, 16) (zero beyond n); one
// thread per eigenvector and 8 observations.
__global__ void __launch_bounds__(256) PackW16Kernel(const float2 *__restrict__ bt, uint16_t *__restrict__ w16,
                                                     uint32_t n, uint32_t rank, uint32_t k_pad)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t pk = airan::AlignUp(n, 16u);
    if (e >= k_pad * (pk / 8u)) return;
    const uint32_t k = e % k_pad, p0 = (e / k_pad) * 8u;
    float2 v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
        v[j] = k < rank && p0 + j < n ? bt[static_cast<size_t>(p0 + j) * rank + k] : make_float2(0.0f, 0.0f);
    airan::StoreFrag8<3>(w16, k, p0, pk / 16u, v);
}

// ---------------------------------------------------------------- apply
struct CeModelDev {
    const float2 *bt;
    const float *a2;
    const uint16_t *a16;
    const uint16_t *w16;
    const float *lam;
    uint32_t n, rank, kept, in_a, in_b;
    float sign_b;
};

struct CeApply {
    uint32_t nr, nl, nm, ndmrs, pilot_pad, used, pad, nsym;
    CeModelDev m[kMaxModels];
    int32_t layer_model[4][2];
    float layer_sign[4];
    float rt[14][4];
    float rtt[4][4];
};

// Workspace layout for `slots` slots (each part 256-byte aligned):
//   stats  [slots][4]                       fp32  relative noise n_q
//   energy [slots][models][NR*D][2]         fp64  |z|^2, |u|^2 per column
//   inv    [slots][models][kMaxRank]        fp32  1 / (lam_k + n_q)
//   hv     [slots][models][used]            fp32  hhat variance
//   t      [slots][models][NR*D][kMaxRank]  c64   U^H z
//   hf     [slots][models][NR*D][used]      c64   frequency estimate (unscaled)
//   tp     [kProjSplits][slots][models][NR*D][kMaxRank]  c64  partial U^H z
//          over one of kProjSplits fixed ranges of the observations
//   zb     [models][zb_stride]               bf16  z in 3-part fragment planes:
//          rows n = slot * NR*D + col < npad (zero beyond slots * NR*D),
//          K = AlignUp(n_q, 16) observations (zero beyond n_q)
//   tb     [models][TbModelElems]            bf16  t / (lam_k + n_q) in fragment
//          planes: rows n = slot * NR*D + col < AlignUp(slots * NR*D,
//          kInterpTile) (zero beyond), K = the model's rank
struct CeWs {
    float *stats;
    double *energy;
    float *inv;
    float *hv;
    float2 *t;
    float2 *hf;
    float2 *tp;
    uint16_t *tb;
    uint16_t *zb;
    size_t zb_stride;  // bf16 per model in zb
    uint32_t npad;     // rows of tb and zb per model
};

size_t Align256(size_t b) { return (b + 255) / 256 * 256; }

uint32_t TbRows(const Geom &g, uint32_t slots) { return airan::AlignUp(slots * g.nr * g.ndmrs, kInterpTile); }

__host__ __device__ inline size_t TbModelElems(uint32_t npad) { return static_cast<size_t>(npad) * kMaxRank * 4; }

// Observations of the largest model: a comb model observes every DMRS RE.
size_t ZbModelElems(const Geom &g, uint32_t npad)
{
    return static_cast<size_t>(npad) * airan::AlignUp(g.n_dmrs_re, 16) * 6;
}

size_t CeWsLayout(const Geom &g, uint32_t slots, uint8_t *base, CeWs *w)
{
    const size_t sm = static_cast<size_t>(slots) * g.nl;  // models <= layers
    const size_t ncol = static_cast<size_t>(g.nr) * g.ndmrs;
    const uint32_t npad = TbRows(g, slots);
    const size_t sizes[9] = {Align256(16u * slots), Align256(sm * ncol * 16), Align256(sm * kMaxRank * 4),
                             Align256(sm * g.used * 4), Align256(sm * kMaxRank * ncol * 8),
                             Align256(sm * ncol * g.used * 8), Align256(kProjSplits * sm * kMaxRank * ncol * 8),
                             Align256(g.nl * TbModelElems(npad) * 2), Align256(g.nl * ZbModelElems(g, npad) * 2)};
    size_t off = 0;
    void *ptrs[9];
    for (int i = 0; i < 9; ++i) {
        ptrs[i] = base ? base + off : nullptr;
        off += sizes[i];
    }
    if (w) {
        w->stats = static_cast<float *>(ptrs[0]);
        w->energy = static_cast<double *>(ptrs[1]);
        w->inv = static_cast<float *>(ptrs[2]);
        w->hv = static_cast<float *>(ptrs[3]);
        w->t = static_cast<float2 *>(ptrs[4]);
        w->hf = static_cast<float2 *>(ptrs[5]);
        w->tp = static_cast<float2 *>(ptrs[6]);
        w->tb = static_cast<uint16_t *>(ptrs[7]);
        w->zb = static_cast<uint16_t *>(ptrs[8]);
        w->zb_stride = ZbModelElems(g, npad);
        w->npad = npad;
    }
    return off;
}

__device__ inline float2 LoadZ(const __half *hr, const __half *hi, const CeApply &c, const CeModelDev &md,
                               uint32_t slot, uint32_t col, uint32_t p)
{
    if (p >= md.n) return make_float2(0.0f, 0.0f);
    const uint32_t rx = col / c.ndmrs, d = col - rx * c.ndmrs;
    const size_t ha = (((static_cast<size_t>(slot) * c.nr + rx) * c.nl + md.in_a) * c.ndmrs + d) * c.pilot_pad + p;
    float2 v = make_float2(__half2float(hr[ha]), __half2float(hi[ha]));
    if (md.sign_b != 0.0f) {
        const size_t hb = (((static_cast<size_t>(slot) * c.nr + rx) * c.nl + md.in_b) * c.ndmrs + d) * c.pilot_pad + p;
        v.x += md.sign_b * __half2float(hr[hb]);
        v.y += md.sign_b * __half2float(hi[hb]);
    }
    return v;
}

// Observations of split `part` of a model: [p0, p0 + len), p0 and len
// multiples of 16 except for the last split's len.
__device__ inline void ProjRange(uint32_t n, uint32_t part, uint32_t *p0, uint32_t *len)
{
    const uint32_t chunk = airan::AlignUp((n + kProjSplits - 1) / kProjSplits, 16u);
    const uint32_t a = min(n, part * chunk), b = min(n, a + chunk);
    *p0 = a;
    *len = b - a;
}

// zb[q] = z[slot, col][p] in 3-part fragment planes (rows n = slot * NR*D +
// col, zero for n >= slots * NR*D and p >= n_q), one thread per row and 8
// observations.  Grid (npad * AlignUp(max n, 16) / 8 / 256, models).
__device__ __forceinline__ void CeZSplitKernelBody(const __half *__restrict__ hr, const __half *__restrict__ hi,
                                                      const CeWs &w, uint32_t slots, const CeApply &c, uint32_t bx, uint32_t by)
{
    const uint32_t q = by;
    const CeModelDev &md = c.m[q];
    const uint32_t e = bx * blockDim.x + threadIdx.x;
    const uint32_t n = e % w.npad, p0 = (e / w.npad) * 8u;
    const uint32_t pk = airan::AlignUp(md.n, 16u);
    if (p0 >= pk) return;
    const uint32_t ncol = c.nr * c.ndmrs;
    float2 v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) v[j] = make_float2(0.0f, 0.0f);
    if (n < slots * ncol) {
        // LoadZ of observations p0 .. p0 + 7 with 16-byte loads (p0 and
        // pilot_pad are multiples of 8 and p0 + 8 <= pk <= pilot_pad).
        const uint32_t slot = n / ncol, col = n - slot * ncol;
        const uint32_t rx = col / c.ndmrs, d = col - rx * c.ndmrs;
        auto row = [&](uint32_t layer) {
            return (((static_cast<size_t>(slot) * c.nr + rx) * c.nl + layer) * c.ndmrs + d) * c.pilot_pad + p0;
        };
        auto h8 = [](const __half *base, size_t off, float *out) {
            const uint4 u = *reinterpret_cast<const uint4 *>(base + off);
            const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 8; ++j)
                out[j] = __half2float(__ushort_as_half(static_cast<unsigned short>(w4[j >> 1] >> ((j & 1) * 16))));
        };
        float ar[8], ai[8];
        h8(hr, row(md.in_a), ar);
        h8(hi, row(md.in_a), ai);
        if (md.sign_b != 0.0f) {
            float br[8], bi[8];
            h8(hr, row(md.in_b), br);
            h8(hi, row(md.in_b), bi);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                ar[j] += md.sign_b * br[j];
                ai[j] += md.sign_b * bi[j];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j)
            if (p0 + j < md.n) v[j] = make_float2(ar[j], ai[j]);
    }
    airan::StoreFrag8<3>(w.zb + q * w.zb_stride, n, p0, pk / 16u, v);
}
__global__ void __launch_bounds__(256) CeZSplitKernel(const __half *__restrict__ hr, const __half *__restrict__ hi,
                                                      CeWs w, uint32_t slots, CeApply c)
{
    CeZSplitKernelBody(hr, hi, w, slots, c, blockIdx.x, blockIdx.y);
}

// tp[part][slot, q][col][k] = sum_{p in range(part)} z[slot, col][p] conj(U[p][k])
// on split bf16 MFMA with 3 parts (fp32-accurate products, which the
// subspace noise estimate needs): C = ZB W^T with rows n (slot, col) and
// columns k.  Workgroup: 64 rows x kProjKTile eigenvectors of one (model,
// split), four waves of 32 x 64 (1 x 2 blocks of 32 x 32).  Linear grid of
// 8 * per_xcd workgroups; each XCD takes a contiguous range of tiles ordered
// (model, split) major, so the operands of one (model, split) -- about 2 MB
// at 273 PRB and 64 slots -- are read through one L2.
__device__ __forceinline__ void CeProjectMfmaKernelBody(const CeWs &w, uint32_t slots, const CeApply &c, uint32_t k_tiles,
                                                              uint32_t per_xcd, uint32_t bx, uint32_t by)
{
    const uint32_t tile = (bx & 7u) * per_xcd + (bx >> 3);
    const uint32_t per_split = (w.npad / 64u) * k_tiles;
    if (tile >= c.nm * kProjSplits * per_split) return;
    const uint32_t qs = tile / per_split, rest = tile - qs * per_split;
    const uint32_t q = qs / kProjSplits, part = qs - q * kProjSplits;
    const CeModelDev &md = c.m[q];
    const uint32_t kt = rest % k_tiles, nt = rest / k_tiles;
    if (kt * kProjKTile >= md.rank) return;
    uint32_t p0, len;
    ProjRange(md.n, part, &p0, &len);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t nb = nt * 2u + (wave & 1u);
    const uint32_t kb0 = kt * (kProjKTile / 32u) + (wave >> 1) * 2u;
    const uint32_t pblocks = airan::AlignUp(md.n, 16u) / 16u;
    const uint16_t *za = w.zb + q * w.zb_stride;
    f32x16_t acc_r[2], acc_i[2];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc_r[b][i] = acc_i[b][i] = 0.0f;
    const uint32_t pb_begin = p0 / 16u, pb_end = (p0 + len + 15u) / 16u;
    auto load = [&](uint32_t pb, airan::CFrag<3> &fa, airan::CFrag<3> *fb) {
        fa = airan::LoadFrag<3>(za + (static_cast<size_t>(nb) * pblocks + pb) * airan::FragBlock<3>(), lane);
#pragma unroll
        for (int b = 0; b < 2; ++b)
            fb[b] = airan::LoadFrag<3>(md.w16 + (static_cast<size_t>(kb0 + b) * pblocks + pb) * airan::FragBlock<3>(),
                                       lane);
    };
    for (uint32_t pb = pb_begin; pb < pb_end; ++pb) {
        airan::CFrag<3> fa, fb[2];
        load(pb, fa, fb);
        const uint4 nai[3] = {airan::NegBf16x8(fa.p[3]), airan::NegBf16x8(fa.p[4]), airan::NegBf16x8(fa.p[5])};
#pragma unroll
        for (int b = 0; b < 2; ++b) airan::CMfmaSplit<3>(acc_r[b], acc_i[b], fa, nai, fb[b]);
    }
    const uint32_t ncol = c.nr * c.ndmrs, N = slots * ncol;
    float2 *tp = w.tp + static_cast<size_t>(part) * slots * c.nm * ncol * kMaxRank;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const uint32_t k = (kb0 + b) * 32u + (lane & 31u);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t n = nb * 32u + 8u * (i / 4) + 4u * (lane >> 5) + (i & 3);
            if (n < N && k < md.rank) {
                const uint32_t slot = n / ncol, col = n - slot * ncol;
                tp[((static_cast<size_t>(slot) * c.nm + q) * ncol + col) * kMaxRank + k] =
                    make_float2(acc_r[b][i], acc_i[b][i]);
            }
        }
    }
}
__global__ void __launch_bounds__(256, 2) CeProjectMfmaKernel(CeWs w, uint32_t slots, CeApply c, uint32_t k_tiles,
                                                              uint32_t per_xcd)
{
    CeProjectMfmaKernelBody(w, slots, c, k_tiles, per_xcd, blockIdx.x, blockIdx.y);
}

// Column energies: |z|^2 over the observations and |u|^2 over the kept
// eigenvectors, fp64 with a fixed reduction tree.  The projection's split-K
// partials are summed here first, t = ((tp[0] + tp[1]) + tp[2]) + tp[3] ...
// in split order, and t is stored for the kernels after.  Grid
// slots * models * NR*D.
__device__ __forceinline__ void CeEnergyKernelBody(const __half *__restrict__ hr, const __half *__restrict__ hi,
                                                      const CeWs &w, const CeApply &c, uint32_t slots, uint32_t bx, uint32_t by)
{
    __shared__ double red[2][4];
    const uint32_t ncol = c.nr * c.ndmrs;
    const uint32_t b = bx;
    const uint32_t col = b % ncol, sq = b / ncol;
    const uint32_t q = sq % c.nm, slot = sq / c.nm;
    const CeModelDev &md = c.m[q];
    double ez = 0.0, eu = 0.0;
    for (uint32_t p = threadIdx.x; p < md.n; p += blockDim.x) {
        const float2 v = LoadZ(hr, hi, c, md, slot, col, p);
        ez += static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
    }
    const size_t total = static_cast<size_t>(slots) * c.nm * kMaxRank * ncol;
    const size_t e0 = (static_cast<size_t>(sq) * ncol + col) * kMaxRank;
    for (uint32_t k = threadIdx.x; k < md.rank; k += blockDim.x) {
        float2 u = w.tp[e0 + k];
        for (uint32_t part = 1; part < kProjSplits; ++part) {
            const float2 v = w.tp[static_cast<size_t>(part) * total + e0 + k];
            u.x += v.x;
            u.y += v.y;
        }
        w.t[e0 + k] = u;
        eu += static_cast<double>(u.x) * u.x + static_cast<double>(u.y) * u.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ez += __shfl_xor(ez, off);
        eu += __shfl_xor(eu, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x / 64] = ez;
        red[1][threadIdx.x / 64] = eu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double *e = w.energy + 2 * (static_cast<size_t>(sq) * ncol + col);
        e[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        e[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}
__global__ void __launch_bounds__(256) CeEnergyKernel(const __half *__restrict__ hr, const __half *__restrict__ hi,
                                                      CeWs w, CeApply c, uint32_t slots)
{
    CeEnergyKernelBody(hr, hi, w, c, slots, blockIdx.x, blockIdx.y);
}

// Noise and per-model statistics of one slot (thread 0, fp64, fixed order),
// then inv[slot, q][k] = 1 / (lam_k + n_q) by all threads.
// Every observation z carries the per-RE noise sigma^2 (comb: y r*; FD-OCC
// pair: h_ls0 +- h_ls1, each the mean of two).  The noise lies in the
// complement of the kept eigenspace: per Rx, sigma^2 = sum_{q,d} (|z|^2 -
// |u|^2) over sum_{q,d} (n_q - kept_q) dimensions; the DMRS-LS estimate,
// which the channel's frequency selectivity inflates, is the fallback below
// kMinNoiseDims dimensions.  P_q = mean |z|^2 - sigma^2 and the filter's
// relative noise n_q = sigma^2 / P_q -> stats[4 slot + q];
// noise_out[slot][rx] = fp16(sigma^2_rx) when non-null.
__device__ __forceinline__ void CeStatsKernelBody(const __half *__restrict__ noise_in,
                                                     __half *__restrict__ noise_out, const CeWs &w, const CeApply &c, uint32_t bx, uint32_t by)
{
    __shared__ float nq[kMaxModels];
    // The slot's energies staged in LDS by all threads, so thread 0's serial
    // sums read LDS instead of waiting on one global load at a time.
    constexpr uint32_t kStage = 2 * kMaxModels * 32;
    __shared__ double stage[kStage];
    const uint32_t slot = bx;
    const uint32_t ncol = c.nr * c.ndmrs;
    const uint32_t n_e = 2 * c.nm * ncol;
    const double *eg = w.energy + 2 * static_cast<size_t>(slot) * c.nm * ncol;
    if (n_e <= kStage) {
        for (uint32_t t = threadIdx.x; t < n_e; t += blockDim.x) stage[t] = eg[t];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double *e = n_e <= kStage ? stage : eg;
        double dims = 0.0;
        for (uint32_t q = 0; q < c.nm; ++q) dims += static_cast<double>(c.m[q].n - c.m[q].kept) * c.ndmrs;
        double s2_sum = 0.0;
        for (uint32_t rx = 0; rx < c.nr; ++rx) {
            double s2;
            if (dims >= kMinNoiseDims) {
                double res = 0.0;
                for (uint32_t q = 0; q < c.nm; ++q)
                    for (uint32_t d = 0; d < c.ndmrs; ++d) {
                        const double *x = e + 2 * (static_cast<size_t>(q) * ncol + rx * c.ndmrs + d);
                        res += fmax(x[0] - x[1], 0.0);
                    }
                s2 = res / dims;
            } else {
                s2 = __half2float(noise_in[static_cast<size_t>(slot) * c.nr + rx]);
            }
            if (noise_out) noise_out[static_cast<size_t>(slot) * c.nr + rx] = __float2half_rn(static_cast<float>(s2));
            s2_sum += s2;
        }
        const double s2 = s2_sum / c.nr;
        for (uint32_t q = 0; q < c.nm; ++q) {
            double eh = 0.0;
            for (uint32_t col = 0; col < ncol; ++col) eh += e[2 * (static_cast<size_t>(q) * ncol + col)];
            const double mean = eh / (static_cast<double>(ncol) * c.m[q].n);
            const double p = fmax(mean - s2, 1e-3 * mean + 1e-30);
            const float v = static_cast<float>(s2 / p);
            w.stats[4 * slot + q] = v;
            nq[q] = v;
        }
    }
    __syncthreads();
    for (uint32_t q = 0; q < c.nm; ++q)
        for (uint32_t k = threadIdx.x; k < c.m[q].rank; k += blockDim.x)
            w.inv[(static_cast<size_t>(slot) * c.nm + q) * kMaxRank + k] = 1.0f / (c.m[q].lam[k] + nq[q]);
}
__global__ void __launch_bounds__(256) CeStatsKernel(const __half *__restrict__ noise_in,
                                                     __half *__restrict__ noise_out, CeWs w, CeApply c)
{
    CeStatsKernelBody(noise_in, noise_out, w, c, blockIdx.x, blockIdx.y);
}

// hhat_var[slot, q][sc] = sum_k |A[k][sc]|^2 / (lam_k + n_q) (ref.py:109-111)
// on fp32 MFMA (v_mfma_f32_32x32x2f32, exact fp32 products): a workgroup per
// 32 slots x 32 subcarriers of one model, its four waves taking fixed
// quarters of the rank (in steps of 16) and their partial sums added in
// wave order.  Lane l takes A[slot l % 32][k + l / 32] = 1 / (lam + n) and
// B[k + l / 32][sc l % 32] = |A|^2; accumulator i holds slot
// 8 (i / 4) + 4 (l / 32) + i % 4.  Rows past `slots` and subcarriers past
// `used` read clamped (valid) addresses and are not stored.  Grid: slot
// tiles fastest (they share the |A|^2 tile), then subcarrier tiles, then
// models.
__device__ __forceinline__ void CeHvMfmaKernelBody(const CeWs &w, uint32_t slots, const CeApply &c, uint32_t slot_tiles,
                                                      uint32_t sc_tiles, uint32_t bx, uint32_t by)
{
    __shared__ float part[3][16][64];
    const uint32_t st = bx % slot_tiles, rest = bx / slot_tiles;
    const uint32_t ct = rest % sc_tiles, q = rest / sc_tiles;
    const CeModelDev &md = c.m[q];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, r = lane & 31u, kh = lane >> 5;
    const uint32_t slot = min(st * 32u + r, slots - 1u), sc = min(ct * 32u + r, c.used - 1u);
    const float *ia = w.inv + (static_cast<size_t>(slot) * c.nm + q) * kMaxRank + kh;
    const float *ba = md.a2 + static_cast<size_t>(kh) * c.used + sc;
    const uint32_t steps = md.rank / 16u;
    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (uint32_t k0 = 16u * (wave * steps / 4u); k0 < 16u * ((wave + 1u) * steps / 4u); k0 += 16) {
        float av[8], bv[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            av[u] = ia[k0 + 2u * u];
            bv[u] = ba[static_cast<size_t>(k0 + 2u * u) * c.used];
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], bv[u], acc, 0, 0, 0);
    }
    if (wave > 0) {
#pragma unroll
        for (int i = 0; i < 16; ++i) part[wave - 1][i][lane] = acc[i];
    }
    __syncthreads();
    if (wave > 0) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = ((acc[i] + part[0][i][lane]) + part[1][i][lane]) + part[2][i][lane];
    const uint32_t col = ct * 32u + r;
    if (col >= c.used) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t s = st * 32u + 8u * (i / 4) + 4u * kh + (i & 3);
        if (s < slots) w.hv[(static_cast<size_t>(s) * c.nm + q) * c.used + col] = acc[i];
    }
}
__global__ void __launch_bounds__(256) CeHvMfmaKernel(CeWs w, uint32_t slots, CeApply c, uint32_t slot_tiles,
                                                      uint32_t sc_tiles)
{
    CeHvMfmaKernelBody(w, slots, c, slot_tiles, sc_tiles, blockIdx.x, blockIdx.y);
}

// tb[q] = t[slot, q][col][k] / (lam_k + n_q) in 2-part fragment planes, row
// n = slot * NR*D + col (zero for n >= slots * NR*D), one thread per row and
// 8 eigenvectors.  Grid (npad * kMaxRank / 8 / 256, models).
__device__ __forceinline__ void CeScaleSplitKernelBody(const CeWs &w, uint32_t slots, const CeApply &c, uint32_t bx, uint32_t by)
{
    const uint32_t q = by;
    const CeModelDev &md = c.m[q];
    const uint32_t e = bx * blockDim.x + threadIdx.x;
    const uint32_t n = e % w.npad, k0 = (e / w.npad) * 8u;
    if (k0 >= md.rank) return;
    const uint32_t ncol = c.nr * c.ndmrs;
    float2 v[8];
    if (n < slots * ncol) {
        const uint32_t slot = n / ncol, col = n - slot * ncol;
        const size_t sq = static_cast<size_t>(slot) * c.nm + q;
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            const float2 x = w.t[(sq * ncol + col) * kMaxRank + k0 + j];
            const float s = w.inv[sq * kMaxRank + k0 + j];
            v[j] = make_float2(x.x * s, x.y * s);
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) v[j] = make_float2(0.0f, 0.0f);
    }
    airan::StoreFrag8<2>(w.tb + q * TbModelElems(w.npad), n, k0, md.rank / 16u, v);
}
__global__ void __launch_bounds__(256) CeScaleSplitKernel(CeWs w, uint32_t slots, CeApply c)
{
    CeScaleSplitKernelBody(w, slots, c, blockIdx.x, blockIdx.y);
}

// hf[slot, q][col][sc] = sum_k (t[slot, q][col][k] / (lam_k + n_q)) A[k][sc]
// on split bf16 MFMA with 2 parts: C = TB A^T with rows n (slot, col) and
// columns sc.  Workgroup: a kInterpTile x kInterpTile tile of one model, four
// waves of 64 x 64 (2 x 2 blocks of 32 x 32).  Per K-step of 16 the tile's
// 4 row blocks and 4 subcarrier blocks (32 planes of 1 KB) are copied to
// LDS by the direct global -> LDS path, double-buffered: the next step's
// copies are in flight while this step's MFMAs run, and every fragment is
// fetched from HBM/L2 once per workgroup instead of once per wave.
// Workgroups are handed out so that each XCD works through a contiguous
// range of tiles with the row tiles of one subcarrier tile adjacent: the
// subcarrier-side operand is read once per XCD, the small row-side operand
// stays in every L2.  Linear grid of 8 * per_xcd.
__device__ __forceinline__ void CeInterpMfmaKernelBody(const CeWs &w, uint32_t slots, const CeApply &c, uint32_t sc_tiles,
                                                          uint32_t per_xcd, uint32_t bx, uint32_t by)
{
    __shared__ uint4 stage[2][8][4][64];  // [buffer][block: 4 row, 4 subcarrier][plane][lane]
    const uint32_t xcd = bx & 7u, j = bx >> 3;
    const uint32_t n_tiles = w.npad / kInterpTile;
    const uint32_t tile = xcd * per_xcd + j;
    if (tile >= c.nm * sc_tiles * n_tiles) return;
    const uint32_t nt = tile % n_tiles, rest = tile / n_tiles;
    const uint32_t st = rest % sc_tiles, q = rest / sc_tiles;
    const CeModelDev &md = c.m[q];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t nrow0 = (nt * kInterpTile) / 32u, nsc0 = (st * kInterpTile) / 32u;
    const uint32_t nb0 = nrow0 + (wave & 1u) * 2u, sb0 = nsc0 + (wave >> 1) * 2u;
    const uint32_t kb_count = md.rank / 16u;
    const uint16_t *ta = w.tb + q * TbModelElems(w.npad);
    f32x16_t acc_r[2][2], acc_i[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc_r[a][b][i] = acc_i[a][b][i] = 0.0f;
    // Wave w copies blocks 2w and 2w + 1 of K-step kb (blocks 0..3: row
    // blocks nrow0 + g, 4..7: subcarrier blocks nsc0 + g - 4): 8 planes.
    auto issue = [&](uint32_t kb, uint32_t buf) {
#pragma unroll
        for (uint32_t jb = 0; jb < 2; ++jb) {
            const uint32_t g = 2u * wave + jb;
            const uint16_t *src =
                g < 4u ? ta + (static_cast<size_t>(nrow0 + g) * kb_count + kb) * airan::FragBlock<2>()
                       : md.a16 + (static_cast<size_t>(nsc0 + g - 4u) * kb_count + kb) * airan::FragBlock<2>();
#pragma unroll
            for (uint32_t p = 0; p < 4; ++p) airan::PlaneToLds(src + p * airan::kFragPlane, &stage[buf][g][p][0], lane);
        }
    };
    issue(0, 0);
    for (uint32_t kb = 0; kb < kb_count; ++kb) {
        const uint32_t cur = kb & 1u;
        if (kb + 1u < kb_count) {
            issue(kb + 1u, cur ^ 1u);
            airan::WaitVm<8>();  // this wave's copies of step kb have landed
        } else {
            airan::WaitVm<0>();
        }
        __syncthreads();
        airan::CFrag<2> fa[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                fa[a].p[p] = stage[cur][(wave & 1u) * 2u + a][p][lane];
                fb[a].p[p] = stage[cur][4u + (wave >> 1) * 2u + a][p][lane];
            }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const uint4 nai[2] = {airan::NegBf16x8(fa[a].p[2]), airan::NegBf16x8(fa[a].p[3])};
#pragma unroll
            for (int b = 0; b < 2; ++b) airan::CMfmaSplit<2>(acc_r[a][b], acc_i[a][b], fa[a], nai, fb[b]);
        }
        __syncthreads();  // buffer cur is refilled by the next step's copies
    }
    const uint32_t ncol = c.nr * c.ndmrs, N = slots * ncol;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const uint32_t sc = (sb0 + b) * 32u + (lane & 31u);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint32_t n = (nb0 + a) * 32u + 8u * (i / 4) + 4u * (lane >> 5) + (i & 3);
                if (n < N && sc < c.used) {
                    const uint32_t slot = n / ncol, col = n - slot * ncol;
                    w.hf[((static_cast<size_t>(slot) * c.nm + q) * ncol + col) * c.used + sc] =
                        make_float2(acc_r[a][b][i], acc_i[a][b][i]);
                }
            }
        }
}
__global__ void __launch_bounds__(256, 2) CeInterpMfmaKernel(CeWs w, uint32_t slots, CeApply c, uint32_t sc_tiles,
                                                          uint32_t per_xcd)
{
    CeInterpMfmaKernelBody(w, slots, c, sc_tiles, per_xcd, blockIdx.x, blockIdx.y);
}

// (R_tt + e I) for D DMRS symbols, eliminated once (no pivoting; the matrix
// is symmetric positive definite) and applied to each symbol's right-hand
// side r = R_t[s][.]: the same operations, in the same order, as eliminating
// the augmented system per symbol.
template <int D>
struct TimeSolver {
    float m[D][D];
    float f[D][D];  // f[row][col]: elimination multipliers

    __device__ inline void Factor(const CeApply &c, float e)
    {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) m[i][j] = c.rtt[i][j] + (i == j ? e : 0.0f);
#pragma unroll
        for (int col = 0; col < D; ++col) {
            const float inv = 1.0f / m[col][col];
#pragma unroll
            for (int row = col + 1; row < D; ++row) {
                f[row][col] = m[row][col] * inv;
#pragma unroll
                for (int k = col; k < D; ++k) m[row][k] -= f[row][col] * m[col][k];
            }
        }
    }

    __device__ inline void Solve(const float *r, float *w) const
    {
        float b[D];
#pragma unroll
        for (int i = 0; i < D; ++i) b[i] = r[i];
#pragma unroll
        for (int col = 0; col < D; ++col)
#pragma unroll
            for (int row = col + 1; row < D; ++row) b[row] -= f[row][col] * b[col];
#pragma unroll
        for (int row = D - 1; row >= 0; --row) {
            float s = b[row];
#pragma unroll
            for (int k = row + 1; k < D; ++k) s -= m[row][k] * w[k];
            w[row] = s / m[row][row];
        }
    }
};

// Rescale of ref.py:109-116: scale = 2 / (hv + 1 - err), err = max(1 - hv, 0),
// and the rescaled error variance used by the time interpolation.
__device__ inline void Rescale(float hv, float *scale, float *err2)
{
    const float err = fmaxf(1.0f - hv, 0.0f);
    const float den = hv + 1.0f - err;
    const float s = den != 0.0f ? 2.0f / den : 0.0f;
    *scale = s;
    *err2 = fmaxf(s * (s - 1.0f) * hv + (1.0f - s) + s * err, 0.0f);
}

// Rescale (per model), pair combination h = (g_a + sign g_b) / 2 with
// relative error (err_a + err_b) / 2, and time interpolation over the D DMRS
// symbols.  One thread per subcarrier of one (slot, layer) = blockIdx.y.
// Writes H[slot][rx][l][sym][sc] fp16 for sc < pad (zero beyond used).
template <int D>
__device__ __forceinline__ void CePostKernelBody(const CeWs &w, __half *__restrict__ out_r, __half *__restrict__ out_i,
                                                    uint32_t out_layers, const CeApply &c, uint32_t bx, uint32_t by)
{
    const uint32_t sc = bx * blockDim.x + threadIdx.x;
    const uint32_t slot = by / c.nl, l = by - slot * c.nl;
    if (sc >= c.pad) return;
    const size_t plane = static_cast<size_t>(c.nsym) * c.pad;
    out_r += static_cast<size_t>(slot) * c.nr * out_layers * plane;
    out_i += static_cast<size_t>(slot) * c.nr * out_layers * plane;
    if (sc >= c.used) {
        for (uint32_t rx = 0; rx < c.nr; ++rx)
            for (uint32_t s = 0; s < c.nsym; ++s) {
                const size_t o = (static_cast<size_t>(rx) * out_layers + l) * plane + static_cast<size_t>(s) * c.pad + sc;
                out_r[o] = __ushort_as_half(0);
                out_i[o] = __ushort_as_half(0);
            }
        return;
    }
    const uint32_t ncol = c.nr * D;
    const int qa = c.layer_model[l][0], qb = c.layer_model[l][1];
    const size_t sqa = static_cast<size_t>(slot) * c.nm + qa;
    float sa, ea;
    Rescale(w.hv[sqa * c.used + sc], &sa, &ea);
    float wa = sa, wb = 0.0f, err2 = ea;
    size_t sqb = sqa;
    if (qb >= 0) {
        sqb = static_cast<size_t>(slot) * c.nm + qb;
        float sb, eb;
        Rescale(w.hv[sqb * c.used + sc], &sb, &eb);
        wa = 0.5f * sa;
        wb = 0.5f * c.layer_sign[l] * sb;
        err2 = 0.5f * (ea + eb);
    }
    TimeSolver<D> ts;
    ts.Factor(c, err2);
    float tw[14][D];  // time weights of every symbol (nsym <= 14)
#pragma unroll
    for (uint32_t s = 0; s < 14; ++s)
        if (s < c.nsym) ts.Solve(c.rt[s], tw[s]);
    for (uint32_t rx = 0; rx < c.nr; ++rx) {
        float2 hf[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const uint32_t col = rx * D + d;
            const float2 ga = w.hf[(sqa * ncol + col) * c.used + sc];
            hf[d] = make_float2(wa * ga.x, wa * ga.y);
            if (qb >= 0) {
                const float2 gb = w.hf[(sqb * ncol + col) * c.used + sc];
                hf[d].x += wb * gb.x;
                hf[d].y += wb * gb.y;
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < 14; ++s) {
            if (s >= c.nsym) break;
            float vr = 0.0f, vi = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                vr += tw[s][d] * hf[d].x;
                vi += tw[s][d] * hf[d].y;
            }
            const size_t o = (static_cast<size_t>(rx) * out_layers + l) * plane + static_cast<size_t>(s) * c.pad + sc;
            out_r[o] = __float2half_rn(vr);
            out_i[o] = __float2half_rn(vi);
        }
    }
}
template <int D>
__global__ void __launch_bounds__(256) CePostKernel(CeWs w, __half *__restrict__ out_r, __half *__restrict__ out_i,
                                                    uint32_t out_layers, CeApply c)
{
    CePostKernelBody<D>(w, out_r, out_i, out_layers, c, blockIdx.x, blockIdx.y);
}

bool CheckRocblas(rocblas_status s) { return s == rocblas_status_success; }

void FreeDev(void *p)
{
    if (p != nullptr) (void)hipFree(p);
}

// Eigen-factors of one observation model (FP64 on the device).
int32_t BuildModel(const Geom &g, const LmmseFilterParamsV1 &prm, hipStream_t stream, CeModelFactors *out)
{
    ObsGeom o{};
    o.n = out->n;
    o.step = out->step;
    o.offset = out->offset;
    o.used = g.used;
    o.scs_tau = prm.scs_hz * prm.tau_s;
    const uint32_t n = o.n;
    if (n == 0 || n > kMaxObs) return AIRAN_UNSUPPORTED_PROFILE;
    double2 *rpp = nullptr, *rhp = nullptr, *av = nullptr;
    float2 *at = nullptr;  // (R_hp U) transposed, [rank][used], packed into a16 / a2
    double *eig = nullptr, *work = nullptr;
    rocblas_int *info = nullptr;
    rocblas_handle handle = nullptr;
    std::vector<double> lam_host(n);
    uint32_t kept = 0;
    auto cleanup = [&]() {
        if (handle) rocblas_destroy_handle(handle);
        FreeDev(rpp);
        FreeDev(rhp);
        FreeDev(av);
        FreeDev(at);
        FreeDev(eig);
        FreeDev(work);
        FreeDev(info);
    };
    auto fail = [&]() {
        cleanup();
        return AIRAN_LAUNCH_FAILED;
    };
    if (hipMalloc(&rpp, sizeof(double2) * n * n) != hipSuccess ||
        hipMalloc(&rhp, sizeof(double2) * static_cast<size_t>(g.used) * n) != hipSuccess ||
        hipMalloc(&eig, sizeof(double) * n) != hipSuccess || hipMalloc(&work, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&info, sizeof(rocblas_int)) != hipSuccess)
        return fail();
    hipLaunchKernelGGL(BuildRppKernel, dim3(airan::CeilDiv(n, 256), n), dim3(256), 0, stream, rpp, o);
    hipLaunchKernelGGL(BuildRhpKernel, dim3(airan::CeilDiv(g.used, 256), n), dim3(256), 0, stream, rhp, o);
    if (!CheckRocblas(rocblas_create_handle(&handle)) || !CheckRocblas(rocblas_set_stream(handle, stream)))
        return fail();
    if (!CheckRocblas(rocsolver_zheevd(handle, rocblas_evect_original, rocblas_fill_lower,
                                       static_cast<rocblas_int>(n), reinterpret_cast<rocblas_double_complex *>(rpp),
                                       static_cast<rocblas_int>(n), eig, work, info)))
        return fail();
    rocblas_int info_h = 0;
    if (hipMemcpyAsync(lam_host.data(), eig, sizeof(double) * n, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(&info_h, info, sizeof(info_h), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess || info_h != 0)
        return fail();
    const double lmax = std::max(lam_host[n - 1], 1e-300);
    for (uint32_t k = 0; k < n && kept < static_cast<uint32_t>(std::min<int>(kMaxRank, prm.max_rank)); ++k) {
        if (lam_host[n - 1 - k] < prm.eig_floor * lmax) break;
        ++kept;
    }
    kept = std::max<uint32_t>(kept, 1);
    const uint32_t rank = airan::AlignUp(kept, 16);
    // av = R_hp (used x n) * V[:, n-kept .. n-1] (n x kept), column-major.
    if (hipMalloc(&av, sizeof(double2) * static_cast<size_t>(g.used) * kept) != hipSuccess) return fail();
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    if (!CheckRocblas(rocblas_zgemm(handle, rocblas_operation_none, rocblas_operation_none,
                                    static_cast<rocblas_int>(g.used), static_cast<rocblas_int>(kept),
                                    static_cast<rocblas_int>(n), &one,
                                    reinterpret_cast<const rocblas_double_complex *>(rhp),
                                    static_cast<rocblas_int>(g.used),
                                    reinterpret_cast<const rocblas_double_complex *>(rpp) + static_cast<size_t>(n - kept) * n,
                                    static_cast<rocblas_int>(n), &zero,
                                    reinterpret_cast<rocblas_double_complex *>(av),
                                    static_cast<rocblas_int>(g.used))))
        return fail();
    out->rank = rank;
    out->kept = kept;
    if (hipMalloc(&out->bt, sizeof(float2) * static_cast<size_t>(n) * rank) != hipSuccess ||
        hipMalloc(&at, sizeof(float2) * static_cast<size_t>(rank) * g.used) != hipSuccess ||
        hipMalloc(&out->a2, sizeof(float) * static_cast<size_t>(rank) * g.used) != hipSuccess ||
        hipMalloc(&out->lam, sizeof(float) * rank) != hipSuccess)
        return fail();
    hipLaunchKernelGGL(PackBtKernel, dim3(airan::CeilDiv(rank, 256), n), dim3(256), 0, stream, rpp, out->bt, n,
                       rank, kept);
    hipLaunchKernelGGL(PackAtKernel, dim3(airan::CeilDiv(g.used, 256), rank), dim3(256), 0, stream, av, at,
                       g.used, rank, kept);
    const uint32_t sc_pad = airan::AlignUp(g.used, kInterpTile);
    if (hipMalloc(&out->a16, sizeof(uint16_t) * 4 * static_cast<size_t>(sc_pad) * rank) != hipSuccess) return fail();
    hipLaunchKernelGGL(PackA16Kernel, dim3(airan::CeilDiv(sc_pad * (rank / 8), 256)), dim3(256), 0, stream, at,
                       out->a16, out->a2, g.used, sc_pad, rank);
    const uint32_t k_pad = airan::AlignUp(rank, kProjKTile), pk = airan::AlignUp(n, 16u);
    if (hipMalloc(&out->w16, sizeof(uint16_t) * 6 * static_cast<size_t>(k_pad) * pk) != hipSuccess) return fail();
    hipLaunchKernelGGL(PackW16Kernel, dim3(airan::CeilDiv(k_pad * (pk / 8), 256)), dim3(256), 0, stream, out->bt,
                       out->w16, n, rank, k_pad);
    std::vector<float> lam_f(rank, 1.0f);  // padded columns: zero factors, unit eigenvalue
    for (uint32_t k = 0; k < kept; ++k) lam_f[k] = static_cast<float>(std::max(lam_host[n - 1 - k], 0.0));
    if (hipMemcpyAsync(out->lam, lam_f.data(), sizeof(float) * rank, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail();
    cleanup();
    return AIRAN_OK;
}

// Observation models of a grant, in layer order: a layer alone on its comb
// gets one comb model (step 2); the first layer of an FD-OCC pair creates the
// pair's two models g+ (offset comb) and g- (offset comb + 2), step 4.  The
// pair's layer with FD-OCC signs (+,+) (port 1000 or 1002) is h_ls "a".
void PlanModels(const Geom &g, const PuschMimoConfig &cfg, LmmseFilter *f)
{
    f->nm = 0;
    bool done[4] = {};
    for (uint32_t l = 0; l < g.nl; ++l) {
        if (done[l]) continue;
        if (!g.occ[l]) {
            CeModelFactors &m = f->model[f->nm];
            m.n = g.n_dmrs_re;
            m.step = 2;
            m.offset = g.comb[l];
            m.in_a = m.in_b = l;
            m.sign_b = 0.0f;
            f->layer_model[l][0] = static_cast<int32_t>(f->nm);
            f->layer_model[l][1] = -1;
            f->layer_sign[l] = 1.0f;
            ++f->nm;
            done[l] = true;
            continue;
        }
        uint32_t partner = l;
        for (uint32_t q = l + 1; q < g.nl; ++q)
            if (g.comb[q] == g.comb[l]) partner = q;
        const bool l_plus = cfg.dmrs_ports[l] == 1000 || cfg.dmrs_ports[l] == 1002;
        const uint32_t lp = l_plus ? l : partner, lm = l_plus ? partner : l;
        for (int sgn = 0; sgn < 2; ++sgn) {
            CeModelFactors &m = f->model[f->nm + sgn];
            m.n = g.n_occ;
            m.step = 4;
            m.offset = g.comb[l] + 2u * sgn;
            m.in_a = lp;
            m.in_b = lm;
            m.sign_b = sgn == 0 ? 1.0f : -1.0f;
        }
        f->layer_model[lp][0] = f->layer_model[lm][0] = static_cast<int32_t>(f->nm);
        f->layer_model[lp][1] = f->layer_model[lm][1] = static_cast<int32_t>(f->nm + 1);
        f->layer_sign[lp] = 1.0f;
        f->layer_sign[lm] = -1.0f;
        f->nm += 2;
        done[lp] = done[lm] = true;
    }
}

CeApply MakeApply(const LmmseFilter &f)
{
    CeApply c{};
    const Geom &g = f.g;
    c.nr = g.nr;
    c.nl = g.nl;
    c.nm = f.nm;
    c.ndmrs = g.ndmrs;
    c.pilot_pad = g.pilot_pad;
    c.used = g.used;
    c.pad = g.pad;
    c.nsym = g.nsym;
    for (uint32_t q = 0; q < f.nm; ++q) {
        const CeModelFactors &m = f.model[q];
        c.m[q] = CeModelDev{m.bt, m.a2, m.a16, m.w16, m.lam, m.n, m.rank, m.kept, m.in_a, m.in_b, m.sign_b};
    }
    std::memcpy(c.layer_model, f.layer_model, sizeof(c.layer_model));
    std::memcpy(c.layer_sign, f.layer_sign, sizeof(c.layer_sign));
    std::memcpy(c.rt, f.rt, sizeof(c.rt));
    std::memcpy(c.rtt, f.rtt, sizeof(c.rtt));
    return c;
}

uint64_t PilotHash(const uint16_t *sc, uint32_t count)
{
    uint64_t h = 1469598103934665603ull;
    for (uint32_t i = 0; i < count; ++i) {
        h ^= static_cast<uint8_t>(sc[i] & 0xffu);
        h *= 1099511628211ull;
        h ^= static_cast<uint8_t>(sc[i] >> 8u);
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

namespace {

// ------------------------------------------------------ grouped launches
// Several CE applications (grant shapes of a ragged batch, each with its own
// filter, slots and workspace) in one launch per step: step k's grid is the
// concatenation of every job's grid for that step; a workgroup finds its job
// by binary search over the block prefix sums and runs the step's body with
// the job's arguments and its block coordinates within the job's grid, so
// each job computes exactly what its own LaunchCeApply would.
constexpr int kCeSteps = 8;  // ZSplit, Project, Energy, Stats, Hv, ScaleSplit, Interp, Post

struct CeJobDev {
    CeApply c;
    CeWs w;
    const __half *hr, *hi, *noise_in;
    __half *noise_out, *out_re, *out_im;
    uint32_t slots, out_layers;
    uint32_t k_tiles, proj_per_xcd, hv_slot_tiles, hv_sc_tiles, sc_tiles, interp_per_xcd;
    uint32_t gx[kCeSteps], gy[kCeSteps];
};

struct CeJobBlock {
    const CeJobDev *job;
    uint32_t bx, by;
};

// starts: [kCeSteps][njobs + 1] block prefix sums (starts[step][njobs] = total).
__device__ inline CeJobBlock LocateJob(const CeJobDev *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                       uint32_t njobs, int step)
{
    const uint32_t *st = starts + static_cast<size_t>(step) * (njobs + 1);
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (st[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    const CeJobDev *j = jobs + lo;
    const uint32_t li = b - st[lo];
    return CeJobBlock{j, li % j->gx[step], li / j->gx[step]};
}

__global__ void __launch_bounds__(256) CeZSplitGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 0);
    const CeJobDev &J = *jb.job;
    CeZSplitKernelBody(J.hr, J.hi, J.w, J.slots, J.c, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256, 2) CeProjectGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 1);
    const CeJobDev &J = *jb.job;
    CeProjectMfmaKernelBody(J.w, J.slots, J.c, J.k_tiles, J.proj_per_xcd, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256) CeEnergyGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 2);
    const CeJobDev &J = *jb.job;
    CeEnergyKernelBody(J.hr, J.hi, J.w, J.c, J.slots, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256) CeStatsGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 3);
    const CeJobDev &J = *jb.job;
    CeStatsKernelBody(J.noise_in, J.noise_out, J.w, J.c, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256) CeHvGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 4);
    const CeJobDev &J = *jb.job;
    CeHvMfmaKernelBody(J.w, J.slots, J.c, J.hv_slot_tiles, J.hv_sc_tiles, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256) CeScaleSplitGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 5);
    const CeJobDev &J = *jb.job;
    CeScaleSplitKernelBody(J.w, J.slots, J.c, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256, 2) CeInterpGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 6);
    const CeJobDev &J = *jb.job;
    CeInterpMfmaKernelBody(J.w, J.slots, J.c, J.sc_tiles, J.interp_per_xcd, jb.bx, jb.by);
}

__global__ void __launch_bounds__(256) CePostGrouped(const CeJobDev *jobs, const uint32_t *starts, uint32_t n)
{
    const CeJobBlock jb = LocateJob(jobs, starts, n, 7);
    const CeJobDev &J = *jb.job;
    switch (J.c.ndmrs) {
    case 1: CePostKernelBody<1>(J.w, J.out_re, J.out_im, J.out_layers, J.c, jb.bx, jb.by); break;
    case 2: CePostKernelBody<2>(J.w, J.out_re, J.out_im, J.out_layers, J.c, jb.bx, jb.by); break;
    case 3: CePostKernelBody<3>(J.w, J.out_re, J.out_im, J.out_layers, J.c, jb.bx, jb.by); break;
    default: CePostKernelBody<4>(J.w, J.out_re, J.out_im, J.out_layers, J.c, jb.bx, jb.by); break;
    }
}

// The grids LaunchCeApply uses for one application.
void StepGrids(const LmmseFilter &f, const CeWs &w, uint32_t slots, CeJobDev *j)
{
    const Geom &g = f.g;
    const uint32_t ncol = g.nr * g.ndmrs;
    uint32_t max_rank = 0, max_n = 0;
    for (uint32_t q = 0; q < f.nm; ++q) {
        max_rank = std::max(max_rank, f.model[q].rank);
        max_n = std::max(max_n, f.model[q].n);
    }
    j->k_tiles = airan::AlignUp(max_rank, kProjKTile) / kProjKTile;
    j->proj_per_xcd = airan::CeilDiv(f.nm * kProjSplits * (w.npad / 64) * j->k_tiles, 8);
    j->hv_slot_tiles = airan::CeilDiv(slots, 32);
    j->hv_sc_tiles = airan::CeilDiv(g.used, 32);
    j->sc_tiles = airan::AlignUp(g.used, kInterpTile) / kInterpTile;
    j->interp_per_xcd = airan::CeilDiv(f.nm * j->sc_tiles * (w.npad / kInterpTile), 8);
    const uint32_t gx[kCeSteps] = {airan::CeilDiv(static_cast<uint64_t>(w.npad) * airan::AlignUp(max_n, 16u) / 8, 256),
                                   8 * j->proj_per_xcd,
                                   slots * f.nm * ncol,
                                   slots,
                                   j->hv_slot_tiles * j->hv_sc_tiles * f.nm,
                                   airan::CeilDiv(static_cast<uint64_t>(w.npad) * kMaxRank / 8, 256),
                                   8 * j->interp_per_xcd,
                                   airan::CeilDiv(g.pad, 256)};
    const uint32_t gy[kCeSteps] = {f.nm, 1, 1, 1, 1, f.nm, 1, slots * g.nl};
    for (int k = 0; k < kCeSteps; ++k) {
        j->gx[k] = gx[k];
        j->gy[k] = gy[k];
    }
}

}  // namespace

namespace airan {

size_t CeWorkspaceBytes(const Geom &g, uint32_t slots) { return CeWsLayout(g, slots, nullptr, nullptr); }

bool CeFilterMatches(const LmmseFilter *f, const Geom &g)
{
    return f != nullptr && g.nl == f->g.nl && g.nr == f->g.nr && g.used == f->g.used && g.pad == f->g.pad &&
           g.ndmrs == f->g.ndmrs && std::memcmp(g.comb, f->g.comb, sizeof(g.comb)) == 0 &&
           std::memcmp(g.occ, f->g.occ, sizeof(g.occ)) == 0 &&
           std::memcmp(g.dmrs_sym, f->g.dmrs_sym, sizeof(g.dmrs_sym)) == 0;
}

int32_t LaunchCeApply(const LmmseFilter *f, const __half *hr, const __half *hi, const __half *noise_in,
                      __half *noise_out, void *ws, size_t ws_bytes, __half *out_re, __half *out_im,
                      uint32_t out_layers, uint32_t slots, hipStream_t s)
{
    const Geom &g = f->g;
    if (slots == 0 || slots > 65535 || out_layers < g.nl) return AIRAN_INVALID_ARGUMENT;
    if (ws_bytes < CeWorkspaceBytes(g, slots)) return AIRAN_RESOURCE_TOO_SMALL;
    const CeApply c = MakeApply(*f);
    CeWs w{};
    CeWsLayout(g, slots, static_cast<uint8_t *>(ws), &w);
    const uint32_t ncol = g.nr * g.ndmrs;
    uint32_t max_rank = 0;
    for (uint32_t q = 0; q < f->nm; ++q) max_rank = std::max(max_rank, f->model[q].rank);
    uint32_t max_n = 0;
    for (uint32_t q = 0; q < f->nm; ++q) max_n = std::max(max_n, f->model[q].n);
    hipLaunchKernelGGL(CeZSplitKernel, dim3(CeilDiv(static_cast<uint64_t>(w.npad) * AlignUp(max_n, 16u) / 8, 256),
                                            f->nm),
                       dim3(256), 0, s, hr, hi, w, slots, c);
    const uint32_t k_tiles = AlignUp(max_rank, kProjKTile) / kProjKTile;
    const uint32_t proj_per_xcd = CeilDiv(f->nm * kProjSplits * (w.npad / 64) * k_tiles, 8);
    hipLaunchKernelGGL(CeProjectMfmaKernel, dim3(8 * proj_per_xcd), dim3(256), 0, s, w, slots, c, k_tiles,
                       proj_per_xcd);
    hipLaunchKernelGGL(CeEnergyKernel, dim3(slots * f->nm * ncol), dim3(256), 0, s, hr, hi, w, c, slots);
    hipLaunchKernelGGL(CeStatsKernel, dim3(slots), dim3(256), 0, s, noise_in, noise_out, w, c);
    const uint32_t hv_slot_tiles = CeilDiv(slots, 32), hv_sc_tiles = CeilDiv(g.used, 32);
    hipLaunchKernelGGL(CeHvMfmaKernel, dim3(hv_slot_tiles * hv_sc_tiles * f->nm), dim3(256), 0, s, w, slots, c,
                       hv_slot_tiles, hv_sc_tiles);
    hipLaunchKernelGGL(CeScaleSplitKernel, dim3(CeilDiv(static_cast<uint64_t>(w.npad) * kMaxRank / 8, 256), f->nm),
                       dim3(256), 0, s, w, slots, c);
    const uint32_t sc_tiles = AlignUp(g.used, kInterpTile) / kInterpTile;
    const uint32_t per_xcd = CeilDiv(f->nm * sc_tiles * (w.npad / kInterpTile), 8);
    hipLaunchKernelGGL(CeInterpMfmaKernel, dim3(8 * per_xcd), dim3(256), 0, s, w, slots, c, sc_tiles, per_xcd);
    const dim3 pgrid(CeilDiv(g.pad, 256), slots * g.nl);
    switch (g.ndmrs) {
    case 1: hipLaunchKernelGGL(CePostKernel<1>, pgrid, dim3(256), 0, s, w, out_re, out_im, out_layers, c); break;
    case 2: hipLaunchKernelGGL(CePostKernel<2>, pgrid, dim3(256), 0, s, w, out_re, out_im, out_layers, c); break;
    case 3: hipLaunchKernelGGL(CePostKernel<3>, pgrid, dim3(256), 0, s, w, out_re, out_im, out_layers, c); break;
    default: hipLaunchKernelGGL(CePostKernel<4>, pgrid, dim3(256), 0, s, w, out_re, out_im, out_layers, c); break;
    }
    return LaunchStatus();
}

struct CeGrouped {
    CeJobDev *jobs = nullptr;
    uint32_t *starts = nullptr;
    uint32_t njobs = 0;
    uint32_t total[kCeSteps] = {};
};

void DestroyCeGrouped(CeGrouped *g)
{
    if (g == nullptr) return;
    FreeDev(g->jobs);
    FreeDev(g->starts);
    delete g;
}

int32_t CreateCeGrouped(const CeGroupJob *jobs, uint32_t n, CeGrouped **out)
{
    if (jobs == nullptr || n == 0 || out == nullptr) return AIRAN_INVALID_ARGUMENT;
    *out = nullptr;
    std::vector<CeJobDev> dev(n);
    std::vector<uint32_t> starts(static_cast<size_t>(kCeSteps) * (n + 1), 0);
    uint64_t acc[kCeSteps] = {};
    for (uint32_t i = 0; i < n; ++i) {
        const CeGroupJob &jb = jobs[i];
        if (jb.filter == nullptr) return AIRAN_INVALID_ARGUMENT;
        const Geom &g = jb.filter->g;
        if (jb.slots == 0 || jb.slots > 65535 || jb.out_layers < g.nl) return AIRAN_INVALID_ARGUMENT;
        if (jb.ws_bytes < CeWorkspaceBytes(g, jb.slots)) return AIRAN_RESOURCE_TOO_SMALL;
        CeJobDev &d = dev[i];
        d.c = MakeApply(*jb.filter);
        CeWsLayout(g, jb.slots, static_cast<uint8_t *>(jb.ws), &d.w);
        d.hr = jb.hr;
        d.hi = jb.hi;
        d.noise_in = jb.noise_in;
        d.noise_out = jb.noise_out;
        d.out_re = jb.out_re;
        d.out_im = jb.out_im;
        d.slots = jb.slots;
        d.out_layers = jb.out_layers;
        StepGrids(*jb.filter, d.w, jb.slots, &d);
        for (int k = 0; k < kCeSteps; ++k) {
            starts[static_cast<size_t>(k) * (n + 1) + i] = static_cast<uint32_t>(acc[k]);
            acc[k] += static_cast<uint64_t>(d.gx[k]) * d.gy[k];
        }
    }
    auto *g = new CeGrouped{};
    g->njobs = n;
    for (int k = 0; k < kCeSteps; ++k) {
        if (acc[k] > 0x7fffffffull) {
            delete g;
            return AIRAN_UNSUPPORTED_PROFILE;
        }
        starts[static_cast<size_t>(k) * (n + 1) + n] = static_cast<uint32_t>(acc[k]);
        g->total[k] = static_cast<uint32_t>(acc[k]);
    }
    if (hipMalloc(reinterpret_cast<void **>(&g->jobs), sizeof(CeJobDev) * n) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&g->starts), sizeof(uint32_t) * starts.size()) != hipSuccess ||
        hipMemcpy(g->jobs, dev.data(), sizeof(CeJobDev) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->starts, starts.data(), sizeof(uint32_t) * starts.size(), hipMemcpyHostToDevice) != hipSuccess) {
        DestroyCeGrouped(g);
        return AIRAN_RESOURCE_TOO_SMALL;
    }
    *out = g;
    return AIRAN_OK;
}

int32_t LaunchCeGrouped(const CeGrouped *g, hipStream_t s)
{
    if (g == nullptr) return AIRAN_INVALID_ARGUMENT;
    const CeJobDev *j = g->jobs;
    const uint32_t *st = g->starts, n = g->njobs;
    if (g->total[0]) hipLaunchKernelGGL(CeZSplitGrouped, dim3(g->total[0]), dim3(256), 0, s, j, st, n);
    if (g->total[1]) hipLaunchKernelGGL(CeProjectGrouped, dim3(g->total[1]), dim3(256), 0, s, j, st, n);
    if (g->total[2]) hipLaunchKernelGGL(CeEnergyGrouped, dim3(g->total[2]), dim3(256), 0, s, j, st, n);
    if (g->total[3]) hipLaunchKernelGGL(CeStatsGrouped, dim3(g->total[3]), dim3(256), 0, s, j, st, n);
    if (g->total[4]) hipLaunchKernelGGL(CeHvGrouped, dim3(g->total[4]), dim3(256), 0, s, j, st, n);
    if (g->total[5]) hipLaunchKernelGGL(CeScaleSplitGrouped, dim3(g->total[5]), dim3(256), 0, s, j, st, n);
    if (g->total[6]) hipLaunchKernelGGL(CeInterpGrouped, dim3(g->total[6]), dim3(256), 0, s, j, st, n);
    if (g->total[7]) hipLaunchKernelGGL(CePostGrouped, dim3(g->total[7]), dim3(256), 0, s, j, st, n);
    return LaunchStatus();
}

}  // namespace airan

extern "C" {

// channel_est_lmmse.cpp:75-87 (FNV-1a-64 over little-endian bytes)
uint64_t airan_channel_est_lmmse_PilotScHash(const uint16_t *pilot_sc, uint32_t count)
{
    if (pilot_sc == nullptr) return 0;
    return PilotHash(pilot_sc, count);
}

// channel_est_lmmse.cpp:89-118
int32_t airan_channel_est_lmmse_BuildWeightModel(const PuschMimoConfig *config, const PuschMimoLayout *layout,
                                                 const uint16_t *pilot_count, const uint16_t *pilot_sc,
                                                 LmmseWeightModelV1 *model)
{
    if (config == nullptr || layout == nullptr || pilot_count == nullptr || pilot_sc == nullptr || model == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (g.ndmrs > 2) return AIRAN_UNSUPPORTED_PROFILE;
    std::memset(model, 0, sizeof(*model));
    model->magic = kWeightModelMagic;
    model->abi_version = AIRAN_PUSCH_MIMO_ABI_VERSION;
    model->struct_size = sizeof(*model);
    model->num_layers = static_cast<uint16_t>(g.nl);
    model->num_dmrs_symbols = static_cast<uint16_t>(g.ndmrs);
    model->rank = 0;
    model->pilot_stride = static_cast<uint16_t>(g.pilot_pad);
    for (uint32_t l = 0; l < g.nl; ++l)
        for (uint32_t d = 0; d < g.ndmrs; ++d) {
            const uint32_t slot = l * g.ndmrs + d;
            model->pilot_count[l][d] = pilot_count[slot];
            model->pilot_sc_hash[l][d] = PilotHash(pilot_sc + static_cast<size_t>(slot) * g.pilot_pad, pilot_count[slot]);
        }
    return AIRAN_OK;
}

// channel_est_lmmse.cpp:120-160
int32_t airan_channel_est_lmmse_ValidateObservationModel(const PuschMimoConfig *config,
                                                         const PuschMimoLayout *layout,
                                                         const uint16_t *pilot_count, const uint16_t *pilot_sc,
                                                         const LmmseWeightModelV1 *model)
{
    if (config == nullptr || layout == nullptr || pilot_count == nullptr || pilot_sc == nullptr || model == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (g.ndmrs > 2) return AIRAN_UNSUPPORTED_PROFILE;  // the model holds two DMRS symbols
    PuschMimoLayout expected{};
    airan::FillLayout(g, &expected);
    if (!airan::SameLayout(expected, *layout)) return AIRAN_LAYOUT_MISMATCH;
    if (model->magic != kWeightModelMagic || model->abi_version != AIRAN_PUSCH_MIMO_ABI_VERSION ||
        model->struct_size < sizeof(LmmseWeightModelV1) || model->num_layers != g.nl ||
        model->num_dmrs_symbols != g.ndmrs || model->pilot_stride != g.pilot_pad)
        return -4;  // CE_WEIGHT_MODEL_MISMATCH (channel_est_lmmse.h:115-122)
    for (uint32_t v : model->reserved)
        if (v != 0) return -4;
    for (uint32_t l = 0; l < g.nl; ++l) {
        const uint32_t expect = g.occ[l] ? g.n_occ : g.n_dmrs_re;
        for (uint32_t d = 0; d < g.ndmrs; ++d) {
            const uint32_t slot = l * g.ndmrs + d;
            if (pilot_count[slot] != expect) return -3;  // CE_OBSERVATION_MODEL_MISMATCH
            const uint16_t *loc = pilot_sc + static_cast<size_t>(slot) * g.pilot_pad;
            for (uint32_t p = 0; p < expect; ++p)
                if (loc[p] != static_cast<uint16_t>(g.comb[l] + (g.occ[l] ? 1u + 4u * p : 2u * p))) return -3;
            if (model->pilot_count[l][d] != expect || model->pilot_sc_hash[l][d] != PilotHash(loc, expect))
                return -4;
        }
    }
    return AIRAN_OK;
}

int32_t airan_channel_est_lmmse_ValidateOpArgs(const ChannelEstLmmseOpArgsV1 *a)
{
    if (!airan::HeaderOk(a, sizeof(ChannelEstLmmseOpArgsV1)) || a->h_ls_re == nullptr || a->h_ls_im == nullptr ||
        a->pilot_count == nullptr || a->pilot_sc == nullptr || a->pilot_count_host == nullptr ||
        a->pilot_sc_host == nullptr || a->weight_model == nullptr || a->h_grid_re == nullptr ||
        a->h_grid_im == nullptr || a->config == nullptr || a->layout == nullptr || a->stream == nullptr)
        return AIRAN_INVALID_ARGUMENT;
    if (a->h_grid_re == a->h_grid_im || a->h_grid_re == a->h_ls_re || a->h_grid_re == a->h_ls_im ||
        a->h_grid_im == a->h_ls_re || a->h_grid_im == a->h_ls_im)
        return AIRAN_INVALID_ARGUMENT;
    return airan_channel_est_lmmse_ValidateObservationModel(a->config, a->layout, a->pilot_count_host,
                                                            a->pilot_sc_host, a->weight_model);
}

int32_t airan_channel_est_lmmse_CreateFilter(const PuschMimoConfig *config, const LmmseFilterParamsV1 *params,
                                             void *stream, LmmseFilter **filter)
{
    if (config == nullptr || params == nullptr || filter == nullptr || stream == nullptr ||
        params->abi_version != AIRAN_PUSCH_MIMO_ABI_VERSION || params->struct_size < sizeof(LmmseFilterParamsV1) ||
        !(params->tau_s >= 0.0) || !(params->fd_hz >= 0.0) || !(params->scs_hz > 0.0) || params->max_rank == 0 ||
        !(params->eig_floor >= 0.0))
        return AIRAN_INVALID_ARGUMENT;
    Geom g;
    int32_t st = airan::DeriveGeom(*config, &g);
    if (st != AIRAN_OK) return st;
    if (g.ndmrs > 4) return AIRAN_UNSUPPORTED_PROFILE;
    auto *f = new LmmseFilter{};
    if (hipGetDevice(&f->device) != hipSuccess) {
        delete f;
        return AIRAN_LAUNCH_FAILED;
    }
    f->g = g;
    f->params = *params;
    PlanModels(g, *config, f);
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (uint32_t q = 0; q < f->nm; ++q) {
        st = BuildModel(g, *params, s, &f->model[q]);
        if (st != AIRAN_OK) {
            airan_channel_est_lmmse_DestroyFilter(f);
            return st;
        }
    }
    // Time prior: J0(2 pi f_d dt), dt in OFDM symbols of duration (1 + 9/128) / scs
    // (channel_est_lmmse_ref.py:38,66-72).
    const double tsym = (1.0 + 9.0 / 128.0) / params->scs_hz;
    for (uint32_t s2 = 0; s2 < g.nsym; ++s2)
        for (uint32_t d = 0; d < g.ndmrs; ++d)
            f->rt[s2][d] = static_cast<float>(
                j0(2.0 * M_PI * params->fd_hz * (static_cast<double>(s2) - g.dmrs_sym[d]) * tsym));
    for (uint32_t i = 0; i < g.ndmrs; ++i)
        for (uint32_t j = 0; j < g.ndmrs; ++j)
            f->rtt[i][j] = static_cast<float>(
                j0(2.0 * M_PI * params->fd_hz * (static_cast<double>(g.dmrs_sym[i]) - g.dmrs_sym[j]) * tsym));
    *filter = f;
    return AIRAN_OK;
}

void airan_channel_est_lmmse_DestroyFilter(LmmseFilter *f)
{
    if (f == nullptr) return;
    for (uint32_t q = 0; q < kMaxModels; ++q) {
        FreeDev(f->model[q].bt);
        FreeDev(f->model[q].a2);
        FreeDev(f->model[q].a16);
        FreeDev(f->model[q].w16);
        FreeDev(f->model[q].lam);
    }
    delete f;
}

uint32_t airan_channel_est_lmmse_FilterRank(const LmmseFilter *f, uint32_t layer)
{
    if (f == nullptr || layer >= f->g.nl) return 0;
    return f->model[f->layer_model[layer][0]].kept;
}

size_t airan_channel_est_lmmse_WorkspaceBytes(const PuschMimoConfig *config)
{
    if (config == nullptr) return 0;
    Geom g;
    if (airan::DeriveGeom(*config, &g) != AIRAN_OK) return 0;
    return airan::CeWorkspaceBytes(g, 1);
}

int32_t airan_channel_est_lmmse_Enqueue(const ChannelEstLmmseOpArgsV1 *a, const LmmseFilter *f,
                                        const void *noise_var_rx, void *noise_out, void *workspace,
                                        size_t workspace_bytes)
{
    int32_t st = airan_channel_est_lmmse_ValidateOpArgs(a);
    if (st != AIRAN_OK) return st;
    if (f == nullptr || noise_var_rx == nullptr || workspace == nullptr) return AIRAN_INVALID_ARGUMENT;
    Geom g;
    airan::DeriveGeom(*a->config, &g);
    if (!airan::CeFilterMatches(f, g)) return AIRAN_LAYOUT_MISMATCH;
    if (noise_out != nullptr && (noise_out == noise_var_rx || noise_out == a->h_grid_re || noise_out == a->h_grid_im))
        return AIRAN_INVALID_ARGUMENT;
    return airan::LaunchCeApply(f, static_cast<const __half *>(a->h_ls_re), static_cast<const __half *>(a->h_ls_im),
                                static_cast<const __half *>(noise_var_rx), static_cast<__half *>(noise_out),
                                workspace, workspace_bytes, static_cast<__half *>(a->h_grid_re),
                                static_cast<__half *>(a->h_grid_im), airan::kDetectLayers, 1,
                                static_cast<hipStream_t>(a->stream));
}

}  // extern "C"
