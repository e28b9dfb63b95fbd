/*
 * cpu_chain.cpp -- the CPU baseline: the plain receive chain restated in
 * C++17 with FP32 arithmetic and OpenMP, timed beside the GPU chain by
 * bench.py's cpu_baseline leg.  TEST / MEASUREMENT INFRASTRUCTURE: nothing in
 * the product library links it.
 *
 * Stages per slot, in the order of chain/pusch_mimo_rx_chain/rx_chain.cpp:
 * 440-640:
 *   OFDM demod       FP32 radix-2 FFT per symbol and Rx (unitary), fftshift
 *                    RE demap to the fp16 grid
 *   DMRS gen + LS    the reference's fp16 arithmetic (or_dmrs_gen,
 *                    or_dmrs_ls: mimo_dmrs_ls_kernel.cpp:146-196) and its
 *                    DMRS-LS noise estimate
 *   LMMSE CE         the GPU chain's estimator (channel_est_lmmse_ref.py:
 *                    50-117 over the joint FD-OCC observation models): per
 *                    model the top eigenpairs (lam_k, v_k) of R_pp, set once
 *                    per grant by the caller (cpu_chain_set_eigen, FP64
 *                    LAPACK) with A = R_hp V formed here in FP64; per slot
 *                    in FP32: u = V^H z, the noise from the energy outside
 *                    the kept eigenspace, n_q = s2 / P_q,
 *                    G = s(sc) A diag(1 / (lam + n_q)) u with the
 *                    reference's rescale s(sc), layers (G+ +- G-) / 2 and
 *                    per-subcarrier time Wiener weights from J0(2 pi f_d dt)
 *   detection        FP32 unbiased per-RE LMMSE (mimo_detect_bri_ref.py:
 *                    72-86) by Cholesky, s2 the mean of the subspace noise
 *   demap .. rate dematch   the reference's integer stages (or_qam_demod,
 *                    or_layer_demap, or_descramble, or_rate_dematch)
 *   LDPC             layered normalised min-sum, alpha 0.75, the active rows
 *                    in order, early stop on a zero syndrome; FP32 posteriors
 *                    and messages, vectorised over the Z rows of a layer
 *   CRC              CRC24B per code block is carried in the TB check:
 *                    CRC24A / CRC16 over the TB (or_crc)
 * The slots of a batch run in parallel (OpenMP, one slot per thread).
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/airan_ldpc_bg1.h"
#include "oracle.h"

using cf = std::complex<float>;
using cd = std::complex<double>;

namespace {

constexpr double kMinNoiseDims = 32.0;  // residual dimensions for the subspace noise estimate

struct Model {
    int n, step, offset, a, b;
    float sb;
    int kept = 0;
    std::vector<float> lam;     // [kept] eigenvalues, descending
    std::vector<float> vr, vi;  // [kept][n] eigenvectors (row k = v_k)
    std::vector<float> ar, ai;  // [used][kept] A = R_hp V
    std::vector<float> a2;      // [used][kept] |A|^2
};

Model MakeModel(int n, int step, int offset, int a, int b, float sb)
{
    Model m;
    m.n = n;
    m.step = step;
    m.offset = offset;
    m.a = a;
    m.b = b;
    m.sb = sb;
    return m;
}

float NoiseFloor(int qm)
{
    const float norm = qm == 2 ? 2.0f : (qm == 4 ? 10.0f : (qm == 6 ? 42.0f : 170.0f));
    return 0.04f * static_cast<float>((1 << (qm / 2)) - 1) / norm;
}

}  // namespace

struct CpuChain {
    PuschMimoConfig cfg;
    OrGeom g;
    OrTbPlan plan;
    int iters, rows, nm;
    Model models[4];
    int lm[4][2];
    float lsgn[4];
    double prior_scs_tau, prior_fd_tsym;
    float rt[14][14];        // R_t = J0(2 pi f_d dt) between symbols
    std::vector<cf> twid;    // FFT twiddles
    std::vector<int> rev;    // bit reversal
    // LDPC edge tables: column base and shift per edge.
    std::vector<int> ecol, eshift;
};

namespace {

// Per-slot work buffers (one set per thread).
struct Work {
    std::vector<cf> x;
    std::vector<uint16_t> rx_re, rx_im, ref_re, ref_im, h_re, h_im, psc, pcnt;
    std::vector<double> noise;
    std::vector<float> hr, hi;      // H [NR][L][14][U] split
    std::vector<float> gr, gi;      // G [nm][NR][D][U], rescaled
    std::vector<float> zr, zi;      // z [NR*D][n]
    std::vector<float> ur, ui;      // inv * V^H z [nm][NR*D][kept]
    std::vector<float> err;         // [nm][U] rescaled error
    std::vector<uint16_t> x_re, x_im, n16;
    std::vector<int16_t> llr, cw, nr, lam;
    std::vector<float> post, msg, ext;
    std::vector<uint8_t> tb;
    explicit Work(const CpuChain &c)
    {
        const OrGeom &g = c.g;
        const size_t grid = (size_t)g.nr * g.nsym * g.pad, nls = (size_t)g.nr * g.nl * g.ndmrs * g.pilot_pad;
        x.resize(g.fft);
        rx_re.assign(grid, 0);
        rx_im.assign(grid, 0);
        ref_re.assign((size_t)g.nl * g.ndmrs * g.ref_pad, 0);
        ref_im.assign(ref_re.size(), 0);
        h_re.assign(nls, 0);
        h_im.assign(nls, 0);
        psc.assign((size_t)g.nl * g.ndmrs * g.pilot_pad, 0);
        pcnt.assign(16, 0);
        noise.assign(g.nr, 0.0);
        hr.assign((size_t)g.nr * g.nl * g.nsym * g.used, 0.0f);
        hi.assign(hr.size(), 0.0f);
        gr.assign((size_t)4 * g.nr * g.ndmrs * g.used, 0.0f);
        gi.assign(gr.size(), 0.0f);
        zr.assign((size_t)g.nr * g.ndmrs * g.n_dmrs_re, 0.0f);
        zi.assign(zr.size(), 0.0f);
        size_t umax = 0;
        for (int q = 0; q < c.nm; ++q) umax = std::max<size_t>(umax, c.models[q].kept);
        ur.assign((size_t)4 * g.nr * g.ndmrs * umax, 0.0f);
        ui.assign(ur.size(), 0.0f);
        err.assign((size_t)4 * g.used, 0.0f);
        x_re.assign((size_t)g.nl * g.nsym * g.pad, 0);
        x_im.assign(x_re.size(), 0);
        n16.assign(x_re.size(), 0);
        llr.assign((size_t)g.nl * g.qm * g.qam_plane, 0);
        cw.assign((size_t)g.qm * g.cw_stride, 0);
        nr.assign(cw.size(), 0);
        lam.assign((size_t)c.plan.C * 68 * c.plan.Z, 0);
        post.assign((size_t)68 * c.plan.Z, 0.0f);
        msg.assign((size_t)AIRAN_BG1_ROW_START[c.rows] * c.plan.Z, 0.0f);
        ext.assign((size_t)AIRAN_BG1_MAX_DEG * c.plan.Z, 0.0f);
        tb.assign((size_t)c.plan.C * c.plan.K, 0);
    }
};

void Fft(cf *x, int n, const std::vector<cf> &tw, const std::vector<int> &rev)
{
    for (int i = 0; i < n; ++i)
        if (i < rev[i]) std::swap(x[i], x[rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len / 2, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < half; ++j) {
                const cf w = tw[(size_t)j * step];
                const cf u = x[i + j], v = x[i + j + half] * w;
                x[i + j] = u + v;
                x[i + j + half] = u - v;
            }
    }
}

// OFDM demod + RE demap to the fp16 grid, DMRS generation and LS; returns
// the slot's DMRS-LS noise (mean over Rx).
double Front(const CpuChain &c, Work &w, const float *iq, const PuschMimoConfig &cfg)
{
    const OrGeom &g = c.g;
    const int N = g.fft, U = g.used;
    const float sc = 1.0f / std::sqrt(static_cast<float>(N));
    for (int r = 0; r < g.nr; ++r)
        for (int s = 0; s < g.nsym; ++s) {
            const float *src = iq + 2 * ((size_t)r * g.samples + g.sym_start[s]);
            for (int i = 0; i < N; ++i) w.x[i] = cf(src[2 * i], src[2 * i + 1]);
            Fft(w.x.data(), N, c.twid, c.rev);
            const size_t o = ((size_t)r * g.nsym + s) * g.pad;
            for (int u = 0; u < U; ++u) {
                const cf v = w.x[u < U / 2 ? N - U / 2 + u : u - U / 2] * sc;
                w.rx_re[o + u] = or_f2h(v.real());
                w.rx_im[o + u] = or_f2h(v.imag());
            }
        }
    or_dmrs_gen(&cfg, w.ref_re.data(), w.ref_im.data());
    or_dmrs_ls(&cfg, w.rx_re.data(), w.rx_im.data(), w.ref_re.data(), w.ref_im.data(), w.h_re.data(), w.h_im.data(),
               w.psc.data(), w.pcnt.data(), w.noise.data());
    double s2 = 0.0;
    for (int r = 0; r < g.nr; ++r) s2 += w.noise[r];
    return s2 / g.nr;
}

// LMMSE CE: subspace projection and noise, filter, rescale, layers and time
// interpolation; returns the detector's noise (mean over Rx).
float ChannelEstimate(const CpuChain &c, Work &w, double s2_ls)
{
    const OrGeom &g = c.g;
    const int U = g.used, D = g.ndmrs, NR = g.nr, L = g.nl, NS = g.nsym;
    const int cols = NR * D;
    double ez[4][64], eu[4][64];  // per model and column: |z|^2, |V^H z|^2
    for (int q = 0; q < c.nm; ++q) {
        const Model &m = c.models[q];
        const int n = m.n, K = m.kept;
        for (int rx = 0; rx < NR; ++rx)
            for (int d = 0; d < D; ++d) {
                const size_t ba = (((size_t)rx * L + m.a) * D + d) * g.pilot_pad;
                const size_t bb = (((size_t)rx * L + m.b) * D + d) * g.pilot_pad;
                float *zr = &w.zr[(size_t)(rx * D + d) * n], *zi = &w.zi[(size_t)(rx * D + d) * n];
                double e = 0.0;
                for (int p = 0; p < n; ++p) {
                    zr[p] = or_h2f(w.h_re[ba + p]) + m.sb * or_h2f(w.h_re[bb + p]);
                    zi[p] = or_h2f(w.h_im[ba + p]) + m.sb * or_h2f(w.h_im[bb + p]);
                    e += (double)zr[p] * zr[p] + (double)zi[p] * zi[p];
                }
                ez[q][rx * D + d] = e;
            }
        // u = V^H z for every column.
        float *ur = &w.ur[(size_t)q * cols * K], *ui = &w.ui[(size_t)q * cols * K];
        for (int k = 0; k < K; ++k) {
            const float *vr = &m.vr[(size_t)k * n], *vi = &m.vi[(size_t)k * n];
            for (int j = 0; j < cols; ++j) {
                const float *zr = &w.zr[(size_t)j * n], *zi = &w.zi[(size_t)j * n];
                float sr = 0.0f, si = 0.0f;
#pragma omp simd reduction(+ : sr, si)
                for (int p = 0; p < n; ++p) {
                    sr += vr[p] * zr[p] + vi[p] * zi[p];
                    si += vr[p] * zi[p] - vi[p] * zr[p];
                }
                ur[(size_t)j * K + k] = sr;
                ui[(size_t)j * K + k] = si;
            }
        }
        for (int j = 0; j < cols; ++j) {
            double e = 0.0;
            for (int k = 0; k < K; ++k)
                e += (double)ur[(size_t)j * K + k] * ur[(size_t)j * K + k] + (double)ui[(size_t)j * K + k] * ui[(size_t)j * K + k];
            eu[q][j] = e;
        }
    }
    // Noise outside the kept eigenspace (the DMRS-LS estimate below
    // kMinNoiseDims residual dimensions), then each model's relative noise.
    double dims = 0.0;
    for (int q = 0; q < c.nm; ++q) dims += (double)(c.models[q].n - c.models[q].kept) * D;
    double s2 = 0.0;
    if (dims >= kMinNoiseDims) {
        for (int rx = 0; rx < NR; ++rx) {
            double res = 0.0;
            for (int q = 0; q < c.nm; ++q)
                for (int d = 0; d < D; ++d) res += std::max(ez[q][rx * D + d] - eu[q][rx * D + d], 0.0);
            s2 += res / dims;
        }
        s2 /= NR;
    } else {
        s2 = s2_ls;
    }
    for (int q = 0; q < c.nm; ++q) {
        const Model &m = c.models[q];
        const int K = m.kept;
        double eh = 0.0;
        for (int j = 0; j < cols; ++j) eh += ez[q][j];
        const double mean = eh / ((double)cols * m.n);
        const double pw = std::max(mean - s2, 1e-3 * mean + 1e-30);
        const float nq = static_cast<float>(s2 / pw);
        float inv[2048];
        for (int k = 0; k < K; ++k) inv[k] = 1.0f / (m.lam[k] + nq);
        float *ur = &w.ur[(size_t)q * cols * K], *ui = &w.ui[(size_t)q * cols * K];
        for (int j = 0; j < cols; ++j)
            for (int k = 0; k < K; ++k) {
                ur[(size_t)j * K + k] *= inv[k];
                ui[(size_t)j * K + k] *= inv[k];
            }
        float *gr = &w.gr[(size_t)q * cols * U], *gi = &w.gi[(size_t)q * cols * U];
        float *er = &w.err[(size_t)q * U];
        for (int sc = 0; sc < U; ++sc) {
            const float *ar = &m.ar[(size_t)sc * K], *ai = &m.ai[(size_t)sc * K], *a2 = &m.a2[(size_t)sc * K];
            float hv = 0.0f;
#pragma omp simd reduction(+ : hv)
            for (int k = 0; k < K; ++k) hv += a2[k] * inv[k];
            // The reference's rescale (ref.py:109-116).
            const float e = std::max(1.0f - hv, 0.0f);
            const float den = hv + 1.0f - e;
            const float sc_ = den != 0.0f ? 2.0f / den : 0.0f;
            const float e2 = sc_ * (sc_ - 1.0f) * hv + (1.0f - sc_) + sc_ * e;
            er[sc] = std::max(e2, 0.0f);
            for (int j = 0; j < cols; ++j) {
                const float *uj = &ur[(size_t)j * K], *vj = &ui[(size_t)j * K];
                float sr = 0.0f, si = 0.0f;
#pragma omp simd reduction(+ : sr, si)
                for (int k = 0; k < K; ++k) {
                    sr += ar[k] * uj[k] - ai[k] * vj[k];
                    si += ar[k] * vj[k] + ai[k] * uj[k];
                }
                gr[(size_t)j * U + sc] = sc_ * sr;
                gi[(size_t)j * U + sc] = sc_ * si;
            }
        }
    }
    // Layers, then per-subcarrier time Wiener weights (R_t over the DMRS
    // symbols + err I)^-1 r_t(s).
    for (int l = 0; l < L; ++l) {
        const int qa = c.lm[l][0], qb = c.lm[l][1];
        for (int sc = 0; sc < U; ++sc) {
            const float e = qb < 0 ? w.err[(size_t)qa * U + sc] : 0.5f * (w.err[(size_t)qa * U + sc] + w.err[(size_t)qb * U + sc]);
            float M[4][4], Mi[4][4];
            for (int i = 0; i < D; ++i)
                for (int j = 0; j < D; ++j) {
                    M[i][j] = c.rt[g.dmrs_sym[i]][g.dmrs_sym[j]] + (i == j ? e : 0.0f);
                    Mi[i][j] = i == j ? 1.0f : 0.0f;
                }
            for (int col = 0; col < D; ++col) {  // Gauss-Jordan (symmetric positive definite)
                const float pv = 1.0f / M[col][col];
                for (int j = 0; j < D; ++j) {
                    M[col][j] *= pv;
                    Mi[col][j] *= pv;
                }
                for (int i = 0; i < D; ++i)
                    if (i != col) {
                        const float f = M[i][col];
                        for (int j = 0; j < D; ++j) {
                            M[i][j] -= f * M[col][j];
                            Mi[i][j] -= f * Mi[col][j];
                        }
                    }
            }
            for (int rx = 0; rx < NR; ++rx) {
                cf hf[4];
                for (int d = 0; d < D; ++d) {
                    const size_t ia = ((size_t)qa * cols + rx * D + d) * U + sc;
                    if (qb < 0) {
                        hf[d] = cf(w.gr[ia], w.gi[ia]);
                    } else {
                        const size_t ib = ((size_t)qb * cols + rx * D + d) * U + sc;
                        hf[d] = 0.5f * cf(w.gr[ia] + c.lsgn[l] * w.gr[ib], w.gi[ia] + c.lsgn[l] * w.gi[ib]);
                    }
                }
                for (int s = 0; s < NS; ++s) {
                    cf v = 0;
                    for (int i = 0; i < D; ++i) {
                        float wt = 0.0f;
                        for (int j = 0; j < D; ++j) wt += Mi[i][j] * c.rt[g.dmrs_sym[j]][s];
                        v += wt * hf[i];
                    }
                    const size_t o = (((size_t)rx * L + l) * NS + s) * U + sc;
                    w.hr[o] = v.real();
                    w.hi[o] = v.imag();
                }
            }
        }
    }
    return static_cast<float>(s2);
}

// FP32 unbiased LMMSE per RE into the demapper's fp16 inputs.
void Detect(const CpuChain &c, Work &w, float s2)
{
    const OrGeom &g = c.g;
    const int U = g.used, L = g.nl, NR = g.nr, NS = g.nsym;
    const uint16_t fl = or_f2h(NoiseFloor(g.qm));
    const float flf = or_h2f(fl);
    for (int ds = 0; ds < g.ndata; ++ds) {
        const int s = g.data_sym[ds];
        for (int k = 0; k < U; ++k) {
            cf A[4][4] = {}, mf[4] = {};
            for (int r = 0; r < NR; ++r) {
                cf h[4];
                for (int l = 0; l < L; ++l) {
                    const size_t i = (((size_t)r * L + l) * NS + s) * U + k;
                    h[l] = cf(w.hr[i], w.hi[i]);
                }
                const size_t iy = ((size_t)r * NS + s) * g.pad + k;
                const cf yy(or_h2f(w.rx_re[iy]), or_h2f(w.rx_im[iy]));
                for (int i = 0; i < L; ++i) {
                    mf[i] += std::conj(h[i]) * yy;
                    for (int j = 0; j <= i; ++j) A[i][j] += std::conj(h[i]) * h[j];
                }
            }
            for (int i = 0; i < L; ++i) A[i][i] += s2;
            cf C[4][4] = {}, Ci[4][4] = {};
            for (int j = 0; j < L; ++j) {
                float d = A[j][j].real();
                for (int q = 0; q < j; ++q) d -= std::norm(C[j][q]);
                d = std::sqrt(std::max(d, 1e-30f));
                C[j][j] = d;
                for (int i = j + 1; i < L; ++i) {
                    cf t = A[i][j];
                    for (int q = 0; q < j; ++q) t -= C[i][q] * std::conj(C[j][q]);
                    C[i][j] = t / d;
                }
            }
            for (int j = 0; j < L; ++j) {
                Ci[j][j] = 1.0f / C[j][j].real();
                for (int i = j + 1; i < L; ++i) {
                    cf t = 0;
                    for (int q = j; q < i; ++q) t -= C[i][q] * Ci[q][j];
                    Ci[i][j] = t / C[i][i].real();
                }
            }
            cf t[4];
            for (int i = 0; i < L; ++i) {
                t[i] = 0;
                for (int q = 0; q <= i; ++q) t[i] += Ci[i][q] * mf[q];
            }
            for (int l = 0; l < L; ++l) {
                cf xl = 0;
                float dinv = 0.0f;
                for (int i = l; i < L; ++i) {
                    xl += std::conj(Ci[i][l]) * t[i];
                    dinv += std::norm(Ci[i][l]);
                }
                const float dd = std::max(1.0f - s2 * dinv, 1e-3f);
                const size_t o = ((size_t)l * NS + s) * g.pad + k;
                w.x_re[o] = or_f2h(xl.real() / dd);
                w.x_im[o] = or_f2h(xl.imag() / dd);
                uint16_t nv = or_f2h((1.0f - dd) / dd);
                if (or_h2f(nv) < flf) nv = fl;
                w.n16[o] = nv;
            }
        }
    }
}

// Layered normalised min-sum over one code block; returns the iterations run.
int DecodeCb(const CpuChain &c, Work &w, const int16_t *lam_in, int8_t *bits)
{
    const int Z = static_cast<int>(c.plan.Z), rows = c.rows;
    float *post = w.post.data(), *ext = w.ext.data();
    for (int v = 0; v < 68 * Z; ++v) post[v] = lam_in[v] * (1.0f / 256.0f);
    std::fill(w.msg.begin(), w.msg.end(), 0.0f);
    int it = 0;
    while (it < c.iters) {
        for (int r = 0; r < rows; ++r) {
            const int e0 = AIRAN_BG1_ROW_START[r], deg = AIRAN_BG1_ROW_START[r + 1] - e0;
            for (int k = 0; k < deg; ++k) {  // ext = posterior - old message, rotated into row order
                const float *pc = post + c.ecol[e0 + k];
                const float *mo = &w.msg[(size_t)(e0 + k) * Z];
                float *ek = ext + (size_t)k * Z;
                const int sh = c.eshift[e0 + k], head = Z - sh;
                for (int i = 0; i < head; ++i) ek[i] = pc[i + sh] - mo[i];
                for (int i = head; i < Z; ++i) ek[i] = pc[i - head] - mo[i];
            }
            float m1[384], m2[384], sp[384];
            for (int i = 0; i < Z; ++i) {
                m1[i] = INFINITY;
                m2[i] = INFINITY;
                sp[i] = 0.75f;
            }
            for (int k = 0; k < deg; ++k) {
                const float *ek = ext + (size_t)k * Z;
#pragma omp simd
                for (int i = 0; i < Z; ++i) {
                    const float a = std::fabs(ek[i]);
                    m2[i] = a < m1[i] ? m1[i] : (a < m2[i] ? a : m2[i]);
                    m1[i] = a < m1[i] ? a : m1[i];
                    sp[i] = ek[i] >= 0.0f ? sp[i] : -sp[i];
                }
            }
            for (int k = 0; k < deg; ++k) {
                const float *ek = ext + (size_t)k * Z;
                float *mo = &w.msg[(size_t)(e0 + k) * Z];
                float *pc = post + c.ecol[e0 + k];
                const int sh = c.eshift[e0 + k], head = Z - sh;
#pragma omp simd
                for (int i = 0; i < Z; ++i) {
                    const float a = std::fabs(ek[i]);
                    const float mag = a == m1[i] ? m2[i] : m1[i];
                    mo[i] = (ek[i] >= 0.0f ? sp[i] : -sp[i]) * mag;
                }
                for (int i = 0; i < head; ++i) pc[i + sh] = ek[i] + mo[i];
                for (int i = head; i < Z; ++i) pc[i - head] = ek[i] + mo[i];
            }
        }
        ++it;
        bool ok = true;
        for (int r = 0; r < rows && ok; ++r) {
            const int e0 = AIRAN_BG1_ROW_START[r], e1 = AIRAN_BG1_ROW_START[r + 1];
            uint8_t par[384] = {};
            for (int e = e0; e < e1; ++e) {
                const float *pc = post + c.ecol[e];
                const int sh = c.eshift[e], head = Z - sh;
                for (int i = 0; i < head; ++i) par[i] ^= pc[i + sh] < 0.0f;
                for (int i = head; i < Z; ++i) par[i] ^= pc[i - head] < 0.0f;
            }
            for (int i = 0; i < Z; ++i) ok &= par[i] == 0;
        }
        if (ok) break;
    }
    for (uint32_t k = 0; k < c.plan.K; ++k) bits[k] = post[k] < 0.0f ? 1 : 0;
    return it;
}

int Slot(const CpuChain &c, Work &w, const float *iq, int slot_number, int8_t *bits, uint8_t *tb_ok)
{
    const OrGeom &g = c.g;
    const OrTbPlan &p = c.plan;
    PuschMimoConfig cfg = c.cfg;
    if (slot_number >= 0) cfg.slot_number = static_cast<uint16_t>(slot_number);
    const double s2_ls = Front(c, w, iq, cfg);
    const float s2 = ChannelEstimate(c, w, s2_ls);
    Detect(c, w, s2);
    or_qam_demod(&cfg, w.x_re.data(), w.x_im.data(), w.n16.data(), w.llr.data());
    or_layer_demap(&cfg, w.llr.data(), w.cw.data());
    or_descramble(&cfg, 1, 0, w.cw.data(), w.nr.data());
    or_rate_dematch(&p, g.nl, g.qm, g.cw_symbols, g.cw_stride, 1, w.nr.data(), w.lam.data());
    for (uint32_t cb = 0; cb < p.C; ++cb) DecodeCb(c, w, &w.lam[(size_t)cb * 68 * p.Z], bits + (size_t)cb * p.K);
    const uint32_t per_cb = p.Kprime - p.crc_cb;
    for (uint32_t r = 0; r < p.C; ++r)
        for (uint32_t k = 0; k < per_cb; ++k) w.tb[(size_t)r * per_cb + k] = static_cast<uint8_t>(bits[(size_t)r * p.K + k]);
    *tb_ok = p.crc_tb == 0 ? 1
                           : static_cast<uint8_t>(or_crc(p.crc_tb == 24 ? OR_CRC24A : OR_CRC16, w.tb.data(),
                                                         (uint64_t)p.tbs + p.crc_tb) == 0);
    return 0;
}

// The observation models of the grant (the oracle's ce_plan): returns the count.
int PlanModels(const PuschMimoConfig &cfg, const OrGeom &g, CpuChain *c)
{
    int comb[4], occ[4], done[4] = {0, 0, 0, 0}, nm = 0;
    for (int l = 0; l < g.nl; ++l) comb[l] = (cfg.dmrs_ports[l] - 1000) / 2;
    for (int l = 0; l < g.nl; ++l) {
        int share = 0;
        for (int q = 0; q < g.nl; ++q) share += comb[q] == comb[l];
        occ[l] = share == 2;
    }
    for (int l = 0; l < g.nl; ++l) {
        if (done[l]) continue;
        if (!occ[l]) {
            c->models[nm] = MakeModel(g.n_dmrs_re, 2, comb[l], l, l, 0.0f);
            c->lm[l][0] = nm;
            c->lm[l][1] = -1;
            c->lsgn[l] = 1.0f;
            ++nm;
            done[l] = 1;
            continue;
        }
        int partner = l;
        for (int q = l + 1; q < g.nl; ++q)
            if (comb[q] == comb[l]) partner = q;
        const bool plus = cfg.dmrs_ports[l] == 1000 || cfg.dmrs_ports[l] == 1002;
        const int lp = plus ? l : partner, lmn = plus ? partner : l;
        c->models[nm] = MakeModel(g.n_occ, 4, comb[l], lp, lmn, 1.0f);
        c->models[nm + 1] = MakeModel(g.n_occ, 4, comb[l] + 2, lp, lmn, -1.0f);
        c->lm[lp][0] = c->lm[lmn][0] = nm;
        c->lm[lp][1] = c->lm[lmn][1] = nm + 1;
        c->lsgn[lp] = 1.0f;
        c->lsgn[lmn] = -1.0f;
        nm += 2;
        done[lp] = done[lmn] = 1;
    }
    return nm;
}

}  // namespace

extern "C" {

/* Sets up a grant; its filters are complete once cpu_chain_set_eigen has
 * been called for every observation model.  Returns 0, or a negative code. */
int cpu_chain_create(const PuschMimoConfig *cfg, const OrTbPlan *plan, const OrCeParams *prm, int ldpc_iters,
                     CpuChain **out)
{
    auto *c = new CpuChain{};
    c->cfg = *cfg;
    c->plan = *plan;
    c->iters = ldpc_iters;
    if (or_geom(cfg, &c->g) != 0 || plan->Z > 384 || c->g.used > 4096 || c->g.nr * c->g.ndmrs > 64) {
        delete c;
        return -1;
    }
    const OrGeom &g = c->g;
    const int N = g.fft;
    c->rows = or_ldpc_active_rows(plan->Z, plan->Kprime, std::max(plan->E_high, plan->E_low), plan->k0, plan->Ncb);
    for (int e = 0; e < AIRAN_BG1_ROW_START[c->rows]; ++e) {
        c->ecol.push_back(static_cast<int>(AIRAN_BG1_COL[e] * plan->Z));
        c->eshift.push_back(static_cast<int>(AIRAN_BG1_SHIFT[e] % plan->Z));
    }
    c->twid.resize(N / 2);
    for (int k = 0; k < N / 2; ++k) c->twid[k] = cf(std::polar(1.0, -2.0 * M_PI * k / N));
    c->rev.resize(N);
    int logn = 0;
    while ((1 << logn) < N) ++logn;
    for (int i = 0; i < N; ++i) {
        int r = 0;
        for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
        c->rev[i] = r;
    }
    c->nm = PlanModels(*cfg, g, c);
    c->prior_scs_tau = prm->scs_hz * prm->tau_s;
    const double tsym = (1.0 + 9.0 / 128.0) / prm->scs_hz;
    for (int a = 0; a < 14; ++a)
        for (int b = 0; b < 14; ++b)  // J0 is even; libstdc++ takes x >= 0 only
            c->rt[a][b] = static_cast<float>(std::cyl_bessel_j(0.0, std::fabs(2.0 * M_PI * prm->fd_hz * (a - b) * tsym)));
    *out = c;
    return 0;
}

int cpu_chain_num_models(const CpuChain *c) { return c->nm; }

/* Model q's observation count, and its covariance R_pp (n x n complex,
 * interleaved, row-major; R_f(dk) = 1 / (1 + j 2 pi dk scs tau)) when rpp is
 * non-null.  Returns n, or -1. */
int cpu_chain_model_rpp(const CpuChain *c, int q, double *rpp)
{
    if (q < 0 || q >= c->nm) return -1;
    const Model &m = c->models[q];
    if (rpp)
        for (int i = 0; i < m.n; ++i)
            for (int j = 0; j < m.n; ++j) {
                const cd v = cd(1.0, 0.0) / cd(1.0, 2.0 * M_PI * (double)(i - j) * m.step * c->prior_scs_tau);
                rpp[2 * ((size_t)i * m.n + j)] = v.real();
                rpp[2 * ((size_t)i * m.n + j) + 1] = v.imag();
            }
    return m.n;
}

/* Model q's kept eigenpairs, descending: lam [kept], vec [kept][n] complex
 * interleaved (row k = eigenvector k).  Forms A = R_hp V in FP64 and stores
 * the factors in FP32.  Returns 0, or -1. */
int cpu_chain_set_eigen(CpuChain *c, int q, int kept, const double *lam, const double *vec)
{
    if (q < 0 || q >= c->nm || kept < 1 || kept > 2048 || kept > c->models[q].n) return -1;
    Model &m = c->models[q];
    const int n = m.n, U = c->g.used;
    m.kept = kept;
    m.lam.resize(kept);
    m.vr.resize((size_t)kept * n);
    m.vi.resize((size_t)kept * n);
    for (int k = 0; k < kept; ++k) {
        m.lam[k] = static_cast<float>(std::max(lam[k], 0.0));
        for (int p = 0; p < n; ++p) {
            m.vr[(size_t)k * n + p] = static_cast<float>(vec[2 * ((size_t)k * n + p)]);
            m.vi[(size_t)k * n + p] = static_cast<float>(vec[2 * ((size_t)k * n + p) + 1]);
        }
    }
    m.ar.assign((size_t)U * kept, 0.0f);
    m.ai.assign(m.ar.size(), 0.0f);
    m.a2.assign(m.ar.size(), 0.0f);
    const double tau = c->prior_scs_tau;
#pragma omp parallel for schedule(static)
    for (int sc = 0; sc < U; ++sc) {
        std::vector<cd> r(n);
        for (int p = 0; p < n; ++p)
            r[p] = cd(1.0, 0.0) / cd(1.0, 2.0 * M_PI * (double)(sc - (m.offset + m.step * p)) * tau);
        for (int k = 0; k < kept; ++k) {
            cd s = 0;
            const double *v = vec + 2 * (size_t)k * n;
            for (int p = 0; p < n; ++p) s += r[p] * cd(v[2 * p], v[2 * p + 1]);
            m.ar[(size_t)sc * kept + k] = static_cast<float>(s.real());
            m.ai[(size_t)sc * kept + k] = static_cast<float>(s.imag());
            m.a2[(size_t)sc * kept + k] = static_cast<float>(std::norm(s));
        }
    }
    return 0;
}

void cpu_chain_destroy(CpuChain *c) { delete c; }

/* Decodes num_slots slots over `threads` OpenMP threads, one slot per
 * thread at a time: slot s reads iq + s * NR * samples * 2 floats and writes
 * bits + s * C * K and tb_ok[s].  slot_numbers NULL: the config's slot
 * number for every slot. */
int cpu_chain_run(const CpuChain *c, const float *iq, int num_slots, const int *slot_numbers, int threads,
                  int8_t *bits, uint8_t *tb_ok)
{
    for (int q = 0; q < c->nm; ++q)
        if (c->models[q].kept == 0) return -1;
    const size_t iq_stride = (size_t)c->g.nr * c->g.samples * 2, bit_stride = (size_t)c->plan.C * c->plan.K;
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) reduction(| : rc)
#endif
    {
        Work w(*c);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int s = 0; s < num_slots; ++s)
            rc |= Slot(*c, w, iq + s * iq_stride, slot_numbers ? slot_numbers[s] : -1, bits + s * bit_stride,
                       tb_ok + s);
    }
    (void)threads;
    return rc;
}

int cpu_chain_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
