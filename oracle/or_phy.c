/*
 * or_phy.c -- oracle: OFDM demodulation, LMMSE channel estimation, LMMSE
 * detection, in double precision.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cd;

/* ------------------------------------------------------------------ OFDM */
/* In-place radix-2 FFT, forward sign, no scaling. */
static void fft_pow2(cd *x, int n)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cd t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                double ang = -2.0 * M_PI * k / len;
                cd w = cos(ang) + I * sin(ang);
                cd u = x[i + k], v = x[i + k + len / 2] * w;
                x[i + k] = u + v;
                x[i + k + len / 2] = u - v;
            }
    }
}

/* CP removal + unitary DFT per symbol: the mathematics of
 * ofdm/ofdm_demod_siso/scripts/ofdm_demod_ref.py:100-136 (the P x Q mixed
 * radix there is one factorisation of this DFT; bin k is natural order). */
void or_ofdm_demod(const PuschMimoConfig *cfg, const float *iq, double *grid)
{
    OrGeom g;
    or_geom(cfg, &g);
    const double scale = 1.0 / sqrt((double)g.fft);
#pragma omp parallel
    {
        cd *buf = (cd *)malloc(sizeof(cd) * (size_t)g.fft);
#pragma omp for collapse(2)
        for (int rx = 0; rx < g.nr; ++rx)
            for (int s = 0; s < g.nsym; ++s) {
                const float *x = iq + 2 * ((size_t)rx * g.samples + g.sym_start[s]);
                for (int n = 0; n < g.fft; ++n) buf[n] = x[2 * n] + I * x[2 * n + 1];
                fft_pow2(buf, g.fft);
                double *o = grid + 2 * (((size_t)rx * g.nsym + s) * g.fft);
                for (int k = 0; k < g.fft; ++k) {
                    o[2 * k] = creal(buf[k]) * scale;
                    o[2 * k + 1] = cimag(buf[k]) * scale;
                }
            }
        free(buf);
    }
}

/* --------------------------------------------------------------- CE LMMSE */
static inline cd rf(double dsc, const OrCeParams *p)
{
    return 1.0 / (1.0 + I * 2.0 * M_PI * dsc * p->scs_hz * p->tau_s);
}

/* Hermitian positive-definite Cholesky, lower factor in place (row-major). */
static int cholesky(cd *a, int n)
{
    for (int j = 0; j < n; ++j) {
        double d = creal(a[(size_t)j * n + j]);
        for (int k = 0; k < j; ++k) d -= creal(a[(size_t)j * n + k] * conj(a[(size_t)j * n + k]));
        if (d <= 0.0) return -1;
        d = sqrt(d);
        a[(size_t)j * n + j] = d;
#pragma omp parallel for schedule(static)
        for (int i = j + 1; i < n; ++i) {
            cd s = a[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= a[(size_t)i * n + k] * conj(a[(size_t)j * n + k]);
            a[(size_t)i * n + j] = s / d;
        }
    }
    return 0;
}

static void forward_sub(const cd *L, int n, cd *x)
{
    for (int i = 0; i < n; ++i) {
        cd s = x[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
}

static void backward_sub_h(const cd *L, int n, cd *x)
{
    for (int i = n - 1; i >= 0; --i) {
        cd s = x[i];
        for (int k = i + 1; k < n; ++k) s -= conj(L[(size_t)k * n + i]) * x[k];
        x[i] = s / conj(L[(size_t)i * n + i]);
    }
}

/* Small dense real solve (D <= 4), Gaussian elimination with pivoting. */
static void solve_real(double *m, double *b, int n)
{
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(m[r * n + c]) > fabs(m[piv * n + c])) piv = r;
        for (int k = 0; k < n; ++k) {
            double t = m[c * n + k];
            m[c * n + k] = m[piv * n + k];
            m[piv * n + k] = t;
        }
        double t = b[c];
        b[c] = b[piv];
        b[piv] = t;
        for (int r = c + 1; r < n; ++r) {
            double f = m[r * n + c] / m[c * n + c];
            for (int k = c; k < n; ++k) m[r * n + k] -= f * m[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int r = n - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < n; ++k) s -= m[r * n + k] * b[k];
        b[r] = s / m[r * n + r];
    }
}

/* Observation models of a grant, in layer order: a layer alone on its comb
 * is observed at comb + 2p (step 2); the first layer of an FD-OCC pair
 * creates the pair's models g+ = h0 + h1 at comb + 4p and g- = h0 - h1 at
 * comb + 2 + 4p (step 4), observed exactly as h_ls0 +- h_ls1 (h_ls0 the
 * layer with FD-OCC signs +,+: port 1000 or 1002).  Returns the count. */
typedef struct {
    int n, step, offset, a, b;
    double sb;
} CeModel;

static int ce_plan(const PuschMimoConfig *cfg, const OrGeom *g, CeModel *m, int lm[4][2], double ls[4])
{
    int comb[4], occ[4], done[4] = {0, 0, 0, 0}, nm = 0;
    for (int l = 0; l < g->nl; ++l) comb[l] = (cfg->dmrs_ports[l] - 1000) / 2;
    for (int l = 0; l < g->nl; ++l) {
        int share = 0;
        for (int p = 0; p < g->nl; ++p) share += comb[p] == comb[l];
        occ[l] = share == 2;
    }
    for (int l = 0; l < g->nl; ++l) {
        if (done[l]) continue;
        if (!occ[l]) {
            m[nm] = (CeModel){g->n_dmrs_re, 2, comb[l], l, l, 0.0};
            lm[l][0] = nm;
            lm[l][1] = -1;
            ls[l] = 1.0;
            ++nm;
            done[l] = 1;
            continue;
        }
        int partner = l;
        for (int q = l + 1; q < g->nl; ++q)
            if (comb[q] == comb[l]) partner = q;
        const int plus = cfg->dmrs_ports[l] == 1000 || cfg->dmrs_ports[l] == 1002;
        const int lp = plus ? l : partner, lmn = plus ? partner : l;
        m[nm] = (CeModel){g->n_occ, 4, comb[l], lp, lmn, 1.0};
        m[nm + 1] = (CeModel){g->n_occ, 4, comb[l] + 2, lp, lmn, -1.0};
        lm[lp][0] = lm[lmn][0] = nm;
        lm[lp][1] = lm[lmn][1] = nm + 1;
        ls[lp] = 1.0;
        ls[lmn] = -1.0;
        nm += 2;
        done[lp] = done[lmn] = 1;
    }
    return nm;
}

/* Relative noise of each observation model: sigma2 / (mean |z_q|^2 - sigma2),
 * the power floored at 1e-3 of the mean; returns the model count. */
int or_ce_noise_rel(const PuschMimoConfig *cfg, const double *h_ls, double sigma2, double *noise_rel)
{
    OrGeom g;
    or_geom(cfg, &g);
    CeModel mod[4];
    int lm[4][2];
    double ls[4];
    const int nm = ce_plan(cfg, &g, mod, lm, ls);
    for (int q = 0; q < nm; ++q) {
        double e = 0.0;
        for (int rx = 0; rx < g.nr; ++rx)
            for (int d = 0; d < g.ndmrs; ++d) {
                const double *sa = h_ls + 2 * ((((size_t)rx * g.nl + mod[q].a) * g.ndmrs + d) * g.pilot_pad);
                const double *sb = h_ls + 2 * ((((size_t)rx * g.nl + mod[q].b) * g.ndmrs + d) * g.pilot_pad);
                for (int p = 0; p < mod[q].n; ++p) {
                    const double re = sa[2 * p] + mod[q].sb * sb[2 * p], im = sa[2 * p + 1] + mod[q].sb * sb[2 * p + 1];
                    e += re * re + im * im;
                }
            }
        const double mean = e / ((double)g.nr * g.ndmrs * mod[q].n);
        double pw = mean - sigma2;
        if (pw < 1e-3 * mean + 1e-30) pw = 1e-3 * mean + 1e-30;
        noise_rel[q] = sigma2 / pw;
    }
    return nm;
}

/* Full-rank FP64 LMMSE channel estimation (channel_est/channel_est_lmmse_mimo/
 * scripts/channel_est_lmmse_ref.py:50-117) over the observation models above:
 *   frequency:  G_q = R_hp (R_pp + n_q I)^-1 z_q, R_f(dk) = 1/(1 + j2pi dk scs tau)
 *   rescale:    scale = 2 / (hhat_var + 1 - err), err = max(1 - hhat_var, 0)
 *               (ref.py:109-116), per model
 *   layer:      H_f = G_a (comb model) or (G_a +- G_b) / 2 (FD-OCC pair),
 *               relative error (err_a + err_b) / 2
 *   time:       per-subcarrier Wiener weights over the DMRS symbols with
 *               R_t = J0(2 pi f_d dt) and the rescaled error (ref.py:75-88).
 * noise_rel[q] is model q's relative noise sigma^2 / P_q.  The reference
 * truncates R_pp to its top RANK eigenpairs; with RANK equal to the
 * observation count the two are identical. */
int or_ce_lmmse(const PuschMimoConfig *cfg, const OrCeParams *prm, const double *noise_rel,
                const double *h_ls, double *h_out)
{
    OrGeom g;
    or_geom(cfg, &g);
    const int U = g.used, D = g.ndmrs, NR = g.nr;
    const double tsym = (1.0 + 9.0 / 128.0) / prm->scs_hz;
    double rt[14][14];
    for (int i = 0; i < g.nsym; ++i)
        for (int j = 0; j < g.nsym; ++j) rt[i][j] = j0(2.0 * M_PI * prm->fd_hz * (i - j) * tsym);
    CeModel mod[4];
    int lm[4][2];
    double lsgn[4];
    const int nm = ce_plan(cfg, &g, mod, lm, lsgn);
    cd *G[4] = {NULL, NULL, NULL, NULL};     /* [NR][D][U] scaled frequency estimates */
    double *err[4] = {NULL, NULL, NULL, NULL};
    int rc = 0;
    for (int q = 0; q < nm && rc == 0; ++q) {
        const CeModel *m = &mod[q];
        const int n = m->n;
        cd *M = (cd *)malloc(sizeof(cd) * (size_t)n * n);
        cd *Rhp = (cd *)malloc(sizeof(cd) * (size_t)U * n);
        double *scale = (double *)malloc(sizeof(double) * U);
        G[q] = (cd *)malloc(sizeof(cd) * (size_t)NR * D * U);
        err[q] = (double *)malloc(sizeof(double) * U);
        for (int p = 0; p < n; ++p)
            for (int r = 0; r < n; ++r)
                M[(size_t)p * n + r] = rf((double)(p - r) * m->step, prm) + (p == r ? noise_rel[q] : 0.0);
        for (int sc = 0; sc < U; ++sc)
            for (int p = 0; p < n; ++p) Rhp[(size_t)sc * n + p] = rf(sc - (m->offset + m->step * p), prm);
        if (cholesky(M, n) != 0) {
            rc = -1;
        } else {
#pragma omp parallel
            {
                cd *y = (cd *)malloc(sizeof(cd) * (size_t)n);
#pragma omp for schedule(static)
                for (int sc = 0; sc < U; ++sc) {
                    for (int p = 0; p < n; ++p) y[p] = conj(Rhp[(size_t)sc * n + p]);
                    forward_sub(M, n, y);
                    double hv = 0.0;
                    for (int p = 0; p < n; ++p) hv += creal(y[p] * conj(y[p]));
                    const double e = 1.0 - hv > 0.0 ? 1.0 - hv : 0.0;
                    const double den = hv + 1.0 - e;
                    const double s = den != 0.0 ? 2.0 / den : 0.0;
                    const double e2 = s * (s - 1.0) * hv + (1.0 - s) + s * e;
                    scale[sc] = s;
                    err[q][sc] = e2 > 0.0 ? e2 : 0.0;
                }
                free(y);
            }
            cd *z = (cd *)malloc(sizeof(cd) * (size_t)n);
            for (int rx = 0; rx < NR; ++rx)
                for (int d = 0; d < D; ++d) {
                    const double *sa = h_ls + 2 * ((((size_t)rx * g.nl + m->a) * D + d) * g.pilot_pad);
                    const double *sb = h_ls + 2 * ((((size_t)rx * g.nl + m->b) * D + d) * g.pilot_pad);
                    for (int p = 0; p < n; ++p)
                        z[p] = (sa[2 * p] + m->sb * sb[2 * p]) + I * (sa[2 * p + 1] + m->sb * sb[2 * p + 1]);
                    forward_sub(M, n, z);
                    backward_sub_h(M, n, z);
#pragma omp parallel for schedule(static)
                    for (int sc = 0; sc < U; ++sc) {
                        cd s = 0;
                        const cd *row = Rhp + (size_t)sc * n;
                        for (int p = 0; p < n; ++p) s += row[p] * z[p];
                        G[q][((size_t)rx * D + d) * U + sc] = s * scale[sc];
                    }
                }
            free(z);
        }
        free(M);
        free(Rhp);
        free(scale);
    }
    for (int l = 0; l < g.nl && rc == 0; ++l) {
        const int qa = lm[l][0], qb = lm[l][1];
        for (int rx = 0; rx < NR; ++rx) {
#pragma omp parallel for schedule(static)
            for (int sc = 0; sc < U; ++sc) {
                cd hf[4];
                double e;
                for (int d = 0; d < D; ++d) {
                    const size_t i = ((size_t)rx * D + d) * U + sc;
                    hf[d] = qb < 0 ? G[qa][i] : 0.5 * (G[qa][i] + lsgn[l] * G[qb][i]);
                }
                e = qb < 0 ? err[qa][sc] : 0.5 * (err[qa][sc] + err[qb][sc]);
                for (int s = 0; s < g.nsym; ++s) {
                    double mm[16], b[4];
                    for (int i = 0; i < D; ++i) {
                        for (int j = 0; j < D; ++j)
                            mm[i * D + j] = rt[g.dmrs_sym[i]][g.dmrs_sym[j]] + (i == j ? e : 0.0);
                        b[i] = rt[g.dmrs_sym[i]][s];
                    }
                    solve_real(mm, b, D); /* symmetric: weights = M^-1 r */
                    cd v = 0;
                    for (int d = 0; d < D; ++d) v += b[d] * hf[d];
                    double *o = h_out + 2 * ((((size_t)rx * g.nl + l) * g.nsym + s) * U + sc);
                    o[0] = creal(v);
                    o[1] = cimag(v);
                }
            }
        }
    }
    for (int q = 0; q < 4; ++q) {
        free(G[q]);
        free(err[q]);
    }
    return rc;
}

/* ------------------------------------------------------------- detection */
static void inv_small(cd *a, int n, cd *inv)
{
    cd m[4][8];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j) m[i][j] = j < n ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (cabs(m[r][c]) > cabs(m[piv][c])) piv = r;
        for (int k = 0; k < 2 * n; ++k) {
            cd t = m[c][k];
            m[c][k] = m[piv][k];
            m[piv][k] = t;
        }
        cd d = m[c][c];
        for (int k = 0; k < 2 * n; ++k) m[c][k] /= d;
        for (int r = 0; r < n; ++r)
            if (r != c) {
                cd f = m[r][c];
                for (int k = 0; k < 2 * n; ++k) m[r][k] -= f * m[c][k];
            }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) inv[i * n + j] = m[i][n + j];
}

/* Unbiased per-RE LMMSE, mimo/mimo_detect_bri_batch/scripts/
 * mimo_detect_bri_ref.py:72-86: A = H^H H + s2 I, x = A^-1 H^H y,
 * d = max(1 - s2 Re diag(A^-1), 1e-3), xhat = x / d, no_eff = (1 - d) / d. */
void or_detect_lmmse(const PuschMimoConfig *cfg, const double *h, const double *y,
                     const double *noise, double *xhat, double *no_eff)
{
    OrGeom g;
    or_geom(cfg, &g);
    const int U = g.used, L = g.nl, R = g.nr;
    const double s2 = noise[0];
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < g.nsym; ++s)
        for (int sc = 0; sc < U; ++sc) {
            cd Hm[64][4], yv[64], A[16], Ai[16], m[4];
            for (int r = 0; r < R; ++r) {
                for (int l = 0; l < L; ++l) {
                    const double *p = h + 2 * ((((size_t)r * L + l) * g.nsym + s) * U + sc);
                    Hm[r][l] = p[0] + I * p[1];
                }
                const double *q = y + 2 * (((size_t)r * g.nsym + s) * U + sc);
                yv[r] = q[0] + I * q[1];
            }
            for (int i = 0; i < L; ++i) {
                cd acc = 0;
                for (int r = 0; r < R; ++r) acc += conj(Hm[r][i]) * yv[r];
                m[i] = acc;
                for (int j = 0; j < L; ++j) {
                    cd a = 0;
                    for (int r = 0; r < R; ++r) a += conj(Hm[r][i]) * Hm[r][j];
                    A[i * L + j] = a + (i == j ? s2 : 0.0);
                }
            }
            inv_small(A, L, Ai);
            for (int i = 0; i < L; ++i) {
                cd x = 0;
                for (int j = 0; j < L; ++j) x += Ai[i * L + j] * m[j];
                double d = 1.0 - s2 * creal(Ai[i * L + i]);
                if (d < 1e-3) d = 1e-3;
                double *o = xhat + 2 * (((size_t)i * g.nsym + s) * U + sc);
                o[0] = creal(x) / d;
                o[1] = cimag(x) / d;
                no_eff[((size_t)i * g.nsym + s) * U + sc] = (1.0 - d) / d;
            }
        }
}
