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

/* Full-rank FP64 restatement of the reference LMMSE estimator
 * (channel_est/channel_est_lmmse_mimo/scripts/channel_est_lmmse_ref.py:
 * 50-117 with the observation model of observation_model_ref.py:41-79):
 *   frequency:  H_f = R_hp (R_pp + n_obs I)^-1 h_obs, n_obs = noise*sum(w^2)
 *   rescale:    scale = 2 raa / (hhat_var + raa - error)   (ref.py:109-116)
 *   time:       per-subcarrier Wiener weights over the DMRS symbols with
 *               R_t = J0(2 pi f_d dt) and the rescaled error (ref.py:75-88).
 * The reference truncates R_pp to its top RANK eigenpairs; with RANK equal to
 * the observation count the two are identical. */
int or_ce_lmmse(const PuschMimoConfig *cfg, const OrCeParams *prm, const double *noise_rel,
                const double *h_ls, double *h_out)
{
    OrGeom g;
    or_geom(cfg, &g);
    const int U = g.used, D = g.ndmrs;
    const double tsym = (1.0 + 9.0 / 128.0) / prm->scs_hz;
    double rt[14][14];
    for (int i = 0; i < g.nsym; ++i)
        for (int j = 0; j < g.nsym; ++j) rt[i][j] = j0(2.0 * M_PI * prm->fd_hz * (i - j) * tsym);
    int comb[4], occ[4];
    for (int l = 0; l < g.nl; ++l) comb[l] = (cfg->dmrs_ports[l] - 1000) / 2;
    for (int l = 0; l < g.nl; ++l) {
        int share = 0;
        for (int p = 0; p < g.nl; ++p) share += comb[p] == comb[l];
        occ[l] = share == 2;
    }
    for (int l = 0; l < g.nl; ++l) {
        const int n = occ[l] ? g.n_occ : g.n_dmrs_re;
        const int W = occ[l] ? 2 : 1;
        int *sup = (int *)malloc(sizeof(int) * (size_t)n * W);
        for (int p = 0; p < n; ++p) {
            if (occ[l]) {
                sup[2 * p] = comb[l] + 4 * p;
                sup[2 * p + 1] = comb[l] + 4 * p + 2;
            } else {
                sup[p] = comb[l] + 2 * p;
            }
        }
        const double w = occ[l] ? 0.5 : 1.0;
        const double nobs = noise_rel[l] * (double)W * w * w;
        cd *M = (cd *)malloc(sizeof(cd) * (size_t)n * n);
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q) {
                cd s = 0;
                for (int a = 0; a < W; ++a)
                    for (int b = 0; b < W; ++b) s += w * w * rf(sup[p * W + a] - sup[q * W + b], prm);
                M[(size_t)p * n + q] = s + (p == q ? nobs : 0.0);
            }
        cd *Rhp = (cd *)malloc(sizeof(cd) * (size_t)U * n);
        for (int sc = 0; sc < U; ++sc)
            for (int p = 0; p < n; ++p) {
                cd s = 0;
                for (int a = 0; a < W; ++a) s += w * rf(sc - sup[p * W + a], prm);
                Rhp[(size_t)sc * n + p] = s;
            }
        if (cholesky(M, n) != 0) return -1;
        double *scale = (double *)malloc(sizeof(double) * U);
        double *err = (double *)malloc(sizeof(double) * U);
#pragma omp parallel
        {
            cd *y = (cd *)malloc(sizeof(cd) * (size_t)n);
#pragma omp for schedule(static)
            for (int sc = 0; sc < U; ++sc) {
                for (int p = 0; p < n; ++p) y[p] = conj(Rhp[(size_t)sc * n + p]);
                forward_sub(M, n, y);
                double hv = 0.0;
                for (int p = 0; p < n; ++p) hv += creal(y[p] * conj(y[p]));
                const double raa = 1.0;
                double e = raa - hv > 0.0 ? raa - hv : 0.0;
                double den = hv + raa - e;
                double s = den != 0.0 ? 2.0 * raa / den : 0.0;
                double e2 = s * (s - 1.0) * hv + (1.0 - s) * raa + s * e;
                scale[sc] = s;
                err[sc] = e2 > 0.0 ? e2 : 0.0;
            }
            free(y);
        }
        cd *hf = (cd *)malloc(sizeof(cd) * (size_t)D * U);
        cd *z = (cd *)malloc(sizeof(cd) * (size_t)n);
        for (int rx = 0; rx < g.nr; ++rx) {
            for (int d = 0; d < D; ++d) {
                const double *src = h_ls + 2 * ((((size_t)rx * g.nl + l) * D + d) * g.pilot_pad);
                for (int p = 0; p < n; ++p) z[p] = src[2 * p] + I * src[2 * p + 1];
                forward_sub(M, n, z);
                backward_sub_h(M, n, z);
#pragma omp parallel for schedule(static)
                for (int sc = 0; sc < U; ++sc) {
                    cd s = 0;
                    const cd *row = Rhp + (size_t)sc * n;
                    for (int p = 0; p < n; ++p) s += row[p] * z[p];
                    hf[(size_t)d * U + sc] = s * scale[sc];
                }
            }
#pragma omp parallel for schedule(static)
            for (int sc = 0; sc < U; ++sc) {
                for (int s = 0; s < g.nsym; ++s) {
                    double m[16], b[4];
                    for (int i = 0; i < D; ++i) {
                        for (int j = 0; j < D; ++j)
                            m[i * D + j] = rt[g.dmrs_sym[i]][g.dmrs_sym[j]] + (i == j ? err[sc] : 0.0);
                        b[i] = rt[g.dmrs_sym[i]][s];
                    }
                    solve_real(m, b, D); /* symmetric: weights = M^-1 r */
                    cd v = 0;
                    for (int d = 0; d < D; ++d) v += b[d] * hf[(size_t)d * U + sc];
                    double *o = h_out + 2 * ((((size_t)rx * g.nl + l) * g.nsym + s) * U + sc);
                    o[0] = creal(v);
                    o[1] = cimag(v);
                }
            }
        }
        free(z);
        free(hf);
        free(scale);
        free(err);
        free(Rhp);
        free(M);
        free(sup);
    }
    return 0;
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
