// detect.h -- exact per-RE LMMSE solve (L <= 4) shared by the detector
// operator and the fused detector of the chain.  Multiply-adds contract only
// within one expression (fmuladd), so every kernel that inlines these
// functions rounds identically.
#pragma once

#include "common.h"

namespace airan {

// ------------------------------------------------------------ small solver
struct Cf {
    float r, i;
};
__device__ inline Cf cmul(Cf a, Cf b)
{
#pragma clang fp contract(on)
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ inline Cf cmulc(Cf a, Cf b)  // conj(a) * b
{
#pragma clang fp contract(on)
    return {a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r};
}
__device__ inline Cf cadd(Cf a, Cf b) { return {a.r + b.r, a.i + b.i}; }
__device__ inline Cf csub(Cf a, Cf b) { return {a.r - b.r, a.i - b.i}; }

// Solve A x = m and return diag(A^-1) for Hermitian PD A (lower triangle
// used), order NL known at compile time.  LDL^H without pivoting.
template <int NL>
__device__ inline void HermSolve(Cf (&A)[4][4], Cf (&m)[4], Cf (&x)[4], float (&dinv)[4])
{
#pragma clang fp contract(on)
    Cf Lm[4][4];
    float D[4], Dinv[4];  // pivots and their reciprocals (one division each)
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        float dj = A[j][j].r;
#pragma unroll
        for (int k = 0; k < j; ++k) dj -= D[k] * (Lm[j][k].r * Lm[j][k].r + Lm[j][k].i * Lm[j][k].i);
        D[j] = dj;
        const float inv = 1.0f / dj;
        Dinv[j] = inv;
#pragma unroll
        for (int i = j + 1; i < NL; ++i) {
            Cf s = A[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) {
                // s -= L_ik D_k conj(L_jk)
                Cf t = cmul(Lm[i][k], {Lm[j][k].r, -Lm[j][k].i});
                s = csub(s, {t.r * D[k], t.i * D[k]});
            }
            Lm[i][j] = {s.r * inv, s.i * inv};
        }
    }
    // forward: L z = m
    Cf z[4];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        Cf s = m[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s = csub(s, cmul(Lm[i][k], z[k]));
        z[i] = s;
    }
    // diagonal: w = D^-1 z, backward L^H x = w
#pragma unroll
    for (int i = NL - 1; i >= 0; --i) {
        Cf s = {z[i].r * Dinv[i], z[i].i * Dinv[i]};
#pragma unroll
        for (int k = i + 1; k < NL; ++k) s = csub(s, cmulc(Lm[k][i], x[k]));
        x[i] = s;
    }
    // diag(A^-1)_i = sum_k |(L^-1)_ki|^2 / D_k, with L^-1 by forward substitution.
#pragma unroll
    for (int c = 0; c < NL; ++c) {
        Cf col[4];
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            if (i < c) {
                col[i] = {0.0f, 0.0f};
                continue;
            }
            Cf s = {i == c ? 1.0f : 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < i; ++k)
                if (k >= c) s = csub(s, cmul(Lm[i][k], col[k]));
            col[i] = s;
            acc += (s.r * s.r + s.i * s.i) * Dinv[i];
        }
        dinv[c] = acc;
    }
}

// Unbiased LMMSE from the Gram matrix A = H^H H (lower triangle) and the
// matched filter m = H^H y: x = (A + no I)^-1 m scaled by 1 / d per layer,
// d = max(1 - no diag((A + no I)^-1), 1e-3), no_eff = (1 - d) / d.
template <int NL>
__device__ inline void DetectFromGram(float no, Cf (&A)[4][4], Cf (&m)[4], Cf (&xh)[4], float (&ne)[4])
{
#pragma clang fp contract(on)
#pragma unroll
    for (int i = 0; i < NL; ++i) A[i][i].r += no;
    Cf x[4];
    float dinv[4];
    HermSolve<NL>(A, m, x, dinv);
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const float d = fmaxf(1.0f - no * dinv[l], 1e-3f);
        const float inv = 1.0f / d;
        xh[l] = {x[l].r * inv, x[l].i * inv};
        ne[l] = (1.0f - d) * inv;
    }
}

// One thread per RE.  H(r, l) and y(r) are fetched by the caller-specific
// loader; the solve is shared.
template <int NL, typename LoadH, typename LoadY>
__device__ inline void Detect(int nr, float no, LoadH loadh, LoadY loady, Cf (&xh)[4], float (&ne)[4])
{
#pragma clang fp contract(on)
    Cf A[4][4], m[4];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        m[i] = {0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < NL; ++j) A[i][j] = {0.0f, 0.0f};
    }
    for (int r = 0; r < nr; ++r) {
        Cf h[4];
#pragma unroll
        for (int l = 0; l < NL; ++l) h[l] = loadh(r, l);
        const Cf yr = loady(r);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            m[i] = cadd(m[i], cmulc(h[i], yr));
#pragma unroll
            for (int j = 0; j <= i; ++j) A[i][j] = cadd(A[i][j], cmulc(h[i], h[j]));  // (H^H H)_ij
        }
    }
    DetectFromGram<NL>(no, A, m, xh, ne);
}

// Single-tap zero forcing (equalization/equalize_zf_siso/equalize_kernel.cpp:
// Process; operator.yaml:72-84) on fp16 inputs in fp32: g = max(hr^2 + hi^2,
// floor), x = conj(h) y / g, no_eff = n0 / g, IEEE divisions.  Products of
// fp16 values are exact in fp32, so every sum rounds once whether or not it
// is contracted (the reference's MulAddDst gives the same value).
__device__ inline void ZfSolve(float hr, float hi, float yr, float yi, float n0, float *xr, float *xi, float *ne)
{
#pragma clang fp contract(off)
    const float g = fmaxf(hr * hr + hi * hi, AIRAN_ZF_G_FLOOR);
    *xr = (hr * yr + hi * yi) / g;
    *xi = (hr * yi - hi * yr) / g;
    *ne = n0 / g;
}

}  // namespace airan
