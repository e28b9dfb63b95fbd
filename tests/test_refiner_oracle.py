"""Pins of the refiner oracle (oracle/or_refiner.c), CPU only: its analytic
gradients against central finite differences of its own loss (exact mode),
and its bf16 rounding against numpy's float32 bit manipulation."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle_lib import load_oracle, ptr

CIN = [32, 64, 64, 64, 64, 64]
COUT = [64, 64, 64, 64, 64, 32]


def refiner_params(oracle, rng, scale=1.0):
    n = oracle.or_refiner_param_offset(6)
    p = np.zeros(n, np.float32)
    for i in range(6):
        o = oracle.or_refiner_param_offset(i)
        nw = 9 * COUT[i] * CIN[i]
        p[o:o + nw] = rng.normal(0, scale * np.sqrt(2.0 / (9 * CIN[i])), nw)
        p[o + nw:o + nw + COUT[i]] = rng.normal(0, 0.05, COUT[i])
    return p


def oracle_loss_grad(oracle, params, x, t, S, W, rounding):
    grad = np.zeros(params.size, np.float64)
    loss = np.zeros(1, np.float64)
    assert oracle.or_refiner_grad(ptr(params), ptr(x), ptr(t), S, W, rounding, ptr(grad), ptr(loss), None) == 0
    return float(loss[0]), grad


@pytest.fixture(scope="module")
def oracle():
    return load_oracle()


def test_param_count_matches_header(oracle):
    import re
    from pathlib import Path
    hdr = (Path(__file__).resolve().parents[1] / "include" / "airan_ce_refiner.h").read_text()
    n = int(re.search(r"#define AIRAN_REFINER_PARAMS (\d+)u", hdr).group(1))
    assert oracle.or_refiner_param_offset(6) == n == sum(9 * a * b + b for a, b in zip(CIN, COUT))


def test_gradients_match_finite_differences(oracle):
    rng = np.random.default_rng(5)
    S, W = 1, 5
    params = refiner_params(oracle, rng)
    x = rng.normal(0, 1, (S, 14, W, 32)).astype(np.float32)
    t = (x + rng.normal(0, 0.3, x.shape)).astype(np.float32)
    _, grad = oracle_loss_grad(oracle, params, x, t, S, W, 0)
    picks = []
    for i in range(6):
        o = oracle.or_refiner_param_offset(i)
        nw = 9 * COUT[i] * CIN[i]
        picks += list(rng.integers(o, o + nw, 6)) + list(rng.integers(o + nw, o + nw + COUT[i], 2))
    ok = 0
    for k in picks:
        eps = 1e-4
        pp, pm = params.copy(), params.copy()
        pp[k] = np.float32(params[k] + eps)
        pm[k] = np.float32(params[k] - eps)
        lp, _ = oracle_loss_grad(oracle, pp, x, t, S, W, 0)
        lm, _ = oracle_loss_grad(oracle, pm, x, t, S, W, 0)
        fd = (lp - lm) / (float(pp[k]) - float(pm[k]))
        ok += abs(fd - grad[k]) <= 1e-5 * max(abs(fd), abs(grad[k])) + 1e-10
    # A ReLU kink inside +-eps can spoil an occasional finite difference.
    assert ok >= len(picks) - 3, (ok, len(picks))


def test_zero_network_is_identity(oracle):
    """With every weight and bias zero the residual network returns its input
    and the loss is the plain mean squared error."""
    rng = np.random.default_rng(2)
    S, W = 2, 4
    params = np.zeros(oracle.or_refiner_param_offset(6), np.float32)
    x = rng.normal(0, 1, (S, 14, W, 32)).astype(np.float32)
    t = rng.normal(0, 1, x.shape).astype(np.float32)
    out = np.zeros_like(x)
    grad = np.zeros(params.size)
    loss = np.zeros(1)
    oracle.or_refiner_grad(ptr(params), ptr(x), ptr(t), S, W, 1, ptr(grad), ptr(loss), ptr(out))
    assert np.array_equal(out, x)
    assert loss[0] == pytest.approx(np.mean((x.astype(np.float64) - t) ** 2), rel=1e-12)
    # Only the last layer's bias sees a gradient (the hidden activations are zero).
    o5 = oracle.or_refiner_param_offset(5) + 9 * 32 * 64
    assert np.count_nonzero(grad[:o5]) == 0
    e = (x.astype(np.float64) - t).reshape(-1, 32)
    want = np.array([oracle.or_bf16_round(v) for v in (2 * e / e.size).reshape(-1)]).reshape(e.shape).sum(0)
    assert np.allclose(grad[o5:], want, rtol=1e-12)


def test_bf16_rounding(oracle):
    rng = np.random.default_rng(9)
    v = rng.normal(0, 10, 2000).astype(np.float32)
    u = v.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)
    got = np.array([oracle.or_bf16_round(float(a)) for a in v], np.float32)
    assert np.array_equal(got, r)
    assert oracle.or_bf16_round(1.0 + 2 ** -8) == 1.0          # tie to even
    assert oracle.or_bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def oracle_loss_grad_dil(oracle, params, x, t, S, W, rounding, dil):
    grad = np.zeros(params.size, np.float64)
    loss = np.zeros(1, np.float64)
    out = np.zeros(x.size, np.float32)
    d = np.asarray(dil, np.int32)
    assert oracle.or_refiner_grad_dil(ptr(params), ptr(x), ptr(t), S, W, rounding, ptr(d), ptr(grad), ptr(loss),
                                      ptr(out)) == 0
    return float(loss[0]), grad, out.reshape(x.shape)


def test_dilated_gradients_match_finite_differences(oracle):
    """Per-layer subcarrier dilations (taps at -d, 0, +d): analytic gradients
    against central differences of the loss, with dilations that reach past
    the edges of a 9-subcarrier input."""
    rng = np.random.default_rng(8)
    S, W = 1, 9
    dil = [1, 2, 3, 1, 4, 2]
    params = refiner_params(oracle, rng)
    x = rng.normal(0, 1, (S, 14, W, 32)).astype(np.float32)
    t = (x + rng.normal(0, 0.3, x.shape)).astype(np.float32)
    _, grad, _ = oracle_loss_grad_dil(oracle, params, x, t, S, W, 0, dil)
    picks = []
    for i in range(6):
        o = oracle.or_refiner_param_offset(i)
        nw = 9 * COUT[i] * CIN[i]
        picks += list(rng.integers(o, o + nw, 6)) + list(rng.integers(o + nw, o + nw + COUT[i], 2))
    ok = 0
    for k in picks:
        eps = 1e-4
        pp, pm = params.copy(), params.copy()
        pp[k] = np.float32(params[k] + eps)
        pm[k] = np.float32(params[k] - eps)
        lp, _, _ = oracle_loss_grad_dil(oracle, pp, x, t, S, W, 0, dil)
        lm, _, _ = oracle_loss_grad_dil(oracle, pm, x, t, S, W, 0, dil)
        fd = (lp - lm) / (float(pp[k]) - float(pm[k]))
        ok += abs(fd - grad[k]) <= 1e-5 * max(abs(fd), abs(grad[k])) + 1e-10
    assert ok >= len(picks) - 3, (ok, len(picks))


def test_uniform_dilation_splits_into_phases(oracle):
    """Every layer dilated by d: the subcarriers of each phase (w mod d) only
    see each other, so the network on the full width equals the undilated
    network on each phase's subsequence; the loss is their size-weighted
    mean and the gradients their sum with the same weights."""
    rng = np.random.default_rng(9)
    S, W, d = 2, 11, 3
    params = refiner_params(oracle, rng)
    x = rng.normal(0, 1, (S, 14, W, 32)).astype(np.float32)
    t = (x + rng.normal(0, 0.3, x.shape)).astype(np.float32)
    loss, grad, out = oracle_loss_grad_dil(oracle, params, x, t, S, W, 0, [d] * 6)
    want_loss, want_grad = 0.0, np.zeros_like(grad)
    for r in range(d):
        xs = np.ascontiguousarray(x[:, :, r::d, :])
        ts = np.ascontiguousarray(t[:, :, r::d, :])
        w = xs.shape[2]
        l_r, g_r, o_r = oracle_loss_grad_dil(oracle, params, xs, ts, S, w, 0, [1] * 6)
        np.testing.assert_allclose(out[:, :, r::d, :], o_r, rtol=1e-6, atol=1e-6)
        want_loss += l_r * w / W
        want_grad += g_r * w / W
    assert loss == pytest.approx(want_loss, rel=1e-10)
    np.testing.assert_allclose(grad, want_grad, rtol=1e-9, atol=1e-12)


def test_residual_scale_gradients_match_finite_differences(oracle):
    """out = x + s_slot z_5 with a scale per slot: analytic gradients against
    central differences, and the scaled network equals the unscaled one with
    the last layer's weights and bias multiplied by s (one slot)."""
    rng = np.random.default_rng(10)
    S, W = 2, 6
    params = refiner_params(oracle, rng)
    x = rng.normal(0, 1, (S, 14, W, 32)).astype(np.float32)
    t = (x + rng.normal(0, 0.3, x.shape)).astype(np.float32)
    sc = np.array([0.25, 0.05], np.float32)
    dil = np.array([1, 2, 1, 1, 3, 1], np.int32)

    def run(p, scale):
        grad = np.zeros(p.size, np.float64)
        loss = np.zeros(1, np.float64)
        out = np.zeros(x.size, np.float32)
        assert oracle.or_refiner_grad_full(ptr(p), ptr(x), ptr(t), S, W, 0, ptr(dil),
                                           None if scale is None else ptr(scale), ptr(grad), ptr(loss), ptr(out)) == 0
        return float(loss[0]), grad, out.reshape(x.shape)

    _, grad, out = run(params, sc)
    picks = []
    for i in range(6):
        o = oracle.or_refiner_param_offset(i)
        nw = 9 * COUT[i] * CIN[i]
        picks += list(rng.integers(o, o + nw, 5)) + list(rng.integers(o + nw, o + nw + COUT[i], 2))
    ok = 0
    for k in picks:
        eps = 1e-4
        pp, pm = params.copy(), params.copy()
        pp[k] = np.float32(params[k] + eps)
        pm[k] = np.float32(params[k] - eps)
        fd = (run(pp, sc)[0] - run(pm, sc)[0]) / (float(pp[k]) - float(pm[k]))
        ok += abs(fd - grad[k]) <= 1e-5 * max(abs(fd), abs(grad[k])) + 1e-10
    assert ok >= len(picks) - 3, (ok, len(picks))
    # Slot 0 alone: scaling the last layer by s is the same network.
    o5 = oracle.or_refiner_param_offset(5)
    p2 = params.copy()
    p2[o5:] = (p2[o5:].astype(np.float64) * float(sc[0])).astype(np.float32)
    x0 = np.ascontiguousarray(x[:1])
    t0 = np.ascontiguousarray(t[:1])
    grad2 = np.zeros(params.size, np.float64)
    loss2 = np.zeros(1, np.float64)
    out2 = np.zeros(x0.size, np.float32)
    assert oracle.or_refiner_grad_full(ptr(p2), ptr(x0), ptr(t0), 1, W, 0, ptr(dil), None, ptr(grad2), ptr(loss2),
                                       ptr(out2)) == 0
    np.testing.assert_allclose(out[0], out2.reshape(x0.shape)[0], rtol=1e-5, atol=1e-6)
