"""GPU parity of the floating-point front end against the FP64 oracle, through
the C ABI: OFDM demodulation (FFT), LMMSE channel estimation, detector I/O
pack and per-RE LMMSE detection.

Tolerances (stated per stage, all against the FP64 oracle):
  OFDM     each fp16 output within 1 fp16 ulp (+1e-5 abs) of the exact DFT,
           NMSE <= -60 dB; fused re-demap output bit-identical to
           re_demap(stage-4 output).
  CE       NMSE <= -50 dB over the used grid against the full-rank FP64
           estimator fed the same relative noise (-40 dB when the filter is
           truncated, which the bound then also covers); the subspace-residual noise
           estimate within 5 % of a numpy eigh restatement (and within 15 %
           of the injected noise); with every eigenvector kept, the DMRS-LS
           noise passes through bit-exact and the statistics match to 1e-4.
  io_pack  bit-exact (layout transform + fp32 mean of fp16 noise).
  detect   NMSE <= -50 dB for xhat and no_eff (FP32 LDL^H vs FP64 inverse).
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from oracle_lib import OracleGeom, OrCeParams, cref, ptr
from puschrx import abi, ops
from puschrx import profiles as P

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def nmse_db(got, want):
    err = np.sum(np.abs(got - want) ** 2)
    ref = np.sum(np.abs(want) ** 2)
    return 10 * np.log10(err / ref + 1e-300)


OFDM_PROFILES = {
    "fft4096_4rx": P.FULL_CARRIER_4RX,
    "fft2048_4rx": P.MID_4RX_RANK2,
    "fft1024_1rx": P.SISO_52PRB,
    "fft512": P.Profile(name="t512", rx=2, rank=1, num_rb=32, fft=512, qm=2, padded_subcarriers=448),
    "fft256": P.Profile(name="t256", rx=3, rank=1, num_rb=16, fft=256, qm=2, padded_subcarriers=256),
    "fft128": P.Profile(name="t128", rx=2, rank=1, num_rb=8, fft=128, qm=2, padded_subcarriers=128),
}


@pytest.mark.parametrize("name", list(OFDM_PROFILES))
@pytest.mark.parametrize("fmt", [abi.IQ_INT16_Q8_8, abi.IQ_COMPLEX64])
def test_ofdm_demod_matches_dft(oracle, gpu_stream, name, fmt):
    cfg = OFDM_PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(11)
    n = g.nr * g.samples * 2
    if fmt == abi.IQ_INT16_Q8_8:
        iq_i = rng.integers(-2000, 2001, n, dtype=np.int16)
        iq_f = iq_i.astype(np.float32) / 256.0
        iq_d = dev(iq_i)
    else:
        iq_f = rng.normal(0, 3.0, n).astype(np.float32)
        iq_d = dev(iq_f)
    want = np.zeros(g.nr * 14 * g.fft * 2)
    oracle.or_ofdm_demod(cref(cfg), ptr(iq_f), ptr(want))
    want = (want[0::2] + 1j * want[1::2]).reshape(g.nr, 14, g.fft)
    sre, sim = ops.ofdm_demod(cfg, iq_d, g.nr, fmt, fused_redemap=False)
    sre = host(sre).astype(np.float64).reshape(g.nr, 14, g.fft)
    sim = host(sim).astype(np.float64).reshape(g.nr, 14, g.fft)
    P_ = g.fft // 64
    k = np.arange(g.fft)
    e = (k % P_) * 64 + k // P_
    got = sre[:, :, e] + 1j * sim[:, :, e]
    assert nmse_db(got, want) <= -60.0
    for part_got, part_want in ((got.real, want.real), (got.imag, want.imag)):
        ulp = np.spacing(np.abs(part_want).astype(np.float16)).astype(np.float64)
        assert np.all(np.abs(part_got - part_want) <= ulp + 1e-5)
    # Fused write-out == re_demap of the stage-4 planes, bit for bit.
    fre, fim = ops.ofdm_demod(cfg, iq_d, g.nr, fmt, fused_redemap=True)
    s4re, s4im = ops.ofdm_demod(cfg, iq_d, g.nr, fmt, fused_redemap=False)
    rre, rim = ops.re_demap(cfg, s4re, s4im)
    assert np.array_equal(host(fre).view(np.uint16), host(rre).view(np.uint16))
    assert np.array_equal(host(fim).view(np.uint16), host(rim).view(np.uint16))


# Stated tolerances of the matrix-core DFT against the FP64 DFT (NMSE; the
# per-rounding error of an 8-bit / 11-bit significand is about -53 / -71 dB,
# and the input, the factors and the twiddled intermediate are each rounded).
MFMA_DFT_NMSE_DB = {abi.DFT_FP16: -60.0, abi.DFT_BF16: -40.0}


@pytest.mark.parametrize("name", ["fft4096_4rx", "fft2048_4rx", "fft1024_1rx"])
@pytest.mark.parametrize("prec", [abi.DFT_FP16, abi.DFT_BF16], ids=["fp16", "bf16"])
@pytest.mark.parametrize("fmt", [abi.IQ_INT16_Q8_8, abi.IQ_COMPLEX64], ids=["int16", "c64"])
def test_ofdm_mfma_dft_precision(oracle, gpu_stream, name, prec, fmt):
    """The BF16 / FP16 matrix-core DFT (block-scaled per symbol, FP32
    accumulation) against the FP64 DFT: NMSE within the stated tolerance on
    every antenna, with antenna levels spread over 40 dB (the block scaling
    keeps weak streams as accurate as strong ones); the fused
    used-subcarrier write-out equals the stage-4 output re-demapped."""
    cfg = OFDM_PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(17)
    n = g.nr * g.samples * 2
    level = 10.0 ** rng.uniform(-1.0, 1.0, (g.nr, 1, 1))  # per-antenna level, +-20 dB
    if fmt == abi.IQ_INT16_Q8_8:
        iq_i = np.clip(np.rint(rng.normal(0, 300.0, (g.nr, g.samples, 2)) * level), -32768, 32767).astype(np.int16)
        iq_i = iq_i.ravel()
        iq_f = iq_i.astype(np.float32) / 256.0
        iq_d = dev(iq_i)
    else:
        iq_f = (rng.normal(0, 1.0, (g.nr, g.samples, 2)) * level).astype(np.float32).ravel()
        iq_d = dev(iq_f)
    want = np.zeros(g.nr * 14 * g.fft * 2)
    oracle.or_ofdm_demod(cref(cfg), ptr(iq_f), ptr(want))
    want = (want[0::2] + 1j * want[1::2]).reshape(g.nr, 14, g.fft)
    sre, sim = ops.ofdm_demod(cfg, iq_d, g.nr, fmt, fused_redemap=False, precision=prec)
    sre = host(sre).astype(np.float64).reshape(g.nr, 14, g.fft)
    sim = host(sim).astype(np.float64).reshape(g.nr, 14, g.fft)
    P_ = g.fft // 64
    k = np.arange(g.fft)
    e = (k % P_) * 64 + k // P_
    got = sre[:, :, e] + 1j * sim[:, :, e]
    worst = max(nmse_db(got[r], want[r]) for r in range(g.nr))
    print(f"{name} precision {prec}: NMSE {nmse_db(got, want):.1f} dB (worst antenna {worst:.1f} dB)")
    assert worst <= MFMA_DFT_NMSE_DB[prec]
    fre, fim = ops.ofdm_demod(cfg, iq_d, g.nr, fmt, fused_redemap=True, precision=prec)
    rre, rim = ops.re_demap(cfg, *ops.ofdm_demod(cfg, iq_d, g.nr, fmt, fused_redemap=False, precision=prec))
    assert np.array_equal(host(fre).view(np.uint16), host(rre).view(np.uint16))
    assert np.array_equal(host(fim).view(np.uint16), host(rim).view(np.uint16))


def test_ofdm_rejects_bad_args(lib, gpu_stream):
    cfg = P.SISO_52PRB.config()
    a = abi.OfdmDemodBatchOpArgsV1()
    a.abi_version, a.struct_size = 1, C.sizeof(a)
    a.iq, a.batch, a.config, a.stream = 0x1000, 1, C.pointer(cfg), 0x1
    a.iq_format = 7
    assert lib.airan_ofdm_demod_batch_ValidateOpArgs(C.byref(a)) == abi.INVALID_ARGUMENT
    a.iq_format = abi.IQ_INT16_Q8_8
    a.precision = 3
    assert lib.airan_ofdm_demod_batch_ValidateOpArgs(C.byref(a)) == abi.UNSUPPORTED_PROFILE
    a.precision = abi.DFT_BF16  # the matrix-core DFT serves FFT 1024 .. 4096; this profile is 1024
    assert lib.airan_ofdm_demod_batch_ValidateOpArgs(C.byref(a)) == abi.OK
    small = OFDM_PROFILES["fft512"].config()
    a.config = C.pointer(small)
    assert lib.airan_ofdm_demod_batch_ValidateOpArgs(C.byref(a)) == abi.UNSUPPORTED_PROFILE
    a.config = C.pointer(cfg)
    a.precision = abi.DFT_FP32
    a.batch = 0
    assert lib.airan_ofdm_demod_batch_ValidateOpArgs(C.byref(a)) == abi.INVALID_ARGUMENT


# ------------------------------------------------------------------------ CE
CE_PROFILES = {
    "siso_52prb": P.SISO_52PRB,
    "mid_4rx_rank2_occ": P.MID_4RX_RANK2,
    "rank2_two_combs": dataclasses.replace(P.MID_4RX_RANK2, name="r2comb", ports=(1000, 1002)),
    "full_carrier_4rx": P.FULL_CARRIER_4RX,
    "reference_fd16x64": P.REFERENCE_FD16X64,
}


def _tdl_freq_response(rng, n_sc, taps, tau_rms, scs):
    delays = np.sort(rng.exponential(tau_rms, taps))
    delays -= delays[0]
    pw = np.exp(-delays / tau_rms)
    pw /= pw.sum()
    a = (rng.normal(size=taps) + 1j * rng.normal(size=taps)) * np.sqrt(pw / 2)
    sc = np.arange(n_sc)
    return (a[None, :] * np.exp(-2j * np.pi * sc[:, None] * scs * delays[None, :])).sum(axis=1)


def _models(cfg, g):
    """Observation models in the library's order: (n, step, offset, a, b, sign_b)."""
    comb = [(cfg.dmrs_ports[l] - 1000) // 2 for l in range(g.nl)]
    out, done = [], set()
    for l in range(g.nl):
        if l in done:
            continue
        mates = [q for q in range(g.nl) if comb[q] == comb[l]]
        if len(mates) == 1:
            out.append((g.n_dmrs_re, 2, comb[l], l, l, 0.0))
        else:
            lp = [q for q in mates if cfg.dmrs_ports[q] in (1000, 1002)][0]
            lm = [q for q in mates if q != lp][0]
            out.append((g.n_occ, 4, comb[l], lp, lm, 1.0))
            out.append((g.n_occ, 4, comb[l] + 2, lp, lm, -1.0))
        done.update(mates)
    return out


@pytest.mark.parametrize("name", list(CE_PROFILES))
@pytest.mark.parametrize("max_rank", [512, 96])
def test_ce_lmmse_matches_fp64(oracle, gpu_stream, name, max_rank):
    cfg = CE_PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(5)
    scs, tau = 30e3, 100e-9
    sigma2 = 0.01
    comb = [(cfg.dmrs_ports[l] - 1000) // 2 for l in range(g.nl)]
    occ = [sum(c == comb[l] for c in comb) == 2 for l in range(g.nl)]
    # Physical DMRS observations: z_m = sum over the comb's layers of
    # w_l(m) h_l(k_m) + noise on k_m = comb + 2m, then the DMRS-LS outputs
    # (pair means for FD-OCC layers, z itself otherwise).
    h_ls = np.zeros((g.nr, g.nl, g.ndmrs, g.pilot_pad), np.complex128)
    for rx in range(g.nr):
        h = [_tdl_freq_response(rng, g.used, 8, tau, scs) for _ in range(g.nl)]
        for d in range(g.ndmrs):
            for c in sorted(set(comb)):
                ls = [l for l in range(g.nl) if comb[l] == c]
                k = c + 2 * np.arange(g.n_dmrs_re)
                sgn = (-1.0) ** np.arange(g.n_dmrs_re)
                z = np.sqrt(sigma2 / 2) * (rng.normal(size=k.size) + 1j * rng.normal(size=k.size))
                for l in ls:
                    odd = cfg.dmrs_ports[l] in (1001, 1003)
                    z = z + (sgn if odd else 1.0) * h[l][k]
                for l in ls:
                    if occ[l]:
                        s = -1.0 if cfg.dmrs_ports[l] in (1001, 1003) else 1.0
                        h_ls[rx, l, d, : g.n_occ] = 0.5 * (z[0::2] + s * z[1::2])
                    else:
                        h_ls[rx, l, d, : g.n_dmrs_re] = z
    h16r = h_ls.real.astype(np.float16)
    h16i = h_ls.imag.astype(np.float16)
    noise16 = np.full(g.nr, sigma2, np.float16)
    # eig_floor keeps the numerically null eigenvectors out of the kept set,
    # so the kept subspace (and the residual noise estimate) is well defined.
    filt = ops.LmmseFilter(cfg, tau_s=tau, fd_hz=0.0, scs_hz=scs, max_rank=max_rank, eig_floor=1e-12)
    models = _models(cfg, g)
    kept_layer = [filt.rank(l) for l in range(g.nl)]
    for l in range(g.nl):
        n = g.n_occ if occ[l] else g.n_dmrs_re
        assert 1 <= kept_layer[l] <= min(max_rank, n)
    sc_d = torch.zeros(g.nl * g.ndmrs * g.pilot_pad, dtype=torch.int16, device="cuda")
    cnt_d = torch.zeros(16, dtype=torch.int16, device="cuda")
    hr, hi, ws, noise_est = ops.ce_lmmse(cfg, filt, dev(h16r.ravel()), dev(h16i.ravel()), sc_d, cnt_d,
                                         dev(noise16))
    hq = h16r.astype(np.float64) + 1j * h16i.astype(np.float64)
    zq = []
    kept = []
    for (n, step, off, a, b, sb) in models:
        zq.append(hq[:, a, :, :n] + sb * hq[:, b, :, :n])
        kept.append(kept_layer[a])  # a model's pair partner has the same covariance
    # Noise: energy outside the top-`kept` eigenspace of each model's R_pp
    # (numpy eigh of the observation covariance restated here), else the
    # DMRS-LS fallback.
    dims = sum((m[0] - kq) * g.ndmrs for m, kq in zip(models, kept))
    if dims >= 32:
        res = np.zeros(g.nr)
        for (n, step, off, a, b, sb), kq, z in zip(models, kept, zq):
            dk = step * (np.arange(n)[:, None] - np.arange(n)[None, :])
            rpp = 1.0 / (1.0 + 2j * np.pi * dk * scs * tau)
            w, v = np.linalg.eigh(rpp)
            u = v[:, ::-1][:, :kq]
            for rx in range(g.nr):
                for d in range(g.ndmrs):
                    res[rx] += np.sum(np.abs(z[rx, d]) ** 2) - np.sum(np.abs(u.conj().T @ z[rx, d]) ** 2)
        s2_rx = res / dims
        assert np.allclose(host(noise_est).astype(np.float64), s2_rx, rtol=0.05)
        # The estimate recovers the injected noise (not the DMRS-LS input).
        assert np.mean(s2_rx) == pytest.approx(sigma2, rel=0.15)
    else:
        s2_rx = noise16.astype(np.float64)
        assert np.array_equal(host(noise_est).view(np.uint16), noise16.view(np.uint16))
    s2 = float(np.mean(s2_rx))
    # Filter statistics: n_q = s2 / P_q, P_q = mean |z_q|^2 - s2.
    stats = host(ws[:16].view(torch.float32))[: len(models)]
    noise_rel = np.zeros(4)
    for q, z in enumerate(zq):
        mean = np.mean(np.abs(z) ** 2)
        p_q = max(mean - s2, 1e-3 * mean)
        assert stats[q] == pytest.approx(s2 / p_q, rel=0.05 if dims >= 32 else 1e-4)
        noise_rel[q] = stats[q]
    prm = OrCeParams(tau, 0.0, scs)
    inter = np.zeros(g.nr * g.nl * g.ndmrs * g.pilot_pad * 2)
    inter[0::2] = hq.real.ravel()
    inter[1::2] = hq.imag.ravel()
    want = np.zeros(g.nr * g.nl * 14 * g.used * 2)
    assert oracle.or_ce_lmmse(cref(cfg), C.byref(prm), ptr(noise_rel), ptr(inter), ptr(want)) == 0
    want = (want[0::2] + 1j * want[1::2]).reshape(g.nr, g.nl, 14, g.used)
    got = (host(hr).astype(np.float64) + 1j * host(hi).astype(np.float64)).reshape(g.nr, 16, 14, g.pad)
    assert np.all(got[:, : g.nl, :, g.used:] == 0)
    # Full rank: the same estimator as the oracle.  Truncated: the oracle is
    # still full rank, so the bound also covers the truncation error.
    full = all(kq == m[0] for m, kq in zip(models, kept))
    assert nmse_db(got[:, : g.nl, :, : g.used], want) <= (-50.0 if full else -40.0)
    filt.close()


# ------------------------------------------------------------------ detection
DET_PROFILES = {"siso_52prb": P.SISO_52PRB, "mid_4rx_rank2": P.MID_4RX_RANK2,
                "full_carrier_4rx": P.FULL_CARRIER_4RX, "reference_fd16x64": P.REFERENCE_FD16X64}


@pytest.mark.parametrize("name", list(DET_PROFILES))
def test_io_pack_and_detect(oracle, gpu_stream, name):
    cfg = DET_PROFILES[name].config()
    g = OracleGeom(cfg)
    rng = np.random.default_rng(3)
    n_re = 14 * g.pad
    h = (rng.normal(size=(g.nr, 16, n_re)) + 1j * rng.normal(size=(g.nr, 16, n_re))) * np.sqrt(0.5)
    y = (rng.normal(size=(g.nr, n_re)) + 1j * rng.normal(size=(g.nr, n_re))) * np.sqrt(0.5 * g.nl)
    hr16, hi16 = h.real.astype(np.float16), h.imag.astype(np.float16)
    hr16[:, g.nl:] = np.float16(np.nan)  # layers >= L are never read
    yr16, yi16 = y.real.astype(np.float16), y.imag.astype(np.float16)
    noise16 = rng.uniform(0.05, 0.2, g.nr).astype(np.float16)
    hrm_r, hrm_i, yv_r, yv_i, no = ops.detect_io_pack(cfg, dev(yr16.ravel()), dev(yi16.ravel()),
                                                      dev(hr16.ravel()), dev(hi16.ravel()), dev(noise16))
    # io_pack layout, bit-exact.
    want_hrm = np.zeros((n_re, g.nr, 16), np.float16)
    want_hrm[:, :, : g.nl] = hr16[:, : g.nl].transpose(2, 0, 1)
    assert np.array_equal(host(hrm_r).view(np.uint16), want_hrm.ravel().view(np.uint16))
    want_yv = np.zeros((n_re // 8, g.nr, 16), np.float16)
    want_yv[:, :, :8] = yr16.reshape(g.nr, n_re // 8, 8).transpose(1, 0, 2)
    assert np.array_equal(host(yv_r).view(np.uint16), want_yv.ravel().view(np.uint16))
    acc = np.float32(0)
    for r in range(g.nr):
        acc = np.float32(acc + np.float32(noise16[r]))
    no_val = np.float16(acc * np.float32(1.0 / g.nr))
    assert np.all(host(no).view(np.uint16) == no_val.view(np.uint16))
    # Detection against the FP64 oracle on the same fp16 inputs.
    xr, xi, ne = ops.detect_lmmse(cfg, hrm_r, hrm_i, yv_r, yv_i, no)
    hq = (hr16[:, : g.nl].astype(np.float64) + 1j * hi16[:, : g.nl].astype(np.float64))
    hq = hq.reshape(g.nr, g.nl, 14, g.pad)[..., : g.used]
    yq = (yr16.astype(np.float64) + 1j * yi16.astype(np.float64)).reshape(g.nr, 14, g.pad)[..., : g.used]
    hin = np.zeros(hq.size * 2)
    hin[0::2], hin[1::2] = hq.real.ravel(), hq.imag.ravel()
    yin = np.zeros(yq.size * 2)
    yin[0::2], yin[1::2] = yq.real.ravel(), yq.imag.ravel()
    xw = np.zeros(g.nl * 14 * g.used * 2)
    nw = np.zeros(g.nl * 14 * g.used)
    oracle.or_detect_lmmse(cref(cfg), ptr(hin), ptr(yin), ptr(np.array([float(no_val)])), ptr(xw), ptr(nw))
    xw = (xw[0::2] + 1j * xw[1::2]).reshape(g.nl, 14, g.used)
    nw = nw.reshape(g.nl, 14, g.used)
    gx = (host(xr).astype(np.float64) + 1j * host(xi).astype(np.float64)).reshape(16, 14, g.pad)
    gn = host(ne).astype(np.float64).reshape(16, 14, g.pad)
    assert nmse_db(gx[: g.nl, :, : g.used], xw) <= -50.0
    assert nmse_db(gn[: g.nl, :, : g.used], nw) <= -50.0
    assert np.all(gx[g.nl:] == 0) and np.all(gn[g.nl:] == 0)
