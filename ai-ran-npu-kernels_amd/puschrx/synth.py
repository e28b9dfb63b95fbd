"""Synthetic PUSCH slots: transmitter (native, airan_pusch_tx_EncodeTbs),
a tapped-delay-line fading channel and OFDM modulation to time-domain IQ.

Channel model (per slot, per Rx/layer link): `taps` taps on a uniform delay
grid spanning 4x the RMS delay spread with an exponential power-delay
profile, Rayleigh gains made time-varying by a sum-of-sinusoids Jakes model
(maximum Doppler `doppler_hz`, evaluated at each OFDM symbol), and
exponential Kronecker receive correlation rho^|i-j|.  Each link has unit
average power, and `snr_db` is the per-Rx, per-layer SNR: complex AWGN of
variance 10^(-snr/10) per resource element.  The channel is applied in the
frequency domain per OFDM symbol (block fading within a symbol; the delay
spread is far below the cyclic prefix), then the grid is OFDM-modulated with a
unitary IFFT and the NR cyclic prefixes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from .abi import load_library


@dataclass
class ChannelSpec:
    taps: int = 16
    delay_spread_s: float = 300e-9
    doppler_hz: float = 100.0
    rx_corr: float = 0.5
    snr_db: float = 28.0
    scs_hz: float = 30e3
    sinusoids: int = 16


@dataclass
class Geometry:
    nr: int
    nl: int
    fft: int
    used: int
    pad: int
    samples: int
    cp_long: int
    cp_short: int
    dmrs_sym: tuple
    data_sym: tuple

    @classmethod
    def of(cls, cfg) -> "Geometry":
        fft = cfg.fft_size
        cs = fft * 9 // 128
        cl = cs + fft // 64
        dm = tuple(s for s in range(14) if (cfg.dmrs_symbol_mask >> s) & 1)
        return cls(cfg.num_rx_antennas, cfg.num_layers, fft, cfg.used_subcarriers, cfg.padded_subcarriers,
                   14 * fft + 13 * cs + cl, cl, cs, dm, tuple(s for s in range(14) if s not in dm))


def random_payload(rng: np.random.Generator, num_tbs: int, tbs: int) -> np.ndarray:
    return rng.integers(0, 2, size=(num_tbs, tbs), dtype=np.uint8)


def tx_grid(cfg, plan, payload: np.ndarray, slots_per_tb: int = 1, want_coded: bool = False,
            slot_numbers=None):
    """Resource grid complex64 [slots, L, 14, used] (and the scrambled coded
    bits [num_tbs, G] when asked).  slot_numbers: the slot number (DMRS) of
    each slot, default cfg.slot_number for all."""
    lib = load_library()
    g = Geometry.of(cfg)
    num_tbs = payload.shape[0]
    payload = np.ascontiguousarray(payload, dtype=np.uint8)
    grid = np.zeros((num_tbs * slots_per_tb, g.nl, 14, g.used), np.complex64)
    coded = np.zeros((num_tbs, plan.G), np.uint8) if want_coded else None
    sn = None
    if slot_numbers is not None:
        sn = np.ascontiguousarray(slot_numbers, dtype=np.uint16)
        if sn.shape != (num_tbs * slots_per_tb,):
            raise ValueError("slot_numbers needs one entry per slot")
    st = lib.airan_pusch_tx_EncodeTbsSlots(C.byref(cfg), C.byref(plan), num_tbs, slots_per_tb,
                                           None if sn is None else sn.ctypes.data_as(C.c_void_p),
                                           payload.ctypes.data_as(C.c_void_p), grid.ctypes.data_as(C.c_void_p),
                                           None if coded is None else coded.ctypes.data_as(C.c_void_p))
    if st != abi.OK:
        raise RuntimeError(f"airan_pusch_tx_EncodeTbsSlots failed with status {st}")
    return (grid, coded) if want_coded else grid


def _rx_corr_sqrt(nr: int, rho: float) -> np.ndarray:
    idx = np.arange(nr)
    r = rho ** np.abs(idx[:, None] - idx[None, :])
    w, v = np.linalg.eigh(r)
    return (v * np.sqrt(np.maximum(w, 0.0))) @ v.conj().T


def channel(rng: np.random.Generator, cfg, spec: ChannelSpec, num_slots: int, slot_time0: float = 0.0):
    """Frequency response H complex64 [slots, NR, L, 14, used]."""
    g = Geometry.of(cfg)
    taps = spec.taps
    delays = np.arange(taps) * (4.0 * spec.delay_spread_s / taps)
    pdp = np.exp(-delays / spec.delay_spread_s)
    pdp /= pdp.sum()
    freqs = (np.arange(g.used) - g.used // 2) * spec.scs_hz
    steer = np.exp(-2j * np.pi * freqs[None, :] * delays[:, None])  # [taps, used]
    tsym = (1.0 + 9.0 / 128.0) / spec.scs_hz
    slot_dur = 14 * tsym
    rsq = _rx_corr_sqrt(g.nr, spec.rx_corr)
    H = np.empty((num_slots, g.nr, g.nl, 14, g.used), np.complex64)
    M = spec.sinusoids
    for s in range(num_slots):
        t = slot_time0 + s * slot_dur + np.arange(14) * tsym  # [14]
        # Jakes sum of sinusoids per (rx, layer, tap).
        alpha = rng.uniform(0, 2 * np.pi, (g.nr, g.nl, taps, M))
        phi = rng.uniform(0, 2 * np.pi, (g.nr, g.nl, taps, M))
        fd = spec.doppler_hz * np.cos(alpha)
        ph = 2 * np.pi * fd[..., None] * t + phi[..., None]  # [nr, nl, taps, M, 14]
        gains = np.exp(1j * ph).sum(axis=3) / np.sqrt(M)  # unit power per tap
        gains *= np.sqrt(pdp)[None, None, :, None]
        gains = np.einsum("ij,jltk->iltk", rsq, gains)  # Rx correlation
        H[s] = np.einsum("rltk,tu->rlku", gains, steer).astype(np.complex64)
    return H


def ofdm_modulate(cfg, y: np.ndarray) -> np.ndarray:
    """Frequency grid [..., 14, used] -> time samples [..., samples] (unitary
    IFFT, NR cyclic prefixes; subcarrier u sits on bin (u - used/2) mod N)."""
    g = Geometry.of(cfg)
    lead = y.shape[:-2]
    bins = np.zeros(lead + (14, g.fft), np.complex64)
    u = np.arange(g.used)
    k = (u - g.used // 2) % g.fft
    bins[..., k] = y
    td = np.fft.ifft(bins, axis=-1) * np.sqrt(g.fft)
    out = np.empty(lead + (g.samples,), np.complex64)
    pos = 0
    for s in range(14):
        cp = g.cp_long if s == 0 else g.cp_short
        out[..., pos:pos + cp] = td[..., s, g.fft - cp:]
        out[..., pos + cp:pos + cp + g.fft] = td[..., s, :]
        pos += cp + g.fft
    return out


def receive(rng: np.random.Generator, cfg, grid: np.ndarray, H: np.ndarray, snr_db: float,
            iq_format: int = abi.IQ_INT16_Q8_8) -> np.ndarray:
    """Time-domain IQ [slots, NR, samples, 2] (int16 Q8.8 or float32)."""
    g = Geometry.of(cfg)
    S = grid.shape[0]
    sigma2 = 10.0 ** (-snr_db / 10.0)
    out_dtype = np.int16 if iq_format == abi.IQ_INT16_Q8_8 else np.float32
    iq = np.empty((S, g.nr, g.samples, 2), out_dtype)
    for s in range(S):
        y = np.einsum("rlku,lku->rku", H[s], grid[s])
        td = ofdm_modulate(cfg, y)
        td += np.sqrt(sigma2 / 2) * (rng.standard_normal(td.shape) + 1j * rng.standard_normal(td.shape))
        if iq_format == abi.IQ_INT16_Q8_8:
            q = np.clip(np.rint(np.stack([td.real, td.imag], -1) * 256.0), -32768, 32767)
            iq[s] = q.astype(np.int16)
        else:
            iq[s] = np.stack([td.real, td.imag], -1).astype(np.float32)
    return iq


def identity_channel(cfg, num_slots: int) -> np.ndarray:
    """H = [I_L; 0] per RE (needs NR >= L): a noiseless loopback."""
    g = Geometry.of(cfg)
    H = np.zeros((num_slots, g.nr, g.nl, 14, g.used), np.complex64)
    for l in range(g.nl):
        H[:, l, l] = 1.0
    return H
