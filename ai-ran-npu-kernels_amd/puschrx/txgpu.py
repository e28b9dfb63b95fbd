"""The synthetic transmitter and fading channel on the device
(airan_pusch_txgpu_*, include/airan_pusch_tx.h): payload bits in device
memory -> resource grid -> tapped-delay-line channel, OFDM modulation and
AWGN -> the chain's IQ input, without a host round trip.

The channel model is synth.channel's: `taps` taps on a uniform delay grid
over 4x the RMS delay spread with an exponential power-delay profile,
Rayleigh gains made time-varying by a sum of sinusoids (Jakes, maximum
Doppler per slot) evaluated at each OFDM symbol, and exponential Kronecker
receive correlation; the gains are drawn on the device with torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi, ops
from .synth import ChannelSpec


def device_channel(gen: torch.Generator, cfg, spec: ChannelSpec, slots: int):
    """Gains complex64 [slots, NR, L, taps, 14] and tap delays float32
    [slots, taps] on the device (one independent channel per slot)."""
    g = ops.Geometry(cfg)
    dev = gen.device
    taps, M = spec.taps, spec.sinusoids
    delays = torch.arange(taps, dtype=torch.float64, device=dev) * (4.0 * spec.delay_spread_s / taps)
    pdp = torch.exp(-delays / spec.delay_spread_s)
    pdp = pdp / pdp.sum()
    t = torch.arange(14, dtype=torch.float64, device=dev) * ((1.0 + 9.0 / 128.0) / spec.scs_hz)
    two_pi = 2.0 * np.pi
    alpha = torch.rand((slots, g.nr, g.nl, taps, M), generator=gen, dtype=torch.float64, device=dev) * two_pi
    phi = torch.rand((slots, g.nr, g.nl, taps, M), generator=gen, dtype=torch.float64, device=dev) * two_pi
    fd = spec.doppler_hz * torch.cos(alpha)
    ph = two_pi * fd[..., None] * t + phi[..., None]
    gains = torch.polar(torch.ones_like(ph), ph).sum(dim=4) / np.sqrt(M)            # [S, r, l, T, 14]
    gains = gains * torch.sqrt(pdp)[None, None, None, :, None]
    idx = torch.arange(g.nr, dtype=torch.float64, device=dev)
    R = spec.rx_corr ** (idx[:, None] - idx[None, :]).abs()
    w, v = torch.linalg.eigh(R)
    rsq = (v * torch.sqrt(torch.clamp_min(w, 0.0))) @ v.T
    gains = torch.einsum("ij,sjltk->siltk", rsq.to(gains.dtype), gains)
    return gains.to(torch.complex64).contiguous(), delays.to(torch.float32).expand(slots, taps).contiguous()


def device_channel_family(gen: torch.Generator, cfg, slots: int, delay_spread_s, doppler_hz, rx_corr,
                          taps: int = 16, sinusoids: int = 16, scs_hz: float = 30e3):
    """device_channel with per-slot parameters (arrays of `slots` values):
    gains complex64 [slots, NR, L, taps, 14], delays float32 [slots, taps]."""
    g = ops.Geometry(cfg)
    dev = gen.device
    f64 = dict(dtype=torch.float64, device=dev)
    ds = torch.as_tensor(np.asarray(delay_spread_s, np.float64), **f64)
    fdmax = torch.as_tensor(np.asarray(doppler_hz, np.float64), **f64)
    rho = torch.as_tensor(np.asarray(rx_corr, np.float64), **f64)
    delays = torch.arange(taps, **f64)[None, :] * (4.0 * ds[:, None] / taps)
    pdp = torch.exp(-delays / ds[:, None])
    pdp = pdp / pdp.sum(dim=1, keepdim=True)
    t = torch.arange(14, **f64) * ((1.0 + 9.0 / 128.0) / scs_hz)
    two_pi = 2.0 * np.pi
    alpha = torch.rand((slots, g.nr, g.nl, taps, sinusoids), generator=gen, **f64) * two_pi
    phi = torch.rand((slots, g.nr, g.nl, taps, sinusoids), generator=gen, **f64) * two_pi
    fd = fdmax[:, None, None, None, None] * torch.cos(alpha)
    ph = two_pi * fd[..., None] * t + phi[..., None]
    gains = torch.polar(torch.ones_like(ph), ph).sum(dim=4) / np.sqrt(sinusoids)
    gains = gains * torch.sqrt(pdp)[:, None, None, :, None]
    idx = torch.arange(g.nr, **f64)
    R = rho[:, None, None] ** (idx[:, None] - idx[None, :]).abs()
    w, v = torch.linalg.eigh(R)
    rsq = (v * torch.sqrt(torch.clamp_min(w, 0.0))[:, None, :]) @ v.transpose(1, 2)
    gains = torch.einsum("sij,sjltk->siltk", rsq.to(gains.dtype), gains)
    return gains.to(torch.complex64).contiguous(), delays.to(torch.float32).contiguous()


class GpuTransmitter:
    """Encoder + channel for one grant shape and up to max_tbs TBs per call."""

    def __init__(self, cfg, plan, max_tbs: int, slots_per_tb: int = 1):
        self.lib = abi.load_library()
        self.cfg, self.plan, self.T = cfg, plan, slots_per_tb
        self.g = ops.Geometry(cfg)
        self.max_tbs = max_tbs
        self.handle = C.c_void_p()
        ops._check("txgpu Create", self.lib.airan_pusch_txgpu_Create(C.byref(cfg), C.byref(plan), max_tbs,
                                                                     slots_per_tb, C.byref(self.handle)))

    def encode(self, payload: torch.Tensor, slot_numbers: torch.Tensor | None = None, want_coded: bool = False):
        """payload uint8 [n, tbs] on the device -> grid complex64 [n T, L, 14,
        used] (and the scrambled coded bits uint8 [n, G])."""
        n = payload.shape[0]
        g = self.g
        grid = torch.empty((n * self.T, g.nl, 14, g.used), dtype=torch.complex64, device=payload.device)
        coded = torch.empty((n, self.plan.G), dtype=torch.uint8, device=payload.device) if want_coded else None
        sn = None if slot_numbers is None else C.c_void_p(slot_numbers.contiguous().data_ptr())
        ops._check("txgpu EncodeTbs", self.lib.airan_pusch_txgpu_EncodeTbs(
            self.handle, C.c_void_p(payload.contiguous().data_ptr()), n, sn, C.c_void_p(grid.data_ptr()),
            None if coded is None else C.c_void_p(coded.data_ptr()), C.c_void_p(ops.stream_ptr())))
        return (grid, coded) if want_coded else grid

    def channel(self, grid: torch.Tensor, gains: torch.Tensor, delays: torch.Tensor, snr_db: float, seed: int,
                iq_format: int = abi.IQ_COMPLEX64, scs_hz: float = 30e3) -> torch.Tensor:
        """IQ [slots, NR, samples, 2] (float32 or int16) on the device; AWGN of
        variance 10^(-snr/10) per sample (per RE after the unitary DFT)."""
        g = self.g
        slots, taps = grid.shape[0], delays.shape[1]
        dtype = torch.float32 if iq_format == abi.IQ_COMPLEX64 else torch.int16
        iq = torch.empty((slots, g.nr, g.samples, 2), dtype=dtype, device=grid.device)
        ops._check("txgpu Channel", self.lib.airan_pusch_txgpu_Channel(
            self.handle, C.c_void_p(grid.data_ptr()), slots, C.c_void_p(gains.contiguous().data_ptr()),
            C.c_void_p(delays.contiguous().data_ptr()), taps, scs_hz, 10.0 ** (-snr_db / 10.0), seed, iq_format,
            C.c_void_p(iq.data_ptr()), C.c_void_p(ops.stream_ptr())))
        return iq

    def close(self):
        if self.handle:
            self.lib.airan_pusch_txgpu_Destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
