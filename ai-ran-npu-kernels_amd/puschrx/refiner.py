"""Learned channel-estimation refiner (include/airan_ce_refiner.h): batches
synthesised on the device, training with the library's BF16 kernels, and
NMSE evaluation against the LMMSE estimator on the same slots.

Training slots draw, per slot, an SNR, an RMS delay spread, a maximum
Doppler and an Rx correlation from the TrainSpec ranges; the tapped-delay-
line gains follow the same model as synth.channel (uniform delay grid over
4x the delay spread, exponential power-delay profile, sum-of-sinusoids
Jakes fading per OFDM symbol, exponential Kronecker Rx correlation).
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import abi, ops
from .chain import _hip

WEIGHTS = Path(__file__).resolve().parents[1] / "weights" / "ce_refiner_4x4.f32"


@dataclass
class TrainSpec:
    snr_db: tuple = (20.0, 46.0)
    delay_spread_s: tuple = (30e-9, 500e-9)
    doppler_hz: tuple = (0.0, 500.0)
    rx_corr: tuple = (0.0, 0.7)
    taps: int = 16
    scs_hz: float = 30e3
    sinusoids: int = 16


def tdl_gains(rng: np.random.Generator, nr: int, nl: int, taps: int, delay_spread: np.ndarray,
              doppler: np.ndarray, rx_corr: np.ndarray, scs_hz: float, sinusoids: int = 16):
    """Per-slot gains complex64 [S][NR][L][taps][14] and tap delays float32 [S][taps]."""
    S = delay_spread.size
    delays = np.arange(taps)[None, :] * (4.0 * delay_spread[:, None] / taps)           # [S, T]
    pdp = np.exp(-delays / delay_spread[:, None])
    pdp /= pdp.sum(axis=1, keepdims=True)
    tsym = (1.0 + 9.0 / 128.0) / scs_hz
    t = np.arange(14) * tsym                                                          # [14]
    M = sinusoids
    alpha = rng.uniform(0, 2 * np.pi, (S, nr, nl, taps, M))
    phi = rng.uniform(0, 2 * np.pi, (S, nr, nl, taps, M))
    fd = doppler[:, None, None, None, None] * np.cos(alpha)
    ph = 2 * np.pi * fd[..., None] * t + phi[..., None]                              # [S,r,l,T,M,14]
    g = np.exp(1j * ph).sum(axis=4) / np.sqrt(M)                                      # [S,r,l,T,14]
    g *= np.sqrt(pdp)[:, None, None, :, None]
    out = np.empty_like(g)
    for s in range(S):
        idx = np.arange(nr)
        R = rx_corr[s] ** np.abs(idx[:, None] - idx[None, :])
        w, v = np.linalg.eigh(R)
        rsq = (v * np.sqrt(np.maximum(w, 0.0))) @ v.T
        out[s] = np.einsum("ij,jltk->iltk", rsq, g[s])
    return out.astype(np.complex64), delays.astype(np.float32)


def tdl_gains_device(gen: torch.Generator, nr: int, nl: int, taps: int, delay_spread: np.ndarray,
                     doppler: np.ndarray, rx_corr: np.ndarray, scs_hz: float, sinusoids: int = 16):
    """tdl_gains computed on the device with torch (the same model; a
    different random stream), returned on the host: a 64-slot batch is
    ~4 M complex exponentials, which numpy takes ~0.2 s for."""
    dev = gen.device
    ds = torch.as_tensor(delay_spread, dtype=torch.float64, device=dev)
    fdmax = torch.as_tensor(doppler, dtype=torch.float64, device=dev)
    rho = torch.as_tensor(rx_corr, dtype=torch.float64, device=dev)
    S = ds.numel()
    delays = torch.arange(taps, dtype=torch.float64, device=dev)[None, :] * (4.0 * ds[:, None] / taps)
    pdp = torch.exp(-delays / ds[:, None])
    pdp = pdp / pdp.sum(dim=1, keepdim=True)
    t = torch.arange(14, dtype=torch.float64, device=dev) * ((1.0 + 9.0 / 128.0) / scs_hz)
    M = sinusoids
    two_pi = 2.0 * np.pi
    alpha = torch.rand((S, nr, nl, taps, M), generator=gen, dtype=torch.float64, device=dev) * two_pi
    phi = torch.rand((S, nr, nl, taps, M), generator=gen, dtype=torch.float64, device=dev) * two_pi
    fd = fdmax[:, None, None, None, None] * torch.cos(alpha)
    ph = two_pi * fd[..., None] * t + phi[..., None]
    g = torch.polar(torch.ones_like(ph), ph).sum(dim=4) / np.sqrt(M)                # [S, r, l, T, 14]
    g = g * torch.sqrt(pdp)[:, None, None, :, None]
    idx = torch.arange(nr, dtype=torch.float64, device=dev)
    R = rho[:, None, None] ** (idx[:, None] - idx[None, :]).abs()                    # [S, r, r]
    w, v = torch.linalg.eigh(R)
    rsq = (v * torch.sqrt(torch.clamp_min(w, 0.0))[:, None, :]) @ v.transpose(1, 2)
    g = torch.einsum("sij,sjltk->siltk", rsq.to(g.dtype), g)
    return g.to(torch.complex64).cpu().numpy(), delays.to(torch.float32).cpu().numpy()


def nmse_db(est: torch.Tensor, ref: torch.Tensor) -> float:
    e = (est.double() - ref.double()).pow(2).sum()
    return float(10.0 * torch.log10(e / ref.double().pow(2).sum()).item())


CIN = (32, 64, 64, 64, 64, 64)
COUT = (64, 64, 64, 64, 64, 32)


def bias_slices():
    """Slices of each layer's biases in the parameter vector."""
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        off += 9 * co * ci
        out.append(slice(off, off + co))
        off += co
    return out


# First parameter of the last convolution (32 <- 64): layer i holds
# 9 cout cin weights then cout biases.
LAST_LAYER_OFFSET = (9 * 64 * 32 + 64) + 4 * (9 * 64 * 64 + 64)


class Refiner:
    """input_mode 0: the network refines the LS interpolation of the DMRS
    estimate; 1: it learns a residual correction of the LMMSE estimate
    (training batches then need `estimator`, an LmmseFilter of the same grant
    shape; in the chain the estimate is the chain's own CE output).  Fresh
    mode-1 weights start with a zero last layer, i.e. at the LMMSE estimate."""

    def __init__(self, cfg, max_slots: int, params: np.ndarray | None = None, seed: int = 1,
                 input_mode: int = 0, estimator: ops.LmmseFilter | None = None, dilation=None,
                 residual_scale: int = 0):
        self.lib = abi.load_library()
        self.cfg = cfg
        self.g = ops.Geometry(cfg)
        self.max_slots = max_slots
        self.input_mode = input_mode
        self.handle = C.c_void_p()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        ops._check("refiner Create", self.lib.airan_ce_refiner_Create(
            C.byref(cfg), max_slots, None if p is None else p.ctypes.data_as(C.c_void_p), seed,
            C.c_void_p(ops.stream_ptr()), C.byref(self.handle)))
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.dilation = [1] * 6 if dilation is None else [int(d) for d in dilation]
        d = (C.c_uint32 * 6)(*self.dilation)
        ops._check("refiner SetDilation", self.lib.airan_ce_refiner_SetDilation(self.handle, d, 6))
        self.residual_scale = int(residual_scale)
        ops._check("refiner SetResidualScale", self.lib.airan_ce_refiner_SetResidualScale(self.handle,
                                                                                          self.residual_scale))
        if input_mode:
            ops._check("refiner SetInput", self.lib.airan_ce_refiner_SetInput(
                self.handle, input_mode, estimator.handle if estimator is not None else None))
            if params is None:
                w = self.params()
                w[LAST_LAYER_OFFSET:] = 0.0
                self.set_params(w)

    # ---------------------------------------------------------------- params
    def params(self) -> np.ndarray:
        p = np.zeros(abi.REFINER_PARAMS, np.float32)
        ops._check("refiner GetParams", self.lib.airan_ce_refiner_GetParams(self.handle,
                                                                          p.ctypes.data_as(C.c_void_p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == abi.REFINER_PARAMS
        ops._check("refiner SetParams", self.lib.airan_ce_refiner_SetParams(self.handle,
                                                                          p.ctypes.data_as(C.c_void_p)))

    def save(self, path: Path = WEIGHTS, extra: dict | None = None) -> None:
        """Weights (fp32 raw) and a JSON sidecar with the input mode, the
        dilations and `extra` (training record)."""
        path.parent.mkdir(parents=True, exist_ok=True)
        self.params().tofile(path)
        meta = {"input_mode": self.input_mode, "dilation": self.dilation, "residual_scale": self.residual_scale,
                **(extra or {})}
        Path(str(path) + ".json").write_text(json.dumps(meta) + "\n")

    @staticmethod
    def load_meta(path: Path = WEIGHTS) -> dict:
        meta = Path(str(path) + ".json")
        m = json.loads(meta.read_text()) if meta.exists() else {}
        m.setdefault("input_mode", 0)
        m.setdefault("dilation", [1] * 6)
        m.setdefault("residual_scale", 0)
        return m

    @classmethod
    def from_file(cls, cfg, max_slots: int, path: Path = WEIGHTS, estimator: ops.LmmseFilter | None = None):
        """A refiner with the weights, input mode and dilations of `path`."""
        meta = cls.load_meta(path)
        return cls(cfg, max_slots=max_slots, params=cls.load_params(path), input_mode=int(meta["input_mode"]),
                   estimator=estimator, dilation=meta["dilation"], residual_scale=int(meta["residual_scale"]))

    @staticmethod
    def load_params(path: Path = WEIGHTS) -> np.ndarray:
        p = np.fromfile(path, np.float32)
        if p.size != abi.REFINER_PARAMS:
            raise ValueError(f"{path}: {p.size} values, expected {abi.REFINER_PARAMS}")
        return p

    # ----------------------------------------------------------------- batches
    def synth(self, rng: np.random.Generator, slots: int, spec: TrainSpec, seed: int,
              snr_db: np.ndarray | None = None, delay_spread: np.ndarray | None = None,
              doppler: np.ndarray | None = None, rx_corr: np.ndarray | None = None,
              gen: torch.Generator | None = None) -> None:
        """A batch from the TrainSpec ranges (or the given values); `gen`: draw
        the fading gains on the device (tdl_gains_device) instead of numpy."""
        u = lambda r: rng.uniform(r[0], r[1], slots)  # noqa: E731
        snr = u(spec.snr_db) if snr_db is None else np.broadcast_to(snr_db, slots)
        ds = u(spec.delay_spread_s) if delay_spread is None else np.broadcast_to(delay_spread, slots)
        fd = u(spec.doppler_hz) if doppler is None else np.broadcast_to(doppler, slots)
        rc = u(spec.rx_corr) if rx_corr is None else np.broadcast_to(rx_corr, slots)
        make = tdl_gains if gen is None else tdl_gains_device
        gains, tau = make(rng if gen is None else gen, self.g.nr, self.g.nl, spec.taps, np.asarray(ds, np.float64),
                          np.asarray(fd, np.float64), np.asarray(rc, np.float64), spec.scs_hz, spec.sinusoids)
        snr = np.ascontiguousarray(snr, np.float32)
        self._keep = (gains, tau, snr)  # host buffers stay alive until the copies complete
        ops._check("refiner SynthBatch", self.lib.airan_ce_refiner_SynthBatch(
            self.handle, gains.ctypes.data_as(C.c_void_p), tau.ctypes.data_as(C.c_void_p), spec.taps,
            snr.ctypes.data_as(C.c_void_p), slots, spec.scs_hz, seed, C.c_void_p(ops.stream_ptr())))
        torch.cuda.current_stream().synchronize()

    def batch(self) -> abi.AiranCeRefinerBatchV1:
        b = abi.AiranCeRefinerBatchV1()
        b.abi_version, b.struct_size = abi.ABI_VERSION, C.sizeof(b)
        ops._check("refiner Batch", self.lib.airan_ce_refiner_Batch(self.handle, C.byref(b)))
        return b

    def fetch(self, name: str, dtype, count: int) -> torch.Tensor:
        src = getattr(self.batch(), name)
        out = torch.empty(count, dtype=dtype, device="cuda")
        rc = _hip().hipMemcpyAsync(out.data_ptr(), src, count * out.element_size(), 3, ops.stream_ptr())
        if rc != 0:
            raise ops.AiranError("hipMemcpyAsync", rc)
        return out

    # ---------------------------------------------------------------- training
    def train_step(self, slots: int, lr: float) -> torch.Tensor:
        ops._check("refiner TrainStep", self.lib.airan_ce_refiner_TrainStep(
            self.handle, slots, lr, C.c_void_p(self._loss.data_ptr()), C.c_void_p(ops.stream_ptr())))
        return self._loss

    def forward(self, slots: int) -> None:
        ops._check("refiner Forward", self.lib.airan_ce_refiner_Forward(self.handle, slots,
                                                                      C.c_void_p(ops.stream_ptr())))

    def set_slot_scale(self, scale) -> None:
        """Residual scale per slot for `gradients` (host float [slots])."""
        sc = np.ascontiguousarray(scale, np.float32)
        ops._check("refiner SetSlotScale", self.lib.airan_ce_refiner_SetSlotScale(
            self.handle, sc.ctypes.data_as(C.c_void_p), sc.size))

    def gradients(self, x: torch.Tensor, target: torch.Tensor, slots: int, width: int):
        grad = torch.zeros(abi.REFINER_PARAMS, dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        ops._check("refiner Gradients", self.lib.airan_ce_refiner_Gradients(
            self.handle, C.c_void_p(x.data_ptr()), C.c_void_p(target.data_ptr()), slots, width,
            C.c_void_p(grad.data_ptr()), C.c_void_p(loss.data_ptr()), C.c_void_p(ops.stream_ptr())))
        return grad, loss

    def train(self, rng: np.random.Generator, steps: int, slots: int, spec: TrainSpec, lr: float = 1e-3,
              log_every: int = 0, seed0: int = 1000, lr_final: float | None = None, schedule=None,
              gen: torch.Generator | None = None):
        """Adam on freshly synthesised batches; returns the loss history.
        schedule(i) -> learning rate of step i (overrides lr / lr_final)."""
        hist = []
        for i in range(steps):
            self.synth(rng, slots, spec, seed0 + i, gen=gen)
            if schedule is not None:
                a = schedule(i)
            else:
                a = lr if lr_final is None else lr * (lr_final / lr) ** (i / max(steps - 1, 1))
            loss = self.train_step(slots, a)
            if log_every and (i % log_every == 0 or i == steps - 1):
                hist.append((i, float(loss.item())))
                print(f"[refiner] step {i} loss {hist[-1][1]:.3e}", flush=True)
        return hist

    # -------------------------------------------------------------- evaluation
    def evaluate(self, slots: int, lmmse: ops.LmmseFilter | None = None) -> dict:
        """NMSE (dB) against the true channel of the synthesised batch: the LS
        interpolation (network input), the refiner, and (with a filter) LMMSE."""
        n = slots * 14 * self.g.used * 32
        self.forward(slots)
        x = self.fetch("x", torch.bfloat16, n).float()
        t = self.fetch("target", torch.float32, n)
        o = self.fetch("out", torch.float32, n)
        res = {"input": nmse_db(x, t), "refiner": nmse_db(o, t)}
        if self.input_mode == 0:
            res["ls_interp"] = res["input"]
        if lmmse is not None:
            res["lmmse"] = nmse_db(self.lmmse_grid(slots, lmmse), t)
        return res

    def lmmse_grid(self, slots: int, filt: ops.LmmseFilter) -> torch.Tensor:
        """LMMSE estimate of the batch in the network layout [S][14][used][32]."""
        g = self.g
        per = g.nr * g.nl * g.ndmrs * g.pilot_pad
        hr = self.fetch("h_ls_re", torch.float16, slots * per)
        hi = self.fetch("h_ls_im", torch.float16, slots * per)
        nz = self.fetch("noise_ls", torch.float16, slots * g.nr)
        sc, cnt = pilot_tables(self.cfg)
        out = torch.empty((slots, 14, g.used, 32), dtype=torch.float32, device="cuda")
        ws = None
        for s in range(slots):
            h_re, h_im, ws, _ = ops.ce_lmmse(self.cfg, filt, hr[s * per:(s + 1) * per], hi[s * per:(s + 1) * per],
                                             sc, cnt, nz[s * g.nr:(s + 1) * g.nr], workspace=ws)
            H = torch.complex(h_re.float(), h_im.float()).reshape(g.nr, 16, 14, g.pad)[:, : g.nl, :, : g.used]
            v = torch.stack([H.real, H.imag], -1)                       # [r, l, 14, used, 2]
            out[s] = v.permute(2, 3, 0, 1, 4).reshape(14, g.used, 32)
        return out.reshape(-1)

    def close(self):
        if self.handle:
            self.lib.airan_ce_refiner_Destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pilot_tables(cfg):
    """Device pilot subcarrier table and counts (the mimo_dmrs_ls profile)."""
    g = ops.Geometry(cfg)
    sc = np.zeros(g.nl * g.ndmrs * g.pilot_pad, np.uint16)
    cnt = np.zeros(16, np.uint16)
    for l in range(g.nl):
        comb = (cfg.dmrs_ports[l] - 1000) // 2
        occ = sum((cfg.dmrs_ports[q] - 1000) // 2 == comb for q in range(g.nl)) == 2
        n = g.n_occ if occ else g.n_dmrs_re
        for d in range(g.ndmrs):
            cnt[l * g.ndmrs + d] = n
            base = (l * g.ndmrs + d) * g.pilot_pad
            sc[base: base + n] = comb + (1 + 4 * np.arange(n) if occ else 2 * np.arange(n))
    return torch.from_numpy(sc.view(np.int16)).cuda(), torch.from_numpy(cnt.view(np.int16)).cuda()
