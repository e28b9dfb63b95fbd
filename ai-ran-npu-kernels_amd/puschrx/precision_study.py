"""Precision study of the batched chain (4 Rx, rank 4, 273 PRB, 256QAM at
R 0.85, 16-tap TDL at 300 ns, 100 Hz, Rx correlation 0.5): the same slots
through every precision variant of the OFDM DFT and the detector's Gram /
matched filter, reporting per variant

  * dft_nmse_db    the received grid against an FP64 DFT (numpy) of the same IQ
  * h_nmse_db      the channel estimate against the true channel
  * llr_changed / llr_sign_flips   codeword LLRs that differ from the FP32
                   chain's, and those whose sign differs
  * bler[snr]      block error rate per SNR point, and the lowest SNR with
                   BLER <= 10 %
  * ms_per_step    one HIP-graph replay of `slots` slots, and the per-stage
                   times (HIP events)

Run: python -m puschrx.precision_study [--slots 64] [--snr 20 22 ... 40]
(prints one JSON object on stdout).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

VARIANTS = {
    # name: (dft precision, gram precision)
    "fp32": (0, 0),
    "fp16": (2, 2),
    "bf16": (1, 1),
    "fp32_dft_fp16_gram": (0, 2),
}


def _log(msg):
    print(f"[precision] {msg}", file=sys.stderr, flush=True)


def run(slots: int = 64, snrs=(20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40), seed: int = 31337, steps: int = 10):
    import torch

    from . import abi, ops, synth
    from . import profiles as P
    from .chain import ChainOptions, NativeChain

    prof = P.FULL_CARRIER_4RX
    cfg = prof.config()
    g = ops.Geometry(cfg)
    plan = ops.tb_plan(ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate), g.cw_symbols * g.qm, g.qm,
                       g.nl)
    spec = synth.ChannelSpec(taps=16, delay_spread_s=300e-9, doppler_hz=100.0, rx_corr=0.5)
    rng = np.random.default_rng(seed)
    payload = synth.random_payload(rng, slots, plan.tbs)
    grid = synth.tx_grid(cfg, plan, payload)
    H = synth.channel(rng, cfg, spec, slots)
    plane = 14 * g.pad
    stream = torch.cuda.Stream()
    out = {"workload": "4Rx rank-4 273-PRB 256QAM R=873/1024, 16-tap 300 ns 100 Hz rho 0.5", "slots": slots,
           "snr_db": list(snrs), "variants": {}}
    with torch.cuda.stream(stream):
        chains = {}
        for name, (dft, gram) in VARIANTS.items():
            opts = ChainOptions(ldpc_iterations=8, early_stop=True, tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz,
                                ce_rank=384, iq_format=abi.IQ_COMPLEX64, dft_precision=dft, gram_precision=gram)
            chains[name] = NativeChain(cfg, plan, max_slots=slots, opts=opts, use_graph=True)
            out["variants"][name] = {"dft_precision": dft, "gram_precision": gram, "bler": {}, "llr_changed": {},
                                     "llr_sign_flips": {}}
        iq_dev = torch.empty((slots, g.nr, g.samples, 2), dtype=torch.float32, device="cuda")
        noise_rng = np.random.default_rng(seed + 1)
        h_true = np.transpose(H, (0, 1, 2, 3, 4))  # [S, NR, L, 14, used]
        for snr in snrs:
            iq = synth.receive(noise_rng, cfg, grid, H, float(snr), iq_format=abi.IQ_COMPLEX64)
            iq_dev.copy_(torch.from_numpy(iq))
            ref_cw = None
            for name, ch in chains.items():
                ch.enqueue(iq_dev, slots)
                ok = ch.fetch("tb_ok", torch.uint8, slots)
                cw = ch.fetch("cw_llr", torch.int16, slots * g.qm * g.cw_stride)
                torch.cuda.synchronize()
                v = out["variants"][name]
                v["bler"][snr] = 1.0 - float(ok.cpu().numpy().mean())
                cw = cw.cpu().numpy()
                if ref_cw is None:
                    ref_cw = cw
                v["llr_changed"][snr] = float(np.mean(cw != ref_cw))
                v["llr_sign_flips"][snr] = float(np.mean(np.sign(cw) != np.sign(ref_cw)))
                if snr == snrs[-1]:
                    # Stage errors at the last (highest) SNR, on the first 4 slots.
                    n = min(4, slots)
                    rr = ch.fetch("rx_grid_re", torch.float16, n * g.nr * plane)
                    ri = ch.fetch("rx_grid_im", torch.float16, n * g.nr * plane)
                    hr = ch.fetch("h_re", torch.float16, n * g.nr * g.nl * plane)
                    hi = ch.fetch("h_im", torch.float16, n * g.nr * g.nl * plane)
                    torch.cuda.synchronize()
                    y = (rr.cpu().numpy().astype(np.float64) + 1j * ri.cpu().numpy().astype(np.float64))
                    y = y.reshape(n, g.nr, 14, g.pad)[..., : g.used]
                    x = iq[:n, ..., 0].astype(np.float64) + 1j * iq[:n, ..., 1].astype(np.float64)
                    want = np.empty_like(y)
                    cl, cs, fft = g.cp_long, g.cp_short, g.fft
                    pos = 0
                    for s in range(14):
                        cp = cl if s == 0 else cs
                        seg = x[..., pos + cp: pos + cp + fft]
                        f = np.fft.fft(seg, axis=-1) / np.sqrt(fft)
                        u = np.arange(g.used)
                        want[:, :, s, :] = f[..., (u - g.used // 2) % fft]
                        pos += cp + fft
                    v["dft_nmse_db"] = float(10 * np.log10(np.sum(np.abs(y - want) ** 2) / np.sum(np.abs(want) ** 2)))
                    h = (hr.cpu().numpy().astype(np.float64) + 1j * hi.cpu().numpy().astype(np.float64))
                    h = h.reshape(n, g.nr, g.nl, 14, g.pad)[..., : g.used]
                    ht = h_true[:n].astype(np.complex128)
                    v["h_nmse_db"] = float(10 * np.log10(np.sum(np.abs(h - ht) ** 2) / np.sum(np.abs(ht) ** 2)))
            _log(f"{snr} dB: " + ", ".join(f"{k} BLER {v['bler'][snr]:.3f} (LLR sign flips "
                                           f"{v['llr_sign_flips'][snr]:.5f})" for k, v in out["variants"].items()))
        # Throughput and stage times at the last SNR point's slots.
        for name, ch in chains.items():
            v = out["variants"][name]
            for _ in range(3):
                ch.enqueue(iq_dev, slots)
            torch.cuda.synchronize()
            s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s0.record(stream)
            for _ in range(steps):
                ch.enqueue(iq_dev, slots)
            s1.record(stream)
            torch.cuda.synchronize()
            v["ms_per_step"] = s0.elapsed_time(s1) / steps
            v["gbps"] = slots * plan.tbs / (v["ms_per_step"] * 1e-3) / 1e9
            ch.set_timing(True)
            acc = {k: 0.0 for k in abi.CHAIN_STAGES}
            for _ in range(steps):
                ch.enqueue(iq_dev, slots)
                for k, t in ch.stage_times().items():
                    acc[k] += t
            ch.set_timing(False)
            v["stage_ms"] = {k: round(t / steps, 4) for k, t in acc.items()}
            ok_snrs = [s for s in snrs if v["bler"][s] <= 0.10]
            v["snr_at_bler_10pct"] = min(ok_snrs) if ok_snrs else None
            _log(f"{name}: {v['ms_per_step']:.3f} ms/step ({v['gbps']:.1f} Gbit/s), DFT NMSE "
                 f"{v['dft_nmse_db']:.1f} dB, H NMSE {v['h_nmse_db']:.2f} dB, SNR at 10% BLER "
                 f"{v['snr_at_bler_10pct']}, stages {v['stage_ms']}")
            ch.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--snr", type=float, nargs="+", default=[20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40])
    ap.add_argument("--seed", type=int, default=31337)
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args()
    import torch
    if not torch.cuda.is_available():
        print("the precision study needs a GPU", file=sys.stderr)
        return 1
    t0 = time.time()
    res = run(args.slots, [float(s) for s in args.snr], args.seed, args.steps)
    res["wall_s"] = round(time.time() - t0, 1)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
