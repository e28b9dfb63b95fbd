"""Where the operating point's dB go: the full-carrier workload (4 Rx, rank
4, 273 PRB, 256QAM at R 0.85, 16-tap TDL at 300 ns, 100 Hz, Rx correlation
0.5) through the operator chain on the GPU with parts of the receiver
replaced by their genie counterparts, on the same slots and noise:

  lmmse             the chain as built: LMMSE CE, detector noise from the CE
                    residual subspace
  lmmse_true_noise  LMMSE CE, detector noise = the true per-RE noise
  genie_h           the true channel (rounded to the chain's fp16 grid) and
                    the true noise into the detector: no CE loss left
  genie_h_20it      genie_h with 20 LDPC iterations instead of 8

Per variant and SNR: BLER over `slots` slots (64 by default), the lowest SNR
with BLER <= 10 %, and for the estimated variants the channel NMSE against
the true channel and the ratio of the noise estimate to the true noise.
The SNR is per receive antenna and layer (every link has unit power, so the
total SNR per receive antenna is 6 dB higher at rank 4).

Run: python -m puschrx.loss_study [--slots 64] [--snr 30 32 ... 44]
(prints one JSON object on stdout).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

VARIANTS = ("lmmse", "lmmse_true_noise", "genie_h", "genie_h_20it")


def _log(msg):
    print(f"[loss] {msg}", file=sys.stderr, flush=True)


def run(slots: int = 64, snrs=(30, 32, 34, 36, 38, 40, 42), seed: int = 4242):
    import torch

    from . import abi, ops, synth
    from . import profiles as P
    from .chain import ChainOptions, OpChain

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
    noise_rng = np.random.default_rng(seed + 1)
    per_cb = plan.Kprime - plan.crc_cb
    out = {"workload": "4Rx rank-4 273-PRB 256QAM R=873/1024, 16-tap 300 ns 100 Hz rho 0.5",
           "snr_definition": "per Rx antenna and layer", "slots": slots, "snr_db": list(snrs),
           "variants": {v: {"bler": {}} for v in VARIANTS}, "h_nmse_db": {}, "noise_ratio": {}}
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        opts = ChainOptions(ldpc_iterations=8, early_stop=True, tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz,
                            ce_rank=384, iq_format=abi.IQ_COMPLEX64)
        ch = OpChain(cfg, plan, opts)
        dmrs_re, dmrs_im = ch.dmrs(cfg.slot_number)
        nplane = g.nr * 16 * 14 * g.pad

        def genie_grid(s):
            """True channel of slot s in the CE output layout [NR][16][14][padded] (fp16 planes)."""
            buf = np.zeros((2, g.nr, 16, 14, g.pad), np.float16)
            buf[0, :, : g.nl, :, : g.used] = H[s].real
            buf[1, :, : g.nl, :, : g.used] = H[s].imag
            t = torch.from_numpy(buf.reshape(-1)).cuda()
            return t[:nplane], t[nplane:]

        def decode(rx_re, rx_im, hg_re, hg_im, noise_det, iters):
            hrm_r, hrm_i, yv_r, yv_i, no = ops.detect_io_pack(cfg, rx_re, rx_im, hg_re, hg_im, noise_det)
            xr, xi, ne = ops.detect_lmmse(cfg, hrm_r, hrm_i, yv_r, yv_i, no)
            llr = ops.qam_demod(cfg, xr, xi, torch.clamp_min(ne, ch.no_eff_floor))
            cw = ops.layer_demap(cfg, llr)
            cw_nr = ops.descramble(cfg, cw, 1, restart_per_slot=True)
            lam = ops.rate_dematch(cfg, cw_nr, 1, plan)
            bits, _, _ = ops.ldpc_decode(lam, plan.C, plan.Z, plan.K, iters, ch.active_rows, opts.ldpc_mode, True)
            _, tb_ok, _ = ops.crc_check(plan, 1, bits)
            return bits, int(tb_ok.item())

        for snr in snrs:
            sigma2 = 10.0 ** (-snr / 10.0)
            iq = synth.receive(noise_rng, cfg, grid, H, float(snr), iq_format=abi.IQ_COMPLEX64)
            ok = {v: 0 for v in VARIANTS}
            err_h, pow_h, ratios = 0.0, 0.0, []
            true_noise = torch.full((g.nr,), sigma2, dtype=torch.float16, device="cuda")
            for s in range(slots):
                x = torch.from_numpy(iq[s]).cuda()
                rx_re, rx_im = ops.ofdm_demod(cfg, x, g.nr, abi.IQ_COMPLEX64, fused_redemap=True)
                h_re, h_im, sc, cnt, noise = ops.dmrs_ls(cfg, rx_re, rx_im, dmrs_re, dmrs_im)
                hg_re, hg_im, _, noise_ce = ops.ce_lmmse(cfg, ch.filter, h_re, h_im, sc, cnt, noise)
                he = (hg_re.float() + 1j * hg_im.float()).view(g.nr, 16, 14, g.pad)[:, : g.nl, :, : g.used]
                ht = torch.from_numpy(H[s]).cuda()
                err_h += float((he - ht).abs().pow(2).sum())
                pow_h += float(ht.abs().pow(2).sum())
                ratios.append(float(noise_ce.float().mean()) / sigma2)
                gr, gi = genie_grid(s)
                runs = {"lmmse": (hg_re, hg_im, noise_ce, 8), "lmmse_true_noise": (hg_re, hg_im, true_noise, 8),
                        "genie_h": (gr, gi, true_noise, 8), "genie_h_20it": (gr, gi, true_noise, 20)}
                for v, (a, b, n, it) in runs.items():
                    bits, tb = decode(rx_re, rx_im, a, b, n, it)
                    if tb:
                        got = bits.cpu().numpy().reshape(plan.C, plan.K)[:, :per_cb].reshape(-1)[: plan.tbs]
                        tb = int(np.array_equal(got.astype(np.uint8), payload[s]))
                    ok[v] += tb
            torch.cuda.synchronize()
            key = str(float(snr))
            for v in VARIANTS:
                out["variants"][v]["bler"][key] = round(1.0 - ok[v] / slots, 4)
            out["h_nmse_db"][key] = round(10 * np.log10(err_h / pow_h), 2)
            out["noise_ratio"][key] = round(float(np.mean(ratios)), 3)
            _log(f"snr {snr} dB: " + ", ".join(f"{v} {1.0 - ok[v] / slots:.3f}" for v in VARIANTS)
                 + f"; H NMSE {out['h_nmse_db'][key]} dB, noise estimate / true {out['noise_ratio'][key]}")
        ch.close()
    for v in VARIANTS:
        b = out["variants"][v]["bler"]
        passing = [float(k) for k, x in b.items() if x <= 0.10]
        out["variants"][v]["snr_at_bler_10pct"] = min(passing) if passing else None
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--snr", type=float, nargs="+", default=[30, 32, 34, 36, 38, 40, 42])
    ap.add_argument("--seed", type=int, default=4242)
    args = ap.parse_args()
    import torch
    if not torch.cuda.is_available():
        print("the loss study needs a GPU", file=sys.stderr)
        return 1
    t0 = time.time()
    res = run(args.slots, [float(s) for s in args.snr], args.seed)
    res["wall_s"] = round(time.time() - t0, 1)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
