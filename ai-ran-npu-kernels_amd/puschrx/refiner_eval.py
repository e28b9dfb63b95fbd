"""The learned CE refiner against the method without it (LMMSE), on
held-out slots: channel NMSE on device-synthesised batches and end-to-end
BLER of the batched chain with and without the refiner, both on

  * the workload channel: 16 taps, 300 ns, 100 Hz, Rx correlation 0.5 --
    the LMMSE prior (300 ns, 100 Hz) matches it exactly;
  * the training family: per slot an RMS delay spread of 30-500 ns, a
    maximum Doppler of 0-500 Hz and an Rx correlation of 0-0.7 -- the
    same fixed LMMSE prior then fits only some of the slots.

Slots for the BLER come from the device transmitter (txgpu) with seeds
disjoint from training; the chain attaches the refiner for every slot
(SNR window 0-60 dB).  Prints one JSON object.

    python -m puschrx.refiner_eval [--weights W] [--slots 64] [--snr 38 40 42]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def run(weights: Path, slots: int, snrs, nmse_snrs, seed: int = 90210) -> dict:
    import torch

    from . import abi, ops
    from . import profiles as P
    from .chain import ChainOptions, NativeChain
    from .refiner import Refiner, TrainSpec
    from .txgpu import GpuTransmitter, device_channel_family

    prof = P.FULL_CARRIER_4RX
    cfg = prof.config()
    g = ops.Geometry(cfg)
    plan = ops.tb_plan(ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate), g.cw_symbols * g.qm, g.qm,
                       g.nl)
    meta = Refiner.load_meta(weights)
    out = {"weights": str(weights), "meta": {k: meta[k] for k in ("input_mode", "dilation", "residual_scale")},
           "slots": slots, "nmse_db": {}, "bler": {}}
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        filt = ops.LmmseFilter(cfg, tau_s=300e-9, fd_hz=100.0, scs_hz=30e3, max_rank=384)
        ref = Refiner.from_file(cfg, max(slots, 8), weights, estimator=filt if meta["input_mode"] else None)
        spec = TrainSpec(snr_db=(0.0, 46.0))
        # NMSE on device-synthesised batches (8 slots per point).
        for fam in ("workload", "family"):
            for snr in nmse_snrs:
                rng = np.random.default_rng(seed + int(snr))
                kw = dict(delay_spread=300e-9, doppler=100.0, rx_corr=0.5) if fam == "workload" else {}
                ref.synth(rng, 8, spec, seed=seed + int(snr), snr_db=snr, **kw)
                r = ref.evaluate(8, lmmse=filt)
                out["nmse_db"][f"{fam}_{snr:g}dB"] = {"refiner": round(r["refiner"], 2), "lmmse": round(r["lmmse"], 2)}
                print(f"[eval] NMSE {fam} {snr} dB: refiner {r['refiner']:.2f}, LMMSE {r['lmmse']:.2f}",
                      file=sys.stderr, flush=True)
        # BLER through the chain.
        opts = ChainOptions(ldpc_iterations=8, early_stop=True, tau_s=300e-9, fd_hz=100.0, ce_rank=384,
                            iq_format=abi.IQ_COMPLEX64)
        chain = NativeChain(cfg, plan, max_slots=slots, opts=opts, use_graph=False)
        tx = GpuTransmitter(cfg, plan, max_tbs=slots)
        gen = torch.Generator(device="cuda")
        per_cb = plan.Kprime - plan.crc_cb
        for fam in ("workload", "family"):
            gen.manual_seed(seed + (0 if fam == "workload" else 1))
            rng = np.random.default_rng(seed + 7)
            if fam == "workload":
                ds, fd, rc = np.full(slots, 300e-9), np.full(slots, 100.0), np.full(slots, 0.5)
            else:
                ds, fd, rc = rng.uniform(30e-9, 500e-9, slots), rng.uniform(0, 500, slots), rng.uniform(0, 0.7, slots)
            payload = torch.randint(0, 2, (slots, plan.tbs), dtype=torch.uint8, device="cuda", generator=gen)
            gains, delays = device_channel_family(gen, cfg, slots, ds, fd, rc)
            grid = tx.encode(payload)
            for snr in snrs:
                iq = tx.channel(grid, gains, delays, float(snr), seed=seed + int(10 * snr))
                res = {}
                for leg in ("lmmse", "refiner"):
                    chain.set_refiner(ref if leg == "refiner" else None, 0.0, 60.0)
                    chain.enqueue(iq, slots)
                    ok = chain.fetch("tb_ok", torch.uint8, slots)
                    bits = chain.fetch("bits", torch.int8, slots * plan.C * plan.K)
                    torch.cuda.synchronize()
                    b = bits.view(slots, plan.C, plan.K)[:, :, :per_cb].reshape(slots, -1)[:, : plan.tbs]
                    good = (ok.bool() & (b.to(torch.uint8) == payload).all(dim=1)).float().mean().item()
                    res[leg] = round(1.0 - good, 4)
                out["bler"][f"{fam}_{snr:g}dB"] = res
                print(f"[eval] BLER {fam} {snr} dB: refiner {res['refiner']:.3f}, LMMSE {res['lmmse']:.3f}",
                      file=sys.stderr, flush=True)
        chain.set_refiner(None)
        chain.close()
        tx.close()
        ref.close()
        filt.close()
    return out


def main() -> int:
    from .refiner import WEIGHTS
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--weights", type=Path, default=WEIGHTS)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--snr", type=float, nargs="+", default=[36.0, 38.0, 40.0, 42.0])
    ap.add_argument("--nmse-snr", type=float, nargs="+", default=[20.0, 30.0, 38.0, 44.0])
    args = ap.parse_args()
    import torch
    if not torch.cuda.is_available():
        print("refiner_eval needs a GPU", file=sys.stderr)
        return 1
    t0 = time.time()
    res = run(args.weights, args.slots, args.snr, args.nmse_snr)
    res["wall_s"] = round(time.time() - t0, 1)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
