"""Train the learned CE refiner on device-synthesised slots and write its
weights (fp32, AIRAN_REFINER_PARAMS values) for the receive chain.

    python -m puschrx.train_refiner --steps 2000 --slots 8 --out weights/ce_refiner_4x4.f32

Every step synthesises a fresh batch (per-slot SNR, delay spread, Doppler
and Rx correlation drawn from TrainSpec), runs the forward and backward
kernels and an Adam update.  Before and after training the NMSE of the LS
interpolation, the refiner and the LMMSE estimator are printed on a held-out
batch at the workload's channel (300 ns, 100 Hz, Rx correlation 0.5).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-4)
    ap.add_argument("--snr-min", type=float, default=20.0)
    ap.add_argument("--snr-max", type=float, default=46.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--init", type=Path, default=None, help="start from these weights")
    ap.add_argument("--out", type=Path, default=None)
    args = ap.parse_args(argv)

    from . import ops
    from . import profiles as P
    from .refiner import WEIGHTS, Refiner, TrainSpec

    if not torch.cuda.is_available():
        print("train_refiner needs a GPU", file=sys.stderr)
        return 1
    cfg = P.FULL_CARRIER_4RX.config()
    spec = TrainSpec(snr_db=(args.snr_min, args.snr_max))
    init = Refiner.load_params(args.init) if args.init else None
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ref = Refiner(cfg, max_slots=max(args.slots, 4), params=init, seed=args.seed)
        filt = ops.LmmseFilter(cfg, tau_s=300e-9, fd_hz=100.0, scs_hz=30e3, max_rank=384)

        def held_out(snr):
            ref.synth(np.random.default_rng(12345), 4, spec, seed=12345, snr_db=snr, delay_spread=300e-9,
                      doppler=100.0, rx_corr=0.5)
            return ref.evaluate(4, lmmse=filt)

        before = {snr: held_out(snr) for snr in (26.0, 34.0, 42.0)}
        print(f"[train] before: {json.dumps(before)}", flush=True)
        t0 = time.perf_counter()
        hist = ref.train(np.random.default_rng(args.seed), args.steps, args.slots, spec, lr=args.lr,
                         lr_final=args.lr_final, log_every=max(args.steps // 20, 1))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        after = {snr: held_out(snr) for snr in (26.0, 34.0, 42.0)}
        print(f"[train] {args.steps} steps of {args.slots} slots in {dt:.1f} s", flush=True)
        print(f"[train] after: {json.dumps(after)}", flush=True)
        out = args.out or WEIGHTS
        ref.save(out)
        print(f"[train] weights -> {out}", flush=True)
        print(json.dumps({"steps": args.steps, "slots": args.slots, "seconds": round(dt, 2), "loss": hist[-1][1],
                          "before": before, "after": after}))
        filt.close()
        ref.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
