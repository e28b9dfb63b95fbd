"""Train the learned CE refiner on device-synthesised slots and write its
weights (fp32, AIRAN_REFINER_PARAMS values, plus a JSON sidecar with the
input mode) for the receive chain.  By default the network learns a residual
correction of the LMMSE estimate (the filter the chain uses: 300 ns, 100 Hz,
rank 384).

    python -m puschrx.train_refiner --steps 2000 --slots 8 --out weights/ce_refiner_4x4.f32

Every step synthesises a fresh batch (per-slot SNR, delay spread, Doppler
and Rx correlation drawn from TrainSpec), runs the forward and backward
kernels and an Adam update.  Before and after training the NMSE of the LS
interpolation, the refiner and the LMMSE estimator are printed on a held-out
batch at the workload's channel (300 ns, 100 Hz, Rx correlation 0.5).
The weights with the best held-out NMSE at 34 dB are kept.
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
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--lr-final", type=float, default=3e-5)
    ap.add_argument("--chunk", type=int, default=500, help="steps between held-out evaluations")
    ap.add_argument("--snr-min", type=float, default=30.0)
    ap.add_argument("--snr-max", type=float, default=46.0)
    ap.add_argument("--delay-spread", type=float, nargs=2, default=(200e-9, 400e-9), metavar=("MIN", "MAX"))
    ap.add_argument("--doppler", type=float, nargs=2, default=(50.0, 200.0), metavar=("MIN", "MAX"))
    ap.add_argument("--rx-corr", type=float, nargs=2, default=(0.3, 0.7), metavar=("MIN", "MAX"))
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--init", type=Path, default=None, help="start from these weights")
    ap.add_argument("--input", choices=("lmmse", "ls"), default="lmmse",
                    help="network input: a residual correction of the LMMSE estimate, or the LS interpolation")
    ap.add_argument("--out", type=Path, default=None)
    args = ap.parse_args(argv)

    from . import ops
    from . import profiles as P
    from .refiner import WEIGHTS, Refiner, TrainSpec

    if not torch.cuda.is_available():
        print("train_refiner needs a GPU", file=sys.stderr)
        return 1
    cfg = P.FULL_CARRIER_4RX.config()
    spec = TrainSpec(snr_db=(args.snr_min, args.snr_max), delay_spread_s=tuple(args.delay_spread),
                     doppler_hz=tuple(args.doppler), rx_corr=tuple(args.rx_corr))
    init = Refiner.load_params(args.init) if args.init else None
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        filt = ops.LmmseFilter(cfg, tau_s=300e-9, fd_hz=100.0, scs_hz=30e3, max_rank=384)
        mode = 1 if args.input == "lmmse" else 0
        ref = Refiner(cfg, max_slots=max(args.slots, 4), params=init, seed=args.seed, input_mode=mode,
                      estimator=filt if mode else None)

        def held_out(snr):
            ref.synth(np.random.default_rng(12345), 4, spec, seed=12345, snr_db=snr, delay_spread=300e-9,
                      doppler=100.0, rx_corr=0.5)
            return ref.evaluate(4, lmmse=filt)

        before = {snr: held_out(snr) for snr in (26.0, 34.0, 42.0)}
        print(f"[train] before: {json.dumps(before)}", flush=True)
        # Exponential learning-rate decay in chunks; the weights with the best
        # held-out NMSE at 34 dB are kept (Adam with ReLUs can collapse into
        # the identity when a step is too large).
        rng = np.random.default_rng(args.seed)
        best, best_nmse, hist = ref.params(), held_out(34.0)["refiner"], []
        t0 = time.perf_counter()
        done = 0
        while done < args.steps:
            n = min(args.chunk, args.steps - done)
            lr0 = args.lr * (args.lr_final / args.lr) ** (done / max(args.steps - 1, 1))
            lr1 = args.lr * (args.lr_final / args.lr) ** ((done + n - 1) / max(args.steps - 1, 1))
            hist += ref.train(rng, n, args.slots, spec, lr=lr0, lr_final=lr1, log_every=n, seed0=1000 + done)
            done += n
            cur = held_out(34.0)["refiner"]
            print(f"[train] {done} steps: held-out NMSE at 34 dB {cur:.2f} dB (best {best_nmse:.2f})", flush=True)
            if cur < best_nmse:
                best, best_nmse = ref.params(), cur
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ref.set_params(best)
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
