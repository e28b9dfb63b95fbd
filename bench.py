#!/usr/bin/env python3
"""Decoded Gbit/s of the batched PUSCH receive chain on one MI355X.

Workload (4 Rx, rank 4, 273 PRB): FFT 4096 at 122.88 Msps, 14 symbols, DMRS
type 1 on symbols {2, 11} with ports 1000-1003, 256QAM at R = 873/1024 (TBS
1,081,512 bits, C = 129 code blocks of Z = 384, 40 filler bits), one TB per
slot.  Synthetic slots from the library's transmitter through a 16-tap TDL
channel (300 ns RMS, exponential PDP, Jakes 100 Hz, Kronecker Rx
correlation 0.5) with complex64 IQ; a step is one pass of the chain over a
batch of slots (IQ already in HBM -> decoded bits and CRC outcomes in HBM),
replayed from a HIP graph.  As a receiver pipelines slots, two batches are in
flight by default (--inflight): two chain instances, each with its own
workspace and stream, take the steps in turn, so that one batch's decoder
tail overlaps the next batch's front end.  The first pipelined batch must
decode to the same bits as the single-stream run, whose time is reported
beside the throughput.

The operating SNR is the lowest point of a 20..46 dB sweep (step 2) at which
the chain's BLER is <= 10 %, measured on 64 other slots per point (4 Rx
carrying 4 layers leaves the LMMSE detector no receive diversity, so plain
LMMSE needs an SNR above the 20..32 dB range quoted for the workload; the
SNR is per receive antenna and layer).  Reported beside the throughput:
BLER and goodput at that SNR, the per-stage times (HIP events on the
chain's stream), the BF16 / FP16 precision variants of the chain on the same
slots, the learned estimator's sweep and throughput (the same batches in
flight, each chain with its own refiner), the roofline of the dominant kernel
and the FP32 C++ CPU chain on the host cores (single-thread and all
threads, p50 of 20 runs).
"""
from __future__ import annotations

import argparse
import dataclasses
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "ai-ran-npu-kernels_amd"
for p in (str(PKG), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "decoded Gbit/s, 4Rx rank-4 273-PRB PUSCH at BLER<=10%; learned vs plain LMMSE"
HBM_PEAK_GBS = 8000.0
# HBM bytes per 64-slot launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by a 256 MiB device copy: x2.0 / x1.0), from
# profiles/chain_kernels_4rx_273prb_64slots.md; reported as roofline.traffic
# when the dominant stage and batch match that profile.
MEASURED_TRAFFIC_64 = {"ldpc_decode": round((176.2 + 74.1) * 1e6), "ce_lmmse": None,
                       "detect_demap": round((205.0 + 161.0) * 1e6)}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(g, plan, S: int, fused_ldpc: bool) -> dict:
    """Algorithmic HBM bytes per launch of each stage for S slots (compulsory
    reads of the stage's inputs and writes of its outputs, in the layouts the
    chain keeps).  With fused_ldpc the decoder reads the codeword LLRs itself
    and the rate-dematch stage moves nothing."""
    nr, nl, qm = g.nr, g.nl, g.qm
    data_re = g.ndata * g.used
    grid16 = S * nr * 14 * g.pad * 4                   # fp16 re+im planes
    iq = S * nr * g.samples * 8                        # complex64
    h_ls = S * nr * nl * g.ndmrs * g.pilot_pad * 4
    h_used = S * nr * nl * 14 * g.used * 4
    cw = S * qm * g.cw_symbols * 2
    lam = S * plan.C * 68 * plan.Z * 2
    bits = S * plan.C * plan.K
    return {
        "ofdm_demod": iq + grid16,
        "dmrs_ls": S * nr * g.ndmrs * g.used * 4 // 2 + h_ls,
        "ce_lmmse": h_ls + h_used,
        "detect_demap": S * nr * data_re * 4 + S * nr * nl * data_re * 4 + cw,
        "rate_dematch": 0 if fused_ldpc else cw + lam,
        "ldpc_decode": (cw if fused_ldpc else lam) + bits,
        "crc_check": bits,
    }


def run_grants(args) -> int:
    """Ragged slots: args.grants grants per slot, each its own cell (PRB
    uniform in 4..273, rank 1..4 on 4 Rx, 16/64/256QAM, code rate 0.3..0.9,
    per-grant SNR 10..30 dB, correlated Rayleigh with 8..16 taps at 50..300
    ns), one batch (one HIP graph) per slot.  value = sum of the slot's TBS /
    the graph replay time of the slot's batch, averaged over the slots; the
    filter builds of new grant shapes (cached for the process) are reported
    apart, they are not on the per-slot path."""
    import numpy as np
    import torch

    from puschrx import abi, autotune, synth
    from puschrx import grants as G

    rng = np.random.default_rng(args.seed + 4)
    opts = dict(iq_format=abi.IQ_COMPLEX64, ldpc_iterations=8, early_stop=True, ce_rank=384, tau_s=300e-9,
                fd_hz=100.0)
    stream = torch.cuda.Stream()
    per_slot = []
    with torch.cuda.stream(stream):
        for slot in range(args.grant_slots):
            grants = G.config4_grants(rng, args.grants, slot_number=slot % 20)
            compiled = G.compile_grants(G.Radio(), grants)
            iqs, payloads = [], []
            for rt, cfg, plan in compiled:
                payload = synth.random_payload(rng, 1, plan.tbs)
                grid = synth.tx_grid(cfg, plan, payload)
                spec = synth.ChannelSpec(taps=int(rng.integers(8, 17)), delay_spread_s=float(rng.uniform(50e-9, 300e-9)),
                                         doppler_hz=10.0, rx_corr=0.5, snr_db=float(rng.uniform(10.0, 30.0)))
                iqs.append(synth.receive(rng, cfg, grid, synth.channel(rng, cfg, spec, 1), spec.snr_db,
                                         iq_format=abi.IQ_COMPLEX64)[0])
                payloads.append(payload[0])
            cache0 = int(abi.load_library().airan_pusch_grants_FilterCacheSize())
            t0 = time.perf_counter()
            batch = G.GrantBatch(compiled, opts)
            create_s = time.perf_counter() - t0
            cache1 = int(abi.load_library().airan_pusch_grants_FilterCacheSize())
            iq_dev = batch.pack_iq(iqs)
            tuned = autotune.tune(batch, grants, iq_dev, args.tune_table or None, reps=args.steps)
            ms = autotune.time_batch(batch, iq_dev, reps=args.steps)
            tb_ok = batch.fetch("tb_ok", torch.uint8, batch.n).cpu().numpy()
            mism = 0
            for i, ((rt, cfg, plan), lo) in enumerate(zip(compiled, batch.layout)):
                if tb_ok[i]:
                    b = batch.fetch("bits", torch.int8, plan.C * plan.K, lo.bits_offset).cpu().numpy()
                    per_cb = plan.Kprime - plan.crc_cb
                    mism += int(not np.array_equal(b.reshape(plan.C, plan.K)[:, :per_cb].reshape(-1)[: plan.tbs],
                                                   payloads[i]))
            tbs_bits = sum(p.tbs for _, _, p in compiled)
            good_bits = sum(p.tbs for (_, _, p), ok in zip(compiled, tb_ok) if ok)
            r = {"ms": ms, "tbs_bits": tbs_bits, "good_bits": good_bits, "bler": 1.0 - float(tb_ok.mean()),
                 "code_blocks": batch.num_cb, "shapes": batch.num_shapes, "new_filters": cache1 - cache0,
                 "create_s": create_s, "tile": tuned["tile"], "tile_candidates_ms": tuned.get("candidates"),
                 "payload_mismatches": mism}
            per_slot.append(r)
            log(f"grant slot {slot}: {args.grants} grants, {batch.num_cb} code blocks, {batch.num_shapes} shapes "
                f"({r['new_filters']} new filters, create {create_s:.2f} s), tile {tuned['tile']} "
                f"{tuned.get('candidates')}: {ms:.3f} ms, {tbs_bits / (ms * 1e-3) / 1e9:.2f} Gbit/s, "
                f"BLER {r['bler']:.3f}, payload mismatches {mism}")
            batch.close()
    ms_mean = float(np.mean([r["ms"] for r in per_slot]))
    bits_mean = float(np.mean([r["tbs_bits"] for r in per_slot]))
    good_mean = float(np.mean([r["good_bits"] for r in per_slot]))
    result = {
        "metric": "decoded Gbit/s, ragged multi-grant slots (4 Rx, ranks 1-4, 16/64/256QAM, PRB 4..273)",
        "value": round(bits_mean / (ms_mean * 1e-3) / 1e9, 4),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 2,
        "ms_per_step": round(ms_mean, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32/fp16 (LMMSE CE and detection fp32, fp16 grids, fp16 LDPC messages)",
        "data": "synthetic (library transmitter + TDL channel per grant, SNR 10..30 dB)",
        "config": {
            "workload": f"{args.grants} grants per slot, each its own cell; one batch (one HIP graph) per slot",
            "slots": args.grant_slots,
            "bler": round(float(np.mean([r["bler"] for r in per_slot])), 4),
            "goodput_gbps": round(good_mean / (ms_mean * 1e-3) / 1e9, 4),
            "per_slot": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for r in per_slot],
        },
    }
    print(json.dumps(result), flush=True)
    return 0


def cpu_baseline(cfg, plan, g, spec, opts, iq, S, args) -> dict:
    """The FP32 C++ chain (oracle/cpu_chain.cpp) on the host cores: the same
    grant, filter factors built once, the same slots.  Returns the
    cpu_baseline record (value: all-threads decoded Gbit/s, p50)."""
    import numpy as np

    from oracle_lib import CpuChain, OrCeParams, host_cpu, load_oracle, or_tb_plan
    info = host_cpu()
    op = or_tb_plan(load_oracle(), plan.tbs, plan.G, g.qm, g.nl)
    t0 = time.perf_counter()
    cc = CpuChain(cfg, op, OrCeParams(spec.delay_spread_s, spec.doppler_hz, 30e3), opts.ldpc_iterations,
                  rank=opts.ce_rank)
    setup_s = time.perf_counter() - t0
    threads_all = cc.max_threads()
    out = {}
    decoded = {}
    for label, th in (("single", 1), ("all", threads_all)):
        times = []
        for r in range(args.cpu_warmup + args.cpu_runs):
            idx = [(r * th + i) % S for i in range(th)]
            t1 = time.perf_counter()
            _, ok = cc.run(iq[idx], th)
            dt = time.perf_counter() - t1
            if r >= args.cpu_warmup:
                times.append(dt)
                for i, s in enumerate(idx):
                    decoded[s] = int(ok[i])
        p50 = float(np.median(times))
        out[label] = {"threads": th, "p50_s": round(p50, 4), "slots_per_s": round(th / p50, 3),
                      "gbps": round(th * plan.tbs / p50 / 1e9, 6)}
        log(f"cpu baseline {label}: {th} threads, p50 {p50:.3f} s per run of {th} slots")
    cc.close()
    bler = 1.0 - sum(decoded.values()) / max(len(decoded), 1)
    return {
        "value": out["all"]["gbps"],
        "unit": "Gbit/s",
        "cores": threads_all,
        "kind": "port",
        "sample": f"{args.cpu_warmup} warm-up + {args.cpu_runs} timed runs of one slot per thread from the same "
                  f"{S} slots, p50; FP32 C++ chain (oracle/cpu_chain.cpp, -O3 -march=x86-64-{cc.isa} -fopenmp)",
        "single_thread_gbps": out["single"]["gbps"],
        "single_thread_slots_per_s": out["single"]["slots_per_s"],
        "all_threads_slots_per_s": out["all"]["slots_per_s"],
        "bler": round(bler, 4),
        "slots_decoded": len(decoded),
        "nproc": info["nproc"],
        "affinity_cpus": info["affinity"],
        "cpu_model": info["model"],
        "isa": f"x86-64-{cc.isa}",
        "setup_s": round(setup_s, 2),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slots", type=int, default=64, help="slots per step")
    ap.add_argument("--snr", type=float, default=None, help="skip the sweep and use this SNR (dB)")
    ap.add_argument("--sweep-slots", type=int, default=64)
    ap.add_argument("--no-variants", action="store_true", help="skip the BF16/FP16 precision variants")
    ap.add_argument("--inflight", type=int, default=2,
                    help="slot batches in flight: chain instances (own workspace, own stream) that take the "
                         "steps in turn, so one batch's decoder tail overlaps the next batch's front end")
    ap.add_argument("--cpu-warmup", type=int, default=3, help="CPU baseline warm-up runs")
    ap.add_argument("--cpu-runs", type=int, default=20, help="CPU baseline timed runs (p50 reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--refiner", default=str(PKG / "weights" / "ce_refiner_4x4.f32"),
                    help="learned CE refiner weights ('' to skip the learned variant)")
    ap.add_argument("--refiner-snr-min", type=float, default=20.0)
    ap.add_argument("--refiner-snr-max", type=float, default=46.0)
    ap.add_argument("--grants", type=int, default=0, metavar="N",
                    help="instead: ragged batches of N grants per slot (mixed PRB / rank / Qm / rate, each its own "
                         "cell), --grant-slots slots; prints that workload's line and exits")
    ap.add_argument("--grant-slots", type=int, default=8)
    ap.add_argument("--tune-table", default="", help="autotuner table (JSON) for --grants")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the chain shards by slot with no exchange, and runs on more than one "
              f"GPU are not part of this build (asked for --gpus {args.gpus}).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from puschrx import abi, ops, synth
    from puschrx import profiles as P
    from puschrx.chain import ChainOptions, NativeChain

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    if args.copy_probe:
        n = args.copy_probe << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 5 copies of {n} bytes")
        return 0
    if args.grants:
        return run_grants(args)
    prof = P.FULL_CARRIER_4RX
    cfg = prof.config()
    g = ops.Geometry(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    plan = ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)
    spec = synth.ChannelSpec(taps=16, delay_spread_s=300e-9, doppler_hz=100.0, rx_corr=0.5)
    opts = ChainOptions(ldpc_iterations=8, early_stop=True, tau_s=spec.delay_spread_s, fd_hz=spec.doppler_hz,
                        ce_rank=384, iq_format=abi.IQ_COMPLEX64)
    rng = np.random.default_rng(args.seed)
    S = args.slots

    def make_slots(n, snr, payload=None, grid=None, H=None):
        if payload is None:
            payload = synth.random_payload(rng, n, plan.tbs)
            grid = synth.tx_grid(cfg, plan, payload)
            H = synth.channel(rng, cfg, spec, n)
        iq = synth.receive(rng, cfg, grid, H, snr, iq_format=abi.IQ_COMPLEX64)
        return payload, grid, H, iq

    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        chain = NativeChain(cfg, plan, max_slots=max(S, args.sweep_slots), opts=opts, use_graph=True)
        fused_ldpc = not chain.outputs().ldpc_llr
        log(f"chain ready, {chain.device_bytes() / 2**20:.0f} MiB on the device, "
            f"{'fused rate dematch + LDPC' if fused_ldpc else 'separate rate dematch'}")

        def run(iq_np, n):
            dev = torch.from_numpy(iq_np).cuda()
            chain.enqueue(dev, n)
            ok = chain.fetch("tb_ok", torch.uint8, n)
            torch.cuda.synchronize()
            return ok.cpu().numpy(), dev

        sweep_rng = np.random.default_rng(args.seed + 1)
        sweep_payload = synth.random_payload(sweep_rng, args.sweep_slots, plan.tbs)
        sweep_grid = synth.tx_grid(cfg, plan, sweep_payload)
        sweep_H = synth.channel(sweep_rng, cfg, spec, args.sweep_slots)

        def operating_point(label):
            """Lowest SNR of the sweep with BLER <= 10 % (same slots, same noise
            seeds for every estimator)."""
            sw = {}
            noise_rng = np.random.default_rng(args.seed + 2)
            for snr in range(20, 48, 2):
                iq = synth.receive(noise_rng, cfg, sweep_grid, sweep_H, float(snr), iq_format=abi.IQ_COMPLEX64)
                ok, _ = run(iq, args.sweep_slots)
                sw[snr] = 1.0 - float(ok.mean())
                log(f"{label} sweep snr {snr} dB: BLER {sw[snr]:.3f}")
                if sw[snr] <= 0.10:
                    return float(snr), sw
            return None, sw

        def measure(snr_op, label):
            """Throughput over `steps` graph replays of S slots at snr_op."""
            payload, grid, H, iq = make_slots(S, snr_op)
            iq_dev = torch.from_numpy(iq).cuda()
            for _ in range(args.warmup):
                chain.enqueue(iq_dev, S)
            torch.cuda.synchronize()
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            start.record(stream)
            for _ in range(args.steps):
                chain.enqueue(iq_dev, S)
            end.record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            ms = start.elapsed_time(end) / args.steps
            tb_ok = chain.fetch("tb_ok", torch.uint8, S)
            bits = chain.fetch("bits", torch.int8, S * plan.C * plan.K)
            iters = chain.fetch("iterations", torch.uint8, S * plan.C)
            torch.cuda.synchronize()
            tb_ok = tb_ok.cpu().numpy()
            b = bits.cpu().numpy().reshape(S, plan.C, plan.K).astype(np.uint8)
            per_cb = plan.Kprime - plan.crc_cb
            mism = sum(int(not np.array_equal(b[s, :, :per_cb].reshape(-1)[: plan.tbs], payload[s]))
                       for s in range(S))
            r = {"snr_db": snr_op, "ms_per_step": ms, "tb_ok": tb_ok, "mean_iters": float(iters.float().mean()),
                 "wall": wall, "mismatches": mism, "iq_dev": iq_dev, "iq": iq, "bits": b, "payload": payload}
            log(f"{label}: {args.steps} steps at {snr_op} dB: {ms:.3f} ms/step, wall {wall:.3f} s, "
                f"tb_ok {int(tb_ok.sum())}/{S}, payload mismatches {mism}, mean LDPC iterations "
                f"{r['mean_iters']:.2f}")
            return r

        # 1. Plain LMMSE chain: operating point, throughput, stage times.
        if args.snr is None:
            snr_op, sweep = operating_point("lmmse")
            if snr_op is None:
                snr_op = 46.0
        else:
            snr_op, sweep = args.snr, {}
        plain = measure(snr_op, "lmmse")
        ms_per_step, tb_ok, mean_iters = plain["ms_per_step"], plain["tb_ok"], plain["mean_iters"]
        iq_dev, iq = plain["iq_dev"], plain["iq"]

        def pipelined(first, label, refiners=None):
            """Throughput with `inflight` batches in flight: chain j (its own
            workspace and stream; with refiners[j] attached for the learned
            estimator) takes steps j, j + n, ...; batch 0 is `first`'s input
            and must decode to the same bits."""
            n_in = args.inflight
            streams = [stream] + [torch.cuda.Stream() for _ in range(n_in - 1)]
            chains = [chain]
            for j in range(1, n_in):
                with torch.cuda.stream(streams[j]):
                    cj = NativeChain(cfg, plan, max_slots=S, opts=opts, use_graph=True)
                    if refiners is not None:
                        cj.set_refiner(refiners[j], args.refiner_snr_min, args.refiner_snr_max)
                    chains.append(cj)
            batches = [(first["payload"], first["iq_dev"])]
            for j in range(1, n_in):
                pl, _, _, iqj = make_slots(S, first["snr_db"])
                batches.append((pl, torch.from_numpy(iqj).cuda()))
            torch.cuda.synchronize()
            for _ in range(max(args.warmup, 1)):
                for j in range(n_in):
                    with torch.cuda.stream(streams[j]):
                        chains[j].enqueue(batches[j][1], S)
            torch.cuda.synchronize()
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record(stream)
            for j in range(1, n_in):
                streams[j].wait_event(start)
            for k in range(args.steps):
                j = k % n_in
                with torch.cuda.stream(streams[j]):
                    chains[j].enqueue(batches[j][1], S)
            for j in range(1, n_in):
                done = torch.cuda.Event()
                done.record(streams[j])
                stream.wait_event(done)
            end.record(stream)
            torch.cuda.synchronize()
            pipe_ms = start.elapsed_time(end) / args.steps
            per_cb = plan.Kprime - plan.crc_cb
            pipe_ok, pipe_mism = 0, 0
            for j in range(n_in):
                with torch.cuda.stream(streams[j]):
                    okj = chains[j].fetch("tb_ok", torch.uint8, S)
                    bj = chains[j].fetch("bits", torch.int8, S * plan.C * plan.K)
                torch.cuda.synchronize()
                bj = bj.cpu().numpy().reshape(S, plan.C, plan.K).astype(np.uint8)
                if j == 0 and not np.array_equal(bj, first["bits"]):
                    raise RuntimeError(f"{label}: pipelined batch 0 decoded differently from the single-stream run")
                pipe_ok += int(okj.cpu().numpy().sum())
                pipe_mism += sum(int(not np.array_equal(bj[s, :, :per_cb].reshape(-1)[: plan.tbs],
                                                        batches[j][0][s])) for s in range(S))
            for j in range(1, n_in):
                chains[j].set_refiner(None)
                chains[j].close()
            log(f"{label}, {n_in} batches in flight: {pipe_ms:.3f} ms/step, tb_ok {pipe_ok}/{n_in * S}, "
                f"payload mismatches {pipe_mism}, batch 0 bit-identical to the single-stream run")
            return {"ms_per_step": pipe_ms, "tb_ok": pipe_ok, "slots": n_in * S, "mismatches": pipe_mism}

        pipe = pipelined(plain, "lmmse") if args.inflight > 1 else None

        # Per-stage times (HIP events on the chain's stream, direct launches).
        chain.set_timing(True)
        acc = {k: 0.0 for k in abi.CHAIN_STAGES}
        for _ in range(args.steps):
            chain.enqueue(iq_dev, S)
            for k, v in chain.stage_times().items():
                acc[k] += v
        chain.set_timing(False)
        torch.cuda.synchronize()
        stage_ms = {k: v / args.steps for k, v in acc.items()}
        log("stage ms: " + ", ".join(f"{k} {v:.3f}" for k, v in stage_ms.items()))

        # 1b. The precision variants of the same chain (OFDM DFT and the
        # detector's Gram on the matrix cores) on the measured slots.
        variants = {}
        if not args.no_variants:
            iq_dev_p = plain["iq_dev"]
            for vname, dft, gram in (("fp16", abi.DFT_FP16, abi.GRAM_FP16), ("bf16", abi.DFT_BF16, abi.GRAM_BF16)):
                vopts = dataclasses.replace(opts, dft_precision=dft, gram_precision=gram)
                vc = NativeChain(cfg, plan, max_slots=S, opts=vopts, use_graph=True)
                for _ in range(args.warmup):
                    vc.enqueue(iq_dev_p, S)
                torch.cuda.synchronize()
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ev0.record(stream)
                for _ in range(args.steps):
                    vc.enqueue(iq_dev_p, S)
                ev1.record(stream)
                vok = vc.fetch("tb_ok", torch.uint8, S)
                torch.cuda.synchronize()
                vms = ev0.elapsed_time(ev1) / args.steps
                variants[vname] = {"dft_precision": dft, "gram_precision": gram, "ms_per_step": round(vms, 4),
                                   "gbps": round(S * plan.tbs / (vms * 1e-3) / 1e9, 4),
                                   "bler": round(1.0 - float(vok.float().mean()), 4)}
                log(f"variant {vname}: {vms:.3f} ms/step, BLER {variants[vname]['bler']:.3f}")
                vc.close()

        # 2. The same chain with the learned estimator (LMMSE fallback outside
        # its training SNR range).
        learned = None
        wpath = Path(args.refiner) if args.refiner else None
        if wpath is not None and wpath.exists():
            from puschrx.refiner import Refiner
            ref = Refiner.from_file(cfg, max(S, args.sweep_slots), wpath)
            chain.set_refiner(ref, args.refiner_snr_min, args.refiner_snr_max)
            l_snr, l_sweep = (operating_point("learned") if args.snr is None else (args.snr, {}))
            learned = {"weights": str(wpath.relative_to(ROOT)) if wpath.is_relative_to(ROOT) else str(wpath),
                       "inflight": 1,
                       "snr_range_db": [args.refiner_snr_min, args.refiner_snr_max],
                       "sweep_bler": {str(k): round(v, 4) for k, v in l_sweep.items()}}
            if l_snr is not None:
                m = measure(l_snr, "learned")
                chain.set_timing(True)
                lacc = 0.0
                for _ in range(args.steps):
                    chain.enqueue(m["iq_dev"], S)
                    lacc += chain.stage_times()["ce_lmmse"]
                chain.set_timing(False)
                l_ms = m["ms_per_step"]
                learned.update({
                    "snr_db": l_snr, "bler": round(1.0 - float(m["tb_ok"].mean()), 4),
                    "single_stream": {"ms_per_step": round(l_ms, 4),
                                      "gbps": round(S * plan.tbs / (l_ms * 1e-3) / 1e9, 4)},
                    "ce_stage_ms": round(lacc / args.steps, 4)})
                if args.inflight > 1:
                    # A refiner's activations live in its handle: one per chain in flight.
                    extra = [Refiner.from_file(cfg, S, wpath) for _ in range(args.inflight - 1)]
                    lp = pipelined(m, "learned", refiners=[ref] + extra)
                    for r_ in extra:
                        r_.close()
                    l_ms = lp["ms_per_step"]
                    learned["inflight"] = args.inflight
                learned.update({"value_gbps": round(S * plan.tbs / (l_ms * 1e-3) / 1e9, 4),
                                "ms_per_step": round(l_ms, 4)})
            else:
                learned.update({"snr_db": None, "value_gbps": 0.0})
            chain.set_refiner(None)
            ref.close()
        chain.close()

    tbs_bits = S * plan.tbs
    single_ms = ms_per_step
    if pipe is not None:
        ms_per_step = pipe["ms_per_step"]
    value = tbs_bits / (ms_per_step * 1e-3) / 1e9
    goodput = float(tb_ok.sum()) * plan.tbs / (ms_per_step * 1e-3) / 1e9
    bytes_per = stage_bytes(g, plan, S, fused_ldpc)
    dominant = max(stage_ms, key=stage_ms.get)
    achieved = bytes_per[dominant] / (stage_ms[dominant] * 1e-3) / 1e9
    chain_bytes = S * (g.nr * g.samples * 8 + plan.C * plan.K)
    result = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FFT, CE, detection) / fp16 (LDPC messages) / int16 (LLRs)",
        "data": "synthetic (library transmitter + TDL channel)",
        "config": {
            "workload": "4Rx rank-4 273-PRB PUSCH, FFT 4096, 256QAM R=873/1024, TBS 1081512, C=129, "
                        "TDL-16 300 ns / 100 Hz / rx-corr 0.5, plain LMMSE",
            "slots_per_step": S,
            "snr_db": snr_op,
            "bler": round(1.0 - float(tb_ok.mean()), 4),
            "goodput_gbps": round(goodput, 4),
            "slots_per_s": round(S / (ms_per_step * 1e-3), 1),
            "inflight": args.inflight if pipe is not None else 1,
            "single_stream": {"ms_per_step": round(single_ms, 4),
                              "gbps": round(tbs_bits / (single_ms * 1e-3) / 1e9, 4)},
            "ldpc_max_iterations": opts.ldpc_iterations,
            "ldpc_mean_iterations": round(mean_iters, 3),
            "ce_rank": opts.ce_rank,
            "sweep_bler": {str(k): round(v, 4) for k, v in sweep.items()},
            "estimator": "lmmse",
            "learned": learned,
            "precision_variants": variants,
            "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
            "chain_hbm_gbps": round(chain_bytes / (ms_per_step * 1e-3) / 1e9, 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": (MEASURED_TRAFFIC_64.get(dominant) if S == 64 and fused_ldpc else None),
        },
        "cpu_baseline": None,
    }

    # 4. CPU baseline: the FP32 C++ chain on the same slots, single-thread
    # and on every thread OpenMP is given, 3 warm-up + 20 timed runs each
    # (a run: one slot per thread), p50.
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(cfg, plan, g, spec, opts, iq, S, args)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline skipped: {e}")
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
