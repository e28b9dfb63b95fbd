"""Tile autotuner for ragged multi-grant batches.

The ragged detector's workgroup covers `detect_tile` subcarriers of one
(grant, data symbol) row (64, 128 or 256).  Small grants waste lanes with
large tiles, large grants pay more workgroups with small ones, so the best
tile depends on the batch's PRB mix.  The tuner replays the batch's HIP
graph with every candidate tile and keeps the fastest, caching the choice per
workload signature (the PRB histogram in octave buckets and the rank mix) in
an optional JSON table so later batches of the same mix skip the search.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

TILES = (64, 128, 256)


def signature(grants) -> str:
    """Workload signature: grants per PRB octave bucket and per rank."""
    buckets = [0] * 8
    ranks = [0] * 4
    for g in grants:
        buckets[min(7, int(math.log2(max(g.num_rb, 1))))] += 1
        ranks[g.rank - 1] += 1
    return "prb" + "-".join(map(str, buckets)) + "_rank" + "-".join(map(str, ranks))


def time_batch(batch, iq_dev, reps: int = 10) -> float:
    """Milliseconds per graph replay of the batch."""
    import torch
    for _ in range(2):
        batch.enqueue(iq_dev)
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    st = torch.cuda.current_stream()
    s.record(st)
    for _ in range(reps):
        batch.enqueue(iq_dev)
    e.record(st)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def tune(batch, grants, iq_dev, table_path: str | None = None, reps: int = 10) -> dict:
    """Pick the batch's detector tile; returns {"tile", "ms", "candidates", "cached"}."""
    sig = signature(grants)
    table = {}
    if table_path and Path(table_path).exists():
        table = json.loads(Path(table_path).read_text())
    if sig in table:
        batch.set_detect_tile(int(table[sig]["tile"]))
        return {**table[sig], "cached": True, "signature": sig}
    cand = {}
    for t in TILES:
        batch.set_detect_tile(t)
        cand[t] = time_batch(batch, iq_dev, reps)
    best = min(cand, key=cand.get)
    batch.set_detect_tile(best)
    entry = {"tile": best, "ms": cand[best], "candidates": {str(k): round(v, 4) for k, v in cand.items()}}
    if table_path:
        table[sig] = entry
        Path(table_path).write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    return {**entry, "cached": False, "signature": sig}
