"""Summarise rocprofv3 CSV output of the receive chain per kernel.

    python -m puschrx.rocprof_summary --trace DIR/kt_kernel_trace.csv \
        [--fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \
         --fetch-scale 2.0 --write-scale 1.0] [--min-grid N]
    python -m puschrx.rocprof_summary --sq P1/p1_counter_collection.csv P2/p2_counter_collection.csv \
        [--min-grid N]

Per kernel name and launch grid: dispatch count and mean / min / max
duration from the kernel trace; with counter files, FETCH_SIZE and
WRITE_SIZE per dispatch (KiB in the CSV) multiplied by the scale factors
measured on a device-to-device copy of known size (the counters' ratio to
the bytes a plain copy moves), as MB per dispatch.  --min-grid keeps only
launches with at least that many work-items (drops the one-slot warm-up
launches).  Prints a Markdown table.

--sq takes counter collections of SQ counters (any split over passes) and
prints per kernel and grid, as means over the dispatches: waves, VALU / LDS /
SALU / MFMA / vector-memory-read instructions per wave, and the ratios
SQ_ACTIVE_INST_VALU, SQ_WAIT_INST_ANY and SQ_WAIT_INST_LDS over
SQ_WAVE_CYCLES, SQ_LDS_BANK_CONFLICT over SQ_ACTIVE_INST_LDS, and the SIMD
VALU occupancy: SQ_INSTS_VALU x 4 cycles (a wave64 VALU instruction on a
16-lane SIMD) over the kernel's wave-cycles per SIMD, i.e.
SQ_INSTS_VALU / SQ_WAVE_CYCLES x (waves per SIMD), with waves per SIMD =
SQ_WAVES / 1024 (256 CUs x 4 SIMDs), for the persistent kernels (all waves
resident for the whole launch; --persistent, a regex on the name).
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys


def _short(name: str) -> str:
    n = re.sub(r"\(anonymous namespace\)::", "", name)
    m = re.match(r"_ZN12_GLOBAL__N_1\d+(\w+?)E", n)
    if m:
        n = m.group(1)
        n = re.sub(r"^\d+", "", n)
    n = n.split("(")[0]
    return n.replace("void ", "")[:48]


def _grid(r: dict) -> int:
    if "Grid_Size_X" in r:
        return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
    return int(r["Grid_Size"])


def sq_table(paths, min_grid: int, persistent: str) -> None:
    per = collections.defaultdict(lambda: collections.defaultdict(float))  # (kernel, grid) -> counter -> sum
    disp = collections.defaultdict(lambda: collections.defaultdict(set))  # (kernel, grid) -> counter -> dispatches
    for path in paths:
        for r in csv.DictReader(open(path)):
            g = _grid(r)
            if g < min_grid:
                continue
            key = (_short(r["Kernel_Name"]), g)
            per[key][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[key][r["Counter_Name"]].add(r["Dispatch_Id"])
    print("| kernel | grid | waves | VALU / wave | LDS / wave | SALU / wave | MFMA / wave | VMEM rd / wave | "
          "ACTIVE_INST_VALU / WAVE_CYCLES | WAIT_INST_ANY / WAVE_CYCLES | WAIT_INST_LDS / WAVE_CYCLES | "
          "LDS_BANK_CONFLICT / ACTIVE_INST_LDS |")
    print("|---|---|---|---|---|---|---|---|---|---|---|---|")

    def mean(key, c):
        n = len(disp[key][c])
        return per[key][c] / n if n else 0.0

    for key in sorted(per, key=lambda k: -mean(k, "SQ_WAVE_CYCLES")):
        w = mean(key, "SQ_WAVES") or 1.0
        wc = mean(key, "SQ_WAVE_CYCLES") or 1.0
        lds = mean(key, "SQ_ACTIVE_INST_LDS") or 1.0
        print(f"| {key[0]} | {key[1]} | {w:.0f} | {mean(key, 'SQ_INSTS_VALU') / w:.0f} | "
              f"{mean(key, 'SQ_INSTS_LDS') / w:.0f} | {mean(key, 'SQ_INSTS_SALU') / w:.0f} | "
              f"{mean(key, 'SQ_INSTS_MFMA') / w:.0f} | {mean(key, 'SQ_INSTS_VMEM_RD') / w:.0f} | "
              f"{mean(key, 'SQ_ACTIVE_INST_VALU') / wc:.3f} | {mean(key, 'SQ_WAIT_INST_ANY') / wc:.3f} | "
              f"{mean(key, 'SQ_WAIT_INST_LDS') / wc:.3f} | {mean(key, 'SQ_LDS_BANK_CONFLICT') / lds:.2f} |")
    print()
    for key in sorted(per, key=lambda k: -mean(k, "SQ_WAVE_CYCLES")):
        w = mean(key, "SQ_WAVES")
        if re.search(persistent, key[0]) and w >= 1024:
            occ = mean(key, "SQ_INSTS_VALU") / (mean(key, "SQ_WAVE_CYCLES") or 1.0) * (w / 1024.0)
            print(f"SIMD VALU occupancy {key[0]} (grid {key[1]}, {w / 1024.0:.1f} waves per SIMD): {occ:.2f}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--sq", nargs="+", help="SQ counter collections (instead of --trace)")
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--fetch-scale", type=float, default=1.0)
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--min-grid", type=int, default=0)
    ap.add_argument("--match", default="", help="regex on the kernel name")
    ap.add_argument("--persistent", default="LdpcPair", help="--sq: regex of the persistent kernels")
    args = ap.parse_args(argv)
    if args.sq:
        sq_table(args.sq, args.min_grid, args.persistent)
        return 0
    if not args.trace:
        ap.error("--trace or --sq is required")

    dur = collections.defaultdict(list)
    for r in csv.DictReader(open(args.trace)):
        g = _grid(r)
        if g < args.min_grid or (args.match and not re.search(args.match, r["Kernel_Name"])):
            continue
        dur[(_short(r["Kernel_Name"]), g)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)

    def counters(path, scale):
        acc = collections.defaultdict(list)
        if not path:
            return acc
        for r in csv.DictReader(open(path)):
            g = _grid(r)
            if g < args.min_grid:
                continue
            acc[(_short(r["Kernel_Name"]), g)].append(float(r["Counter_Value"]) * 1024.0 * scale / 1e6)
        return acc

    fetch = counters(args.fetch, args.fetch_scale)
    write = counters(args.write, args.write_scale)
    print("| kernel | grid | calls | mean us | min us | max us | fetch MB | write MB | achieved GB/s |")
    print("|---|---|---|---|---|---|---|---|---|")
    total = 0.0
    for (name, g), v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
        mean = sum(v) / len(v)
        total += sum(v)
        f = fetch.get((name, g))
        w = write.get((name, g))
        fm = sum(f) / len(f) if f else None
        wm = sum(w) / len(w) if w else None
        bw = (fm + wm) / mean * 1e3 if fm is not None and wm is not None else None
        print(f"| {name} | {g} | {len(v)} | {mean:.1f} | {min(v):.1f} | {max(v):.1f} | "
              f"{'' if fm is None else f'{fm:.1f}'} | {'' if wm is None else f'{wm:.1f}'} | "
              f"{'' if bw is None else f'{bw:.0f}'} |")
    print(f"\nTotal kernel time of the rows: {total / 1e3:.3f} ms")
    return 0


if __name__ == "__main__":
    sys.exit(main())
