"""Summarise rocprofv3 CSV output of the receive chain per kernel.

    python -m puschrx.rocprof_summary --trace DIR/kt_kernel_trace.csv \
        [--fetch DIR/fetch_counter_collection.csv --write DIR/write_counter_collection.csv \
         --fetch-scale 2.0 --write-scale 1.0] [--min-grid N]

Per kernel name and launch grid: dispatch count and mean / min / max
duration from the kernel trace; with counter files, FETCH_SIZE and
WRITE_SIZE per dispatch (KiB in the CSV) multiplied by the scale factors
measured on a device-to-device copy of known size (the counters' ratio to
the bytes a plain copy moves), as MB per dispatch.  --min-grid keeps only
launches with at least that many work-items (drops the one-slot warm-up
launches).  Prints a Markdown table.
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--fetch-scale", type=float, default=1.0)
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--min-grid", type=int, default=0)
    ap.add_argument("--match", default="", help="regex on the kernel name")
    args = ap.parse_args(argv)

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
