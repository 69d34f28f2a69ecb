"""Short text summaries of rocprofv3 CSV output for this library's kernels.

    python -m ptb_amd.rocprof_summary stats DIR [--top N]
        kernel table from DIR/*_kernel_stats.csv (rocprofv3 --kernel-trace --stats)
    python -m ptb_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR --kernel NAME --copy-mib MIB [--first N]
        HBM bytes per dispatch of kernel NAME from two rocprofv3 --pmc runs
        (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch), each scaled by what the
        same counter reads for the plain MIB-MiB device copy in the same run;
        --first N averages only the kernel's first N dispatches (one build's
        launches when the run makes several kinds of builds).

    python -m ptb_amd.rocprof_summary counters DIR [--top N] [--cus 256] [--xcds 8]
        per-kernel per-dispatch averages of every counter in
        DIR/*counter_collection.csv (one rocprofv3 --pmc run), with the MFMA
        figures derived from them where their counters are present:
        MFMA utilisation = SQ_VALU_MFMA_BUSY_CYCLES (SIMD cycles with the
        matrix core busy, summed over the SIMDs) / (GRBM_GUI_ACTIVE / XCDs x
        CUs x 4 SIMDs); GRBM_GUI_ACTIVE reads as the sum over the XCDs' clocks
        (per XCD it matches the kernel's duration), and matrix-core work =
        SQ_INSTS_VALU_MFMA_MOPS_<type> x 512 flop.

Prints markdown on stdout; the traffic mode also prints one line
`bytes_per_launch=<float>` that bench.py reads back from the committed summary.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sys


def short_name(name: str) -> str:
    """Kernel name without namespace, template arguments and parameter list."""
    s = re.sub(r"\(anonymous namespace\)::", "", name)
    s = re.sub(r"^void ", "", s)
    head = s.split("(")[0]
    return head if len(head) < 60 else head[:57] + "..."


def _one(pattern: str) -> str:
    files = sorted(glob.glob(pattern))
    if not files:
        sys.exit(f"no file matches {pattern}")
    return files[0]


def stats(d: str, top: int) -> None:
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*kernel_stats.csv")))))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | calls | total ms | avg us | min us | max us | % |")
    print("|---|---|---|---|---|---|---|")
    for r in rows[:top]:
        print(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} "
              f"| {float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} "
              f"| {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    print(f"\nall kernels: {total / 1e6:.2f} ms in {sum(int(r['Calls']) for r in rows)} dispatches")


def _counter(d: str, counter: str):
    """{dispatch id: (kernel name, value in KiB)} for one counter."""
    out = {}
    for r in csv.DictReader(open(_one(os.path.join(d, "*counter_collection.csv")))):
        if r["Counter_Name"] != counter:
            continue
        did = int(r["Dispatch_Id"])
        name, val = r["Kernel_Name"], float(r["Counter_Value"])
        if did in out:  # one row per agent/instance: sum them
            val += out[did][1]
        out[did] = (name, val)
    return out


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_mib: int, first: int = 0) -> None:
    expect_kib = copy_mib * 1024.0
    scales = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _counter(d, counter)
        copies = [v for n, v in rows.values() if "copyBuffer" in n and v > 0.25 * expect_kib]
        if not copies:
            sys.exit(f"{counter}: no calibration copy of {copy_mib} MiB in {d}")
        calib = max(copies)
        mine = [v for _, (n, v) in sorted(rows.items()) if short_name(n) == kernel]
        if first > 0:
            mine = mine[:first]
        if not mine:
            sys.exit(f"{counter}: kernel {kernel} not found in {d}")
        scales[counter] = (expect_kib / calib, calib, sum(mine) / len(mine), len(mine))
    print(f"| counter | copy reads (KiB) | scale | {kernel} KiB/dispatch (raw) | scaled GB/dispatch | dispatches |")
    print("|---|---|---|---|---|---|")
    total = 0.0
    for counter, (scale, calib, avg, cnt) in scales.items():
        b = avg * scale * 1024.0
        total += b
        print(f"| {counter} | {calib:.0f} (expect {expect_kib:.0f}) | {scale:.3f} | {avg:.0f} "
              f"| {b / 1e9:.4f} | {cnt} |")
    print(f"\nHBM traffic per dispatch of {kernel} (fetch + write, scaled): {total / 1e9:.4f} GB\n")
    print(f"bytes_per_launch={total:.6e}")


def counters(d: str, top: int, cus: int, xcds: int) -> None:
    per = {}  # kernel -> counter -> {dispatch: value}
    for r in csv.DictReader(open(_one(os.path.join(d, "*counter_collection.csv")))):
        k = short_name(r["Kernel_Name"])
        c = per.setdefault(k, {}).setdefault(r["Counter_Name"], {})
        did = int(r["Dispatch_Id"])
        c[did] = c.get(did, 0.0) + float(r["Counter_Value"])
    names = sorted({c for k in per.values() for c in k})
    order = sorted(per, key=lambda k: -sum(per[k].get("GRBM_GUI_ACTIVE", {0: 0}).values()))
    print("| kernel | dispatches | " + " | ".join(f"{n} / dispatch" for n in names) +
          " | MFMA util | MFMA Tflop (F64, BF16) / dispatch |")
    print("|---|---|" + "---|" * len(names) + "---|---|")
    for k in order[:top]:
        cs = per[k]
        nd = max(len(v) for v in cs.values())
        avg = {n: (sum(cs[n].values()) / len(cs[n]) if n in cs else float("nan")) for n in names}
        util = "—"
        if "SQ_VALU_MFMA_BUSY_CYCLES" in cs and "GRBM_GUI_ACTIVE" in cs and avg["GRBM_GUI_ACTIVE"] > 0:
            util = f"{avg['SQ_VALU_MFMA_BUSY_CYCLES'] / (avg['GRBM_GUI_ACTIVE'] / xcds * cus * 4):.3f}"
        fl = []
        for t in ("F64", "BF16"):
            n = f"SQ_INSTS_VALU_MFMA_MOPS_{t}"
            fl.append(f"{avg[n] * 512 / 1e12:.4f}" if n in cs else "—")
        print(f"| `{k}` | {nd} | " + " | ".join(f"{avg[n]:.4g}" for n in names) +
              f" | {util} | {', '.join(fl)} |")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--first", type=int, default=0)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--top", type=int, default=20)
    c.add_argument("--cus", type=int, default=256)
    c.add_argument("--xcds", type=int, default=8)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        stats(a.dir, a.top)
    elif a.mode == "counters":
        counters(a.dir, a.top, a.cus, a.xcds)
    else:
        traffic(a.fetch_dir, a.write_dir, a.kernel, a.copy_mib, a.first)


if __name__ == "__main__":
    main()
