"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m hastycompute_amd.profiling stats  <kernel_stats.csv>
    python -m hastycompute_amd.profiling traffic --fetch <counter_collection.csv> --write <...>
        --copy-fetch <...> --copy-write <...> --copy-bytes <bytes> [--json out.json --key-suffix S]

`traffic` turns FETCH_SIZE / WRITE_SIZE (KiB per dispatch, one counter per run) into
HBM bytes per launch, each scaled by what the same counter reads for a plain device
copy of known size (bytes counted / bytes moved), and prints a Markdown table; the JSON
entries carry the SHA-256 of the kernel's source file, so a reader can tell when a
stored figure no longer describes the code.
"""
from __future__ import annotations

import argparse
import collections
import csv
import hashlib
import json
import os
import re
import sys

# kernel symbol -> name used by the library's event timing
KERNEL_NAMES = {"fwd_xy_kernel": "sense_fwd_xy", "inv_xy_kernel": "sense_inv_xy", "fwd_xy_w_kernel": "sense_fwd_xy",
                "inv_xy_w_kernel": "sense_inv_xy", "inv_xy_d_kernel": "sense_inv_xy", "fwd_xy_r16_kernel": "sense_fwd_xy",
                "inv_xy_r16_kernel": "sense_inv_xy", "z_conv_r16_kernel": "sense_z_conv",
                "z_conv_r20_kernel": "sense_z_conv",
                "z_conv_ct_kernel": "sense_z_conv", "z_conv_kernel": "sense_z_conv", "spread_kernel": "nufft_spread",
                "spread_rows_kernel": "nufft_spread", "spread_planes_kernel": "nufft_spread",
                "gather_spread_kernel": "nufft_spread", "zero_kernel": "fill_zero",
                "interp_kernel": "nufft_interp", "axis_fft_kernel": "fft_axis", "axis_fft_r16_kernel": "fft_axis",
                "gram_kernel": "llr_gram",
                "eig_w_kernel": "llr_eig", "apply_kernel": "llr_apply"}


def short_name(full: str) -> str:
    m = re.search(r"(\w+)(?:<[^()]*>)?\(", full.replace("void ", ""))
    return m.group(1) if m else full[:60]


def kernel_stats(path: str):
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows:
        out.append({"kernel": short_name(r["Name"]), "full": r["Name"], "calls": int(r["Calls"]),
                    "avg_us": float(r["AverageNs"]) / 1e3, "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "percent": float(r["Percentage"])})
    return out


def counter_per_kernel(path: str):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        agg[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return agg


# kernel symbol -> the source file it is compiled from (for the staleness stamp)
KERNEL_SOURCES = {"fwd_xy_r16_kernel": "sense_fused.hip", "inv_xy_r16_kernel": "sense_fused.hip",
                  "z_conv_r16_kernel": "sense_fused.hip", "fwd_xy_w_kernel": "sense_fused.hip",
                  "inv_xy_w_kernel": "sense_fused.hip", "fwd_xy_kernel": "sense_fused.hip",
                  "inv_xy_kernel": "sense_fused.hip", "z_conv_r20_kernel": "sense_fused.hip",
                  "spread_kernel": "nufft.hip", "interp_kernel": "nufft.hip",
                  "spread_rows_kernel": "nufft.hip", "spread_planes_kernel": "nufft.hip",
                  "gather_spread_kernel": "nufft.hip", "zero_kernel": "nufft.hip",
                  "axis_fft_kernel": "fft.hip", "axis_fft_r16_kernel": "fft.hip", "gram_kernel": "llr.hip",
                  "eig_w_kernel": "llr.hip",
                  "apply_kernel": "llr.hip"}
SRC_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "hip")


def source_sha256(kernel_symbol: str):
    """SHA-256 of the source file a kernel is built from (None if unknown): a stored
    counter figure is only valid for the source it was measured on."""
    f = KERNEL_SOURCES.get(kernel_symbol)
    if f is None:
        return None
    with open(os.path.join(SRC_DIR, f), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def copy_factor(path: str, known_bytes: int) -> float:
    """Counted bytes per moved byte for the largest dispatch of the calibration run other
    than the streaming-ceiling probes (that is the known-size copy)."""
    best = 0.0
    for name, vals in counter_per_kernel(path).items():
        if "stream_probe" in name:
            continue
        best = max(best, max(vals))
    return best * 1024.0 / known_bytes


def cmd_stats(a):
    st = kernel_stats(a.csv)
    print("| kernel | calls | avg (us) | total (ms) | % |")
    print("|---|---:|---:|---:|---:|")
    for s in st[:a.top]:
        print(f"| {s['kernel']} | {s['calls']} | {s['avg_us']:.1f} | {s['total_ms']:.2f} | {s['percent']:.2f} |")


def cmd_traffic(a):
    fr = copy_factor(a.copy_fetch, a.copy_bytes)
    fw = copy_factor(a.copy_write, a.copy_bytes)
    fetch = counter_per_kernel(a.fetch)
    write = counter_per_kernel(a.write)
    names = [k for k in fetch if k in KERNEL_NAMES and k in write]
    print(f"calibration: FETCH_SIZE reads {fr:.4f} and WRITE_SIZE {fw:.4f} of the bytes of a "
          f"{a.copy_bytes} B device copy\n")
    print("| kernel | dispatches | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | HBM read B/launch | "
          "HBM write B/launch | traffic B/launch |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    js = {}
    for k in names:
        f = sum(fetch[k]) / len(fetch[k])
        w = sum(write[k]) / len(write[k])
        rb, wb = f * 1024 / fr, w * 1024 / fw
        print(f"| {k} | {len(fetch[k])} | {f:.1f} | {w:.1f} | {rb:.4g} | {wb:.4g} | {rb + wb:.4g} |")
        ent = {"kernel_symbol": k, "hbm_read_bytes_per_launch": rb, "hbm_write_bytes_per_launch": wb,
               "hbm_bytes_per_launch": rb + wb, "fetch_factor": fr, "write_factor": fw,
               "source_sha256": source_sha256(k)}
        if a.units_per_launch:
            ent["units_per_launch"] = a.units_per_launch
            ent["hbm_bytes_per_unit"] = (rb + wb) / a.units_per_launch
        js[KERNEL_NAMES[k] + a.key_suffix] = ent
    if a.json:
        old = {}
        try:
            old = json.load(open(a.json))
        except (OSError, ValueError):
            pass
        old.update(js)
        json.dump(old, open(a.json, "w"), indent=1, sort_keys=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m hastycompute_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    for n in ("--fetch", "--write", "--copy-fetch", "--copy-write"):
        t.add_argument(n, required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--json")
    t.add_argument("--key-suffix", default="")
    t.add_argument("--units-per-launch", type=float, default=0.0,
                   help="work units (problems, transforms) per launch of the measured run")
    a = ap.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic}[a.cmd](a)


if __name__ == "__main__":
    main(sys.argv[1:])
