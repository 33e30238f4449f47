"""Per-kernel durations measured with HIP events inside libprt, and the
summary of rocprofv3 counter runs that bench.py reports beside them.

    with KernelTimes() as kt:
        ... calls into libprt ...
    kt.stats  # {kernel name: (launches, total ms)}

    python -m parlllelrtree_amd.profiling counters --run NAME=DIR ... [--calib-bytes B]
        per-kernel averages per dispatch of rocprofv3 --pmc CSV output
        (markdown table on stdout, JSON with --json OUT)
    python -m parlllelrtree_amd.profiling trace KERNEL_TRACE.csv
        per-kernel durations of a rocprofv3 --kernel-trace CSV

source_digest() names the kernel sources a recorded counter figure belongs
to: bench.py reports a recorded figure only while the sources are unchanged.
"""
from __future__ import annotations

import argparse
import csv
import ctypes
import hashlib
import json
import sys
from collections import defaultdict
from pathlib import Path

import numpy as np

from ._lib import lib

CSRC = Path(__file__).resolve().parent / "csrc"


# the sources each recorded counter set depends on
DIGEST_SETS = {
    "build": ("prt_common.hpp", "radix.hip", "bu_build.hip", "td_build.hip"),
    "clip": ("prt_common.hpp", "clip_lane.hpp", "clip_team.hpp", "clip_plain.hip", "clip_batched.hip"),
}


def source_digest(which: str = "build") -> str:
    """sha256 (first 16 hex digits) of the libprt sources a recorded counter
    figure depends on (DIGEST_SETS[which], in that order)."""
    h = hashlib.sha256()
    for name in DIGEST_SETS[which]:
        h.update(name.encode())
        h.update((CSRC / name).read_bytes())
    return h.hexdigest()[:16]


class KernelTimes:
    def __init__(self):
        self.stats: dict[str, tuple[int, float]] = {}

    def __enter__(self):
        lib().prt_profile_read(None, 0, None, None, 0)  # drop stale records
        lib().prt_profile_enable(1)
        return self

    def __exit__(self, *exc):
        lib().prt_profile_enable(0)
        names = ctypes.create_string_buffer(1 << 16)
        cnt = np.zeros(256, np.int64)
        ms = np.zeros(256, np.float64)
        k = lib().prt_profile_read(names, len(names), ctypes.c_void_p(cnt.ctypes.data),
                                   ctypes.c_void_p(ms.ctypes.data), 256)
        if k < 0:
            raise RuntimeError(f"prt_profile_read failed ({k})")
        labels = names.raw.split(b"\0")[:k]
        for i, lab in enumerate(labels):
            self.stats[lab.decode()] = (int(cnt[i]), float(ms[i]))
        return False

    def avg_ms(self, name: str) -> float:
        c, t = self.stats[name]
        return t / c

    def dominant(self) -> str:
        return max(self.stats, key=lambda k: self.stats[k][1])


# ---------------------------------------------------------------- counters
def read_counter_csv(run_dir: str | Path) -> dict:
    """{(kernel, grid): {counter: [value per dispatch]}} from the
    *counter_collection.csv files rocprofv3 --pmc --output-format csv wrote
    under run_dir (one row per dispatch and counter)."""
    out: dict = defaultdict(lambda: defaultdict(list))
    files = sorted(Path(run_dir).rglob("*counter_collection.csv"))
    if not files:
        raise FileNotFoundError(f"no *counter_collection.csv under {run_dir}")
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "").strip()
                grid = int(float(row.get("Grid_Size", 0) or 0))
                out[(name, grid)][row["Counter_Name"].strip()].append(
                    float(row["Counter_Value"]))
    return out


def short_name(kernel: str) -> str:
    """Kernel name without namespace, argument list and return type."""
    k = kernel.replace("prt::(anonymous namespace)::", "").replace("prt::", "")
    k = k.split("(")[0].strip()
    for pre in ("void ", "__global__ "):
        if k.startswith(pre):
            k = k[len(pre):]
    return k


def counters_main(argv: list[str]) -> int:
    ap = argparse.ArgumentParser(prog="profiling counters")
    ap.add_argument("--run", action="append", required=True, metavar="NAME=DIR",
                    help="a rocprofv3 --pmc output directory (one counter per run)")
    ap.add_argument("--calib", action="append", default=[], metavar="COUNTER=DIR",
                    help="the same counter's run over a device copy of --calib-bytes")
    ap.add_argument("--calib-bytes", type=float, default=0.0)
    ap.add_argument("--json", default=None)
    ap.add_argument("--traffic-json", default=None,
                    help="also write, per kernel name, the scaled bytes per dispatch averaged "
                         "over all its dispatches (the form bench.py reads)")
    ap.add_argument("--items", type=int, default=0, help="items per launch, for --traffic-json")
    ap.add_argument("--digest", default="build", choices=sorted(DIGEST_SETS),
                    help="the source set the figures are stamped with")
    args = ap.parse_args(argv)
    per = {}
    for spec in args.run:
        name, d = spec.split("=", 1)
        per[name] = read_counter_csv(d)
    scale = {}
    for spec in args.calib:
        cname, d = spec.split("=", 1)
        data = read_counter_csv(d)
        # the copy kernel: the dispatch with the largest value of the counter
        best = max((max(v.get(cname, [0.0])) for v in data.values()), default=0.0)
        if best > 0 and args.calib_bytes > 0:
            scale[cname] = args.calib_bytes / (best * 1024.0)
    rows = {}
    for run, data in per.items():
        for (kern, grid), cnts in data.items():
            for cname, vals in cnts.items():
                key = (short_name(kern), grid)
                r = rows.setdefault(key, {"dispatches": len(vals)})
                r[cname] = float(np.mean(vals))
    print("| kernel | grid | dispatches | " + " | ".join(
        f"{c} KiB/dispatch" for c in sorted({c for r in rows.values() for c in r
                                               if c != "dispatches"})) + " | scaled bytes |")
    cols = sorted({c for r in rows.values() for c in r if c != "dispatches"})
    print("|" + "---|" * (4 + len(cols)))
    res = {}
    for (k, g), r in sorted(rows.items(), key=lambda kv: -sum(v for c, v in kv[1].items()
                                                               if c != "dispatches")):
        scaled = sum(r.get(c, 0.0) * 1024.0 * scale.get(c, 1.0) for c in cols)
        print(f"| {k} | {g} | {r['dispatches']} | " +
              " | ".join(f"{r.get(c, float('nan')):.1f}" for c in cols) + f" | {scaled:.0f} |")
        res[f"{k}@{g}"] = {"kernel": k, "grid": g, "dispatches": r["dispatches"],
                           "kib_per_dispatch": {c: r.get(c) for c in cols},
                           "scaled_bytes_per_dispatch": scaled}
    if args.json:
        Path(args.json).write_text(json.dumps({"scale": scale, "kernels": res,
                                               "source_digest": source_digest(args.digest)},
                                              indent=1))
    if args.traffic_json:
        per_kernel: dict = {}
        for r in res.values():
            d = per_kernel.setdefault(r["kernel"], [0, 0.0])
            d[0] += r["dispatches"]
            d[1] += r["dispatches"] * r["scaled_bytes_per_dispatch"]
        Path(args.traffic_json).write_text(json.dumps(
            {"source_digest": source_digest(args.digest), "scale": scale,
             **{k: {"traffic_bytes": int(round(t / c)), "dispatches": c,
                    "items_per_launch": args.items} for k, (c, t) in per_kernel.items()}},
            indent=1))
    print(f"\nscale factors: {scale}; source digest ({args.digest}) {source_digest(args.digest)}")
    return 0


def valu_main(argv: list[str]) -> int:
    """FP64 VALU of the batched overlay kernels from two rocprofv3 --pmc runs
    (the FP64 instruction counters; the VALU activity counters), per bucket
    launch and per call (the sum over one call's launches); markdown on
    stdout, JSON with --json (stamped with the clip sources' digest).
    FP64 counters count wave instructions: SQ_INSTS_VALU_FLOPS_FP64 = ADD +
    MUL + 2 FMA + TRANS per wave instruction, so x64 lanes bounds the flops.
    VALUBusy = SQ_ACTIVE_INST_VALU / (CUs x GRBM_GUI_ACTIVE / XCDs);
    VALUUtilization = SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU)."""
    ap = argparse.ArgumentParser(prog="profiling valu")
    ap.add_argument("--fp64", required=True, help="run directory of the FP64 counters")
    ap.add_argument("--active", required=True, help="run directory of the activity counters")
    ap.add_argument("--kernel", default="k_clip_batched")
    ap.add_argument("--calls", type=int, required=True, help="batched calls in each run")
    ap.add_argument("--call-ms", type=float, required=True, help="wall time of one call (ms)")
    ap.add_argument("--pairs", type=int, required=True)
    ap.add_argument("--cus", type=int, default=256)
    ap.add_argument("--xcds", type=int, default=8)
    ap.add_argument("--peak-tflops", type=float, default=78.6)
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    data = {}
    for d in (args.fp64, args.active):
        for key, cnts in read_counter_csv(d).items():
            if short_name(key[0]).startswith(args.kernel):
                rec = data.setdefault((short_name(key[0]), key[1]), {})
                for c, v in cnts.items():
                    rec[c] = (float(np.sum(v)), len(v))
    rows, tot_flops, tot_valu = [], 0.0, 0.0
    for (k, g), r in sorted(data.items()):
        get = lambda c: r[c][0] / args.calls if c in r else 0.0  # noqa: E731
        flops = get("SQ_INSTS_VALU_FLOPS_FP64") * 64
        valu = get("SQ_INSTS_VALU")
        act = get("SQ_ACTIVE_INST_VALU")
        thr = get("SQ_THREAD_CYCLES_VALU")
        grbm = get("GRBM_GUI_ACTIVE") / args.xcds
        busy = 100.0 * act / (args.cus * grbm) if grbm else 0.0
        util = 100.0 * thr / (64 * act) if act else 0.0
        fp64_share = 100.0 * get("SQ_INSTS_VALU_FLOPS_FP64") / valu if valu else 0.0
        rows.append((k, g, flops, valu, fp64_share, busy, util))
        tot_flops += flops
        tot_valu += valu
    print("| kernel (bucket launch) | grid | FP64 flops per call (64 lanes x wave instr) | VALU wave-instr per call | FP64 share | VALUBusy % | VALUUtilization % |")
    print("|---|---|---|---|---|---|---|")
    for k, g, fl, va, sh, bu, ut in rows:
        print(f"| {k} | {g} | {fl:.3e} | {va:.3e} | {sh:.1f} % | {bu:.1f} | {ut:.1f} |")
    tflops = tot_flops / (args.call_ms * 1e-3) / 1e12
    print(f"\nper call ({args.call_ms} ms, {args.pairs} pairs): {tot_flops:.3e} FP64 flops at most "
          f"(every lane of every FP64 wave instruction) = {tflops:.3f} TFLOP/s, "
          f"{100 * tflops / args.peak_tflops:.2f} % of {args.peak_tflops} TFLOP/s")
    if args.json:
        Path(args.json).write_text(json.dumps({
            "source_digest": source_digest("clip"),
            "grid_batched": {
                "pairs": args.pairs, "call_ms": args.call_ms,
                "fp64_flops_upper_per_call": tot_flops, "valu_wave_instr_per_call": tot_valu,
                "fp64_tflops_upper": round(tflops, 4),
                "frac_of_peak_upper": round(tflops / args.peak_tflops, 5),
                "launches": [{"kernel": k, "grid": g, "fp64_share_pct": round(sh, 2),
                              "valu_busy_pct": round(bu, 1), "valu_utilization_pct": round(ut, 1)}
                             for k, g, _, _, sh, bu, ut in rows]}}, indent=1))
    return 0


def trace_main(argv: list[str]) -> int:
    """Per (kernel, grid): dispatches, average / median / total duration of a
    rocprofv3 --kernel-trace CSV (markdown on stdout)."""
    ap = argparse.ArgumentParser(prog="profiling trace")
    ap.add_argument("trace_csv")
    args = ap.parse_args(argv)
    groups: dict = defaultdict(list)
    with open(args.trace_csv, newline="") as fh:
        for row in csv.DictReader(fh):
            k = short_name(row["Kernel_Name"])
            g = int(row.get("Grid_Size_X") or row.get("Grid_Size") or 0)
            blk = int(row.get("Workgroup_Size_X") or row.get("Workgroup_Size") or 0)
            groups[(k, g, blk)].append((int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
                                       / 1000.0)
    print("| kernel | grid | block | dispatches | avg us | median us | total ms |")
    print("|---|---|---|---|---|---|---|")
    for (k, g, b), v in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        print(f"| {k} | {g} | {b} | {len(v)} | {np.mean(v):.2f} | {np.median(v):.2f} | "
              f"{sum(v) / 1e3:.2f} |")
    return 0


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "counters":
        sys.exit(counters_main(sys.argv[2:]))
    if len(sys.argv) > 1 and sys.argv[1] == "trace":
        sys.exit(trace_main(sys.argv[2:]))
    if len(sys.argv) > 1 and sys.argv[1] == "valu":
        sys.exit(valu_main(sys.argv[2:]))
    print(__doc__)
