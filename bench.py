#!/usr/bin/env python3
"""bench.py -- R-tree build throughput on one MI355X (and the join / overlay
phases), printed as one JSON line.

Workload (the step): the largest single-GPU configuration, 16,777,216 fp32
rects from a 64-component Gaussian mixture (means U[0,1]^2, sigma 0.01,
sides U[1e-5, 1e-4]), inputs resident in HBM; one step builds BOTH trees of
that configuration over the rects, fanout 16: the bottom-up tree (CENTER30
Morton keys, stable LSD radix sort, level packing) and the top-down tree
(per-level segmented sorts, slot layout).  value = rects entered into a tree
per second = 2 x 16,777,216 / step time (Mrect/s); each build's own rate is
in phases.w4_builds.
roofline: the dominant kernel of the step (k_onesweep, one digit pass of the
radix sort), its average duration from HIP events that hipExtLaunchKernel
stamps with each launch's own start and end over 3 steps, against its
algorithmic bytes per launch (every item read once and written once; the
passes that also move the rects count those bytes).  traffic: the HBM bytes
per launch that rocprofv3 FETCH_SIZE / WRITE_SIZE recorded (profiles/), when
recorded for these exact kernel sources.
cpu_baseline: the oracle's serial bottom-up and top-down builds of the same
16,777,216 rects on one host core.  Every phase carries its own CPU figures
(the oracle's builds, DFS joins and FP64 clips, on 1 core and on all cores
of the host share), and the CPU leg then compares the phases' GPU results
with the oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "R-tree build Mrect/s; overlay pairs/s batched vs plain at ≤1e-12 area err, 1 GPU"
W4_N = 1 << 24


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cores() -> int:
    """Host cores the CPU baselines use: this process's CPU set, at most 16
    (a GPU box's share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def nproc() -> int:
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        return os.cpu_count() or 1


# algorithmic HBM bytes per processed item, per kernel (minimal traffic:
# every input read once, every output written once)
BYTES_PER_ITEM = {
    "k_onesweep": 16,       # one digit pass: 8-B item read + 8-B item written (+ GATHER_BYTES)
    "k_hist_all<Src>": 24,  # first radix level: 16-B rect read, 8-B item written (8-B read after)
    "k_td_gather": 48,      # 8-B item, 16-B rect + 4-B id read; 16 + 4 written
    "k_td_subtree<8>": 52,  # 8-B item, 16-B rect, 4-B id read; 16-B box + 4-B child + 4-B perm written
    "k_td_subtree<4>": 52,  # the same kernel when each block has a CU to itself
    "k_bu_leaves": 52,      # 8-B item, 16-B rect read; 4+4 key/id, 16+4 entry written
}
# the last digit pass of the top-down radix level before the last one also
# moves the rects and ids into the sorted order (16-B rect read, 16 + 4 B
# written per item; the id is the item's payload); the last digit pass of a
# level before that reads each item's rect for the next level's key (16 B)
GATHER_BYTES = 36
NEXT_KEY_BYTES = 16
TD_LDS_SLOTS = 4096  # segments up to this size are sorted in LDS by the subtree kernel
BU_KEY_BITS = 30     # CENTER30 keys: 4 digit passes of at most 8 bits
TD_KEY_BITS = 32     # ord32 coordinates


def td_radix_levels(n: int, f: int) -> int:
    """Levels of a top-down build sorted by the segmented radix sort (their
    segments hold more than TD_LDS_SLOTS rects)."""
    from parlllelrtree_amd import rtree
    L = rtree.td_levels(n, f)
    count = 0
    for p in range(L):
        a = 1
        for _ in range(L - p):
            a *= f
            if a >= n:
                a = n
                break
        count += a > TD_LDS_SLOTS
    return count


def onesweep_mix(n: int, builds: list[str], f: int = 16) -> tuple[int, int]:
    """(digit-pass launches, algorithmic bytes) of k_onesweep for one
    bottom-up ("bu") / top-down ("td") build of n rects each."""
    launches, nbytes = 0, 0
    for b in builds:
        if b == "bu":
            p = (BU_KEY_BITS + 7) // 8
            launches += p
            nbytes += p * 16 * n
        else:
            p = td_radix_levels(n, f) * ((TD_KEY_BITS + 7) // 8)
            launches += p
            nbytes += p * 16 * n + td_extra_bytes(n, f)
    return launches, nbytes


def td_extra_bytes(n: int, f: int = 16) -> int:
    """Bytes the last digit passes of a top-down build's radix levels move
    besides the items: NEXT_KEY_BYTES per item for each level followed by two
    or more radix levels, GATHER_BYTES for the level before the last."""
    lv = td_radix_levels(n, f)
    return (max(0, lv - 2) * NEXT_KEY_BYTES + (1 if lv >= 2 else 0) * GATHER_BYTES) * n


def digit_pass_bytes(stats: dict, runs: int, n: int, builds: list[str], f: int = 16) -> dict:
    """Algorithmic bytes per launch of k_onesweep over `runs` runs of the
    builds (n rects each), the mean over the digit passes: every pass moves
    16 B per item, the last passes of the top-down radix levels more
    (td_extra_bytes)."""
    launches, nbytes = onesweep_mix(n, builds, f)
    c = stats.get("k_onesweep", (0, 0.0))[0] // runs
    assert c == launches, (c, launches)
    return {"k_onesweep": nbytes // launches}


def timed(fn, steps: int, warmup: int) -> float:
    """Average ms per call of fn, bracketed by device syncs and HIP events on
    the current stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def cpu_median(fn, reps: int, warm: bool = True) -> float:
    """Median wall time (s) of fn over reps runs, after one warm-up run
    (when `warm`)."""
    if warm:
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def cpu_rate(fn, units: float, unit: str, reps: int, cores: int, sample: str,
             fn_n=None) -> dict:
    """CPU baseline of one phase: units / median time on 1 core (fn) and on
    `cores` cores (fn_n, when given)."""
    t1 = cpu_median(fn, reps)
    out = {"kind": "port", "unit": unit, "1_core": round(units / t1 / 1e6, 4),
           "sample": sample, "median_of": reps}
    if fn_n is not None and cores > 1:
        tn = cpu_median(fn_n, reps)
        out[f"{cores}_cores"] = round(units / tn / 1e6, 4)
    return out


def recorded(name: str, key: str, which: str = "build"):
    """A recorded counter figure from profiles/<name>, only while the kernel
    sources it depends on are the ones it was recorded for (else None)."""
    from parlllelrtree_amd.profiling import source_digest
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    rec = json.loads(p.read_text())
    if rec.get("source_digest") != source_digest(which):
        return None
    if key in rec:
        return rec[key]
    # the profiler names template instances (k_onesweep<8>); the HIP-event
    # record names the kernel: take its one instance
    inst = [k for k in rec if k.startswith(key + "<")]
    return rec[inst[0]] if len(inst) == 1 else None


# ---------------------------------------------------------------- headline
def headline(dev, args, cpu_jobs):
    """The step: bottom-up + top-down build of the 16,777,216 mixture rects."""
    import torch

    from parlllelrtree_amd import rtree, workloads
    from parlllelrtree_amd.profiling import KernelTimes

    n = W4_N
    r = workloads.w4_rects(n)
    d = torch.from_numpy(r).to(dev)
    tb = rtree.build_bottom_up(d, 16)
    tt = rtree.build_top_down(d, 16)
    torch.cuda.synchronize()

    def step():
        rtree.rebuild(tb, d)
        rtree.rebuild(tt, d)

    log(f"timing bottom-up + top-down build of {n} mixture rects: {args.warmup} warmup + "
        f"{args.steps} steps")
    ms = timed(step, args.steps, args.warmup)
    rtree.check_device_status()
    value = 2 * n / (ms * 1e-3) / 1e6
    log(f"step {ms:.3f} ms -> {value:.1f} Mrect/s (each rect into both trees)")

    with KernelTimes() as kt:
        for _ in range(3):
            step()
        torch.cuda.synchronize()
    dom = max((k for k in kt.stats if k in BYTES_PER_ITEM), key=lambda k: kt.stats[k][1])
    cnt, tot = kt.stats[dom]
    avg_ms = tot / cnt
    passes = digit_pass_bytes(kt.stats, 3, n, ["bu", "td"])
    alg_bytes = passes[dom] if dom in passes else BYTES_PER_ITEM[dom] * n
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    # every digit pass of the step (both kernels): bytes / time
    p_launch, p_bytes = onesweep_mix(n, ["bu", "td"])
    p_ms = sum(kt.stats[k][1] for k in passes) / 3
    digit = {"launches": p_launch, "ms": round(p_ms, 4),
             "achieved": round(p_bytes / (p_ms * 1e-3) / 1e9, 1),
             "frac": round(p_bytes / (p_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "per_kernel": {k: {"launches": kt.stats[k][0] // 3,
                                "avg_us": round(kt.stats[k][1] / kt.stats[k][0] * 1e3, 2),
                                "bytes_per_launch": b,
                                "GBs": round(b / (kt.stats[k][1] / kt.stats[k][0] * 1e-3) / 1e9, 1)}
                            for k, b in passes.items()}}
    traffic = recorded("hbm_traffic.json", dom)
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic["traffic_bytes"] if traffic else None,
                "avg_us": round(avg_ms * 1e3, 2), "bytes_per_launch": alg_bytes,
                "launches_per_step": cnt // 3, "digit_passes": digit,
                "kernels_ms_per_step": {k: round(v[1] / 3, 4) for k, v in
                                        sorted(kt.stats.items(), key=lambda kv: -kv[1][1])}}
    if traffic:
        roofline["traffic_ratio"] = round(traffic["traffic_bytes"] / alg_bytes, 3)
    log(f"dominant kernel {dom}: {avg_ms * 1e3:.1f} us avg over {cnt // 3} launches a step, "
        f"{achieved:.0f} GB/s algorithmic ({achieved / HBM_PEAK_GBS:.3f} of peak)")

    # each build alone
    ms_bu = timed(lambda: rtree.rebuild(tb, d), args.steps, args.warmup)
    ms_td = timed(lambda: rtree.rebuild(tt, d), args.steps, args.warmup)
    with KernelTimes() as kb:
        rtree.rebuild(tb, d)
        torch.cuda.synchronize()
    with KernelTimes() as ktd:
        rtree.rebuild(tt, d)
        torch.cuda.synchronize()
    rtree.check_device_status()
    phase = {"rects": n, "bu_ms": round(ms_bu, 4), "bu_Mrect_s": round(n / ms_bu / 1e3, 1),
             "td_ms": round(ms_td, 4), "td_Mrect_s": round(n / ms_td / 1e3, 1),
             "bu_kernels_ms": {k: round(v[1], 4) for k, v in
                               sorted(kb.stats.items(), key=lambda kv: -kv[1][1])},
             "td_kernels_ms": {k: round(v[1], 4) for k, v in
                               sorted(ktd.stats.items(), key=lambda kv: -kv[1][1])}}
    log(f"w4 builds alone: bu {ms_bu:.3f} ms ({n / ms_bu / 1e3:.0f} Mrect/s), "
        f"td {ms_td:.3f} ms ({n / ms_td / 1e3:.0f} Mrect/s)")
    del tb, tt, d
    torch.cuda.empty_cache()

    cpu = {}

    def cpu_job(O):
        reps = args.cpu_reps_large
        t_bu = cpu_median(lambda: O.bu_build(r, 16), reps, warm=False)
        t_td = cpu_median(lambda: O.td_build(r, 16), reps, warm=False)
        cpu.update({"value": round(2 * n / (t_bu + t_td) / 1e6, 3), "unit": "Mrect/s",
                    "cores": 1, "kind": "port",
                    "sample": f"the whole workload: the oracle's serial bottom-up build "
                              f"(Morton keys, stable sort, packing) {t_bu:.2f} s and serial "
                              f"top-down build {t_td:.2f} s of the 16,777,216 rects, median of "
                              f"{reps} each; host {cpu_model()}, nproc {nproc()}"})
        phase["cpu_baseline"] = {"kind": "port", "unit": "Mrect/s",
                                 "bu_1_core": round(n / t_bu / 1e6, 3),
                                 "td_1_core": round(n / t_td / 1e6, 3), "median_of": reps}
        log(f"cpu baseline: bu {t_bu:.2f} s, td {t_td:.2f} s -> {cpu['value']} Mrect/s")
    cpu_jobs.append(cpu_job)
    line = {"value": value, "ms": ms, "roofline": roofline, "cpu": cpu}
    return line, phase


# ---------------------------------------------------------------- phases
def layers_phase(dev, args, cpu_jobs, checks):
    """Top-down builds of two 1,048,576-rect uniform layers in one launch
    sequence and one layer per call; the layer-layer join; plain vs batched
    overlay of the first 100,000 sorted pairs; their contours."""
    import numpy as np
    import torch

    from parlllelrtree_amd import rtree, workloads
    from parlllelrtree_amd.profiling import KernelTimes

    n = 1 << 20
    ra, rb = workloads.w1(n)
    da, db = torch.from_numpy(ra).to(dev), torch.from_numpy(rb).to(dev)
    TL = rtree.TopDownLayers([da, db], 16)
    ta = TL.trees[0]
    ms = timed(lambda: TL.rebuild([da, db]), args.steps, args.warmup)
    with KernelTimes() as kt:
        for _ in range(3):
            TL.rebuild([da, db])
        torch.cuda.synchronize()
    c, t = kt.stats["k_onesweep"]
    # both layers per launch: a top-down build of 2n items, segmented per layer
    per_step = td_radix_levels(n, 16) * 4
    alg = (per_step * 16 * 2 * n + 2 * td_extra_bytes(n, 16)) // per_step
    t1a, t1b = rtree.build_top_down(da, 16), rtree.build_top_down(db, 16)

    def per_layer():
        rtree.rebuild(t1a, da)
        rtree.rebuild(t1b, db)
    ms1 = timed(per_layer, args.steps, args.warmup)
    del t1a, t1b
    out = {"td_build_2x1M": {
        "ms": round(ms, 4), "Mrect_s": round(2 * n / ms / 1e3, 1),
        "one_layer_per_call_ms": round(ms1, 4),
        "one_layer_per_call_Mrect_s": round(2 * n / ms1 / 1e3, 1),
        "k_onesweep_avg_us": round(t / c * 1e3, 2),
        "k_onesweep_GBs": round(alg / (t / c * 1e-3) / 1e9, 1),
        "note": "working set inside the 256 MB Infinity Cache: latency-bound, not an HBM figure"}}
    log(f"td 2 x 1M layers: {ms:.3f} ms ({2 * n / ms / 1e3:.0f} Mrect/s), one layer per call "
        f"{ms1:.3f} ms")

    def cpu_td(O):
        out["td_build_2x1M"]["cpu_baseline"] = cpu_rate(
            lambda: O.td_build(ra, 16), n, "Mrect/s", args.cpu_reps, 1,
            "the oracle's serial top-down build of one 1,048,576-rect layer")
    cpu_jobs.append(cpu_td)

    # layer-layer join (top-down tree of A, queries = B)
    cap1 = rtree.join(ta, db).shape[0]
    j1 = rtree.Joiner(ta, n, cap1, dev)
    ms_j1 = timed(lambda: j1.run(db), max(args.steps // 2, 3), 1)
    pairs_qr = j1.result().cpu().numpy().view(np.uint32)
    out["join_layers_1M_x_1M"] = {"ms": round(ms_j1, 4), "pairs": cap1,
                                  "Mpairs_s": round(cap1 / ms_j1 / 1e3, 1)}
    log(f"layer join {ms_j1:.3f} ms ({cap1} pairs)")

    def cpu_join(O):
        ot = O.td_build(ra, 16)
        out["join_layers_1M_x_1M"]["cpu_baseline"] = cpu_join_rate(
            O, ot, rb, "td", cap1, args.cpu_reps,
            "the oracle's DFS search (RtreeSearch semantics) of all 1,048,576 B rects "
            "in the top-down tree of A")
    cpu_jobs.append(cpu_join)

    cp, A, B, _, _ = workloads.clip_inputs_from_pairs(pairs_qr, ra, rb, 100_000)
    out["overlay_layers_100k"] = overlay_phase(cp, A, B, dev, args, checks, cpu_jobs)
    out["overlay_layers_100k"].pop("_tune")
    out["contours_layers_100k"] = contours_phase(cp, A, B, dev, checks, cpu_jobs)
    return out


def cpu_join_rate(O, tree, q, kind, pairs, reps, sample):
    """The oracle's DFS join on 1 core and on all cores (queries split in
    contiguous chunks over threads; the C search releases the GIL)."""
    import numpy as np
    fn = O.join_bu if kind == "bu" else O.join_td
    cores = host_cores()
    chunks = np.array_split(np.ascontiguousarray(q), cores)

    def par():
        with ThreadPoolExecutor(cores) as ex:
            return sum(len(p) for p, _ in ex.map(lambda c: fn(tree, c), chunks))
    got = par()
    assert got == pairs, (got, pairs)
    return cpu_rate(lambda: fn(tree, q), pairs, "Mpairs/s", reps, cores, sample, par)


def bu_phase(dev, args, cpu_jobs):
    """Bottom-up build of 1,048,576 uniform rects + 65,536 windows."""
    import torch

    from parlllelrtree_amd import rtree, workloads

    n = 1 << 20
    r0, q0 = workloads.w0()
    d0, dq0 = torch.from_numpy(r0).to(dev), torch.from_numpy(q0).to(dev)
    t0 = rtree.build_bottom_up(d0, 16)
    ms_bu = timed(lambda: rtree.rebuild(t0, d0), args.steps, args.warmup)
    out = {"bu_build_1M": {"ms": round(ms_bu, 4), "Mrect_s": round(n / ms_bu / 1e3, 1)}}
    cap0 = rtree.join(t0, dq0).shape[0]
    j0 = rtree.Joiner(t0, q0.shape[0], cap0, dev)
    ms_j0 = timed(lambda: j0.run(dq0), max(args.steps // 2, 3), 1)
    out["join_windows_65536"] = {"ms": round(ms_j0, 4), "pairs": cap0,
                                 "Mpairs_s": round(cap0 / ms_j0 / 1e3, 1)}
    log(f"bu build {ms_bu:.3f} ms, window join {ms_j0:.3f} ms ({cap0} pairs)")

    def cpu(O):
        out["bu_build_1M"]["cpu_baseline"] = cpu_rate(
            lambda: O.bu_build(r0, 16), n, "Mrect/s", args.cpu_reps, 1,
            "the oracle's serial bottom-up build of the 1,048,576 rects")
        ot = O.bu_build(r0, 16)
        out["join_windows_65536"]["cpu_baseline"] = cpu_join_rate(
            O, ot, q0, "bu", cap0, args.cpu_reps,
            "the oracle's DFS search of the 65,536 windows in the bottom-up tree")
    cpu_jobs.append(cpu)
    return out


def overlay_phase(cp, A, B, dev, args, checks, cpu_jobs, check_sample=None,
                  plain_reps=(2, 1), cpu_sample=20_000):
    """Plain vs batched FP64 overlay of pairs cp over CSR sets A, B.  Appends
    to `checks` the comparison of the areas with the oracle (on
    `check_sample` pairs if given) and the oracle's FP64 flop count, and to
    `cpu_jobs` the oracle's clip rate on `cpu_sample` pairs (1 core, all
    cores)."""
    import numpy as np
    import torch

    from parlllelrtree_amd import rtree
    from parlllelrtree_amd._lib import lib
    from parlllelrtree_amd.autotune import autotune, tune_to_json

    pd = torch.from_numpy(np.ascontiguousarray(cp).view(np.int32)).to(dev)
    Ad, Bd = rtree.to_device_polys(A, dev), rtree.to_device_polys(B, dev)
    mv = Ad.max_verts + Bd.max_verts
    npairs = len(cp)
    out = {"pairs": npairs, "max_verts_pair": mv}
    area_p = torch.empty(npairs, dtype=torch.float64, device=dev)
    ws_p = torch.empty(lib().prt_overlay_plain_workspace_size(npairs, mv), dtype=torch.uint8,
                       device=dev)
    ms_p = timed(lambda: rtree.overlay_plain(pd, Ad, Bd, area_p, ws_p), *plain_reps)
    del ws_p
    tune, tstats = autotune(cp, A, B, Ad, Bd)
    area_b = torch.empty(npairs, dtype=torch.float64, device=dev)
    ws_b = torch.empty(lib().prt_overlay_batched_workspace_size(npairs, mv), dtype=torch.uint8,
                       device=dev)
    ms_b = timed(lambda: rtree.overlay_batched(pd, Ad, Bd, tune, area_b, ws_b), 5, 1)
    got_p, got_b = area_p.cpu().numpy(), area_b.cpu().numpy()
    out.update({
        "plain_ms": round(ms_p, 3), "plain_pairs_s": round(npairs / ms_p * 1e3, 1),
        "batched_ms": round(ms_b, 3), "batched_pairs_s": round(npairs / ms_b * 1e3, 1),
        "batched_vs_plain": round(ms_p / ms_b, 2),
        "nonempty": int((got_b > 0).sum()),
        "plain_equals_batched": bool(np.array_equal(got_p, got_b)),
        "total_area": float(got_b.sum()),
        "tune": tune_to_json(tune, tstats)["buckets"], "_tune": tune})
    if "whole" in tstats:  # the tuned shapes and the defaults timed on every pair
        out["tune_check"] = tstats["whole"]
    log(f"overlay {npairs} pairs: plain {ms_p:.2f} ms, batched {ms_b:.2f} ms "
        f"({ms_p / ms_b:.2f}x)")

    def check(O):
        sel = np.arange(npairs) if check_sample is None or check_sample >= npairs else \
            np.sort(np.random.default_rng(0).choice(npairs, check_sample, replace=False))
        ref, flops = O.overlay(cp[sel], A, B, nthreads=host_cores(), with_flops=True)
        pa = np.abs([O.poly_area(A.poly(i)) for i in cp[sel, 0]])
        pb = np.abs([O.poly_area(B.poly(i)) for i in cp[sel, 1]])
        den = np.maximum(np.abs(ref), 1e-12 * np.minimum(pa, pb))
        den = np.where(den == 0, 1.0, den)
        rel_p = np.abs(got_p[sel] - ref) / den
        rel_b = np.abs(got_b[sel] - ref) / den
        fpp = flops / len(sel)
        tflops = fpp * npairs / (ms_b * 1e-3) / 1e12
        out.update({
            "checked_pairs": int(len(sel)),
            "max_rel_err_plain": float(rel_p.max()), "max_rel_err_batched": float(rel_b.max()),
            "bit_identical": bool(np.array_equal(got_p[sel], ref)
                                  and np.array_equal(got_b[sel], ref)),
            "fp64_valu": {"oracle_flops_per_pair": round(fpp, 1),
                          "batched_tflops": round(tflops, 4), "peak_tflops": FP64_PEAK_TFLOPS,
                          "frac_of_peak": round(tflops / FP64_PEAK_TFLOPS, 6)}})
        log(f"overlay {npairs} pairs against the oracle: max rel err "
            f"{max(rel_p.max(), rel_b.max()):.1e} on {len(sel)} pairs")
    checks.append(check)

    def cpu(O):
        k = min(cpu_sample, npairs)
        sel = np.sort(np.random.default_rng(1).choice(npairs, k, replace=False))
        cps = np.ascontiguousarray(cp[sel])
        cores = host_cores()
        out["cpu_baseline"] = cpu_rate(
            lambda: O.overlay(cps, A, B, nthreads=1), k, "Mpairs/s", args.cpu_reps, cores,
            f"the oracle's FP64 scanbeam clip (the area of gpc_polygon_clip's GPC_INT sweep) "
            f"of {k} random pairs of the {npairs}",
            lambda: O.overlay(cps, A, B, nthreads=cores))
    cpu_jobs.append(cpu)
    return out


def contours_phase(cp, A, B, dev, checks, cpu_jobs, sample=2000):
    """Result contours (intersection) of the pairs cp: count pass, scan,
    write pass.  Appends to `checks` the vertex-for-vertex comparison of the
    first `sample` pairs with the oracle, timed (the CPU figure)."""
    import numpy as np
    import torch

    from parlllelrtree_amd import rtree

    pd = torch.from_numpy(np.ascontiguousarray(cp).view(np.int32)).to(dev)
    Ad, Bd = rtree.to_device_polys(A, dev), rtree.to_device_polys(B, dev)
    ms = timed(lambda: rtree.overlay_contours(pd, Ad, Bd, "int"), 3, 1)
    got = rtree.overlay_contours(pd, Ad, Bd, "int")
    k = min(sample, len(cp))
    mine = [got.pair(t) for t in range(k)]
    out = {"pairs": len(cp), "ms": round(ms, 3), "pairs_s": round(len(cp) / ms * 1e3, 1),
           "contours": int(got.cont_off[-1]), "vertices": int(got.vert_off[-1])}
    log(f"contours {len(cp)} pairs: {ms:.2f} ms, {out['contours']} contours, "
        f"{out['vertices']} vertices")

    def check(O):
        same = True
        polys = [(A.poly(int(cp[t, 0])), B.poly(int(cp[t, 1]))) for t in range(k)]
        t0 = time.perf_counter()
        refs = [O.clip_contours(P, Q, "int") for P, Q in polys]
        t_cpu = time.perf_counter() - t0
        for t in range(k):
            ref, holes = refs[t]
            same &= len(mine[t]) == len(ref) and all(
                h == bool(rh) and np.array_equal(c, r)
                for (c, h), r, rh in zip(mine[t], ref, holes))
        out.update({"checked_pairs": int(k), "equal_to_oracle": bool(same),
                    "cpu_baseline": {"kind": "port", "unit": "Mpairs/s",
                                     "1_core": round(k / t_cpu / 1e6, 4),
                                     "sample": f"the oracle's contour assembly of the first {k} "
                                               f"pairs, one run (the comparison's own pass)"}})
        log(f"contours against the oracle on {k} pairs: equal {same}")
    checks.append(check)
    return out


def grid_phase(dev, args, checks, cpu_jobs):
    """2 x 500,000 stars on the jittered 708x708 grid: bottom-up tree of
    layer A's MBRs, join with layer B's MBRs, overlay of every pair, and the
    whole pipeline as one graph."""
    import numpy as np
    import torch

    from parlllelrtree_amd import gen, rtree, workloads

    A, B = workloads.w2()
    ma, mb = gen.poly_mbr(A), gen.poly_mbr(B)
    da, db = torch.from_numpy(ma).to(dev), torch.from_numpy(mb).to(dev)
    t = rtree.build_bottom_up(da, 16)
    cap = rtree.join(t, db).shape[0]
    j = rtree.Joiner(t, mb.shape[0], cap, dev)
    ms_j = timed(lambda: j.run(db), 3, 1)
    pqr = j.result().cpu().numpy().view(np.uint32)
    cp = np.ascontiguousarray(pqr[:, ::-1])      # (A index, B index)
    log(f"grid join {ms_j:.3f} ms ({cap} pairs)")
    out = overlay_phase(cp, A, B, dev, args, checks, cpu_jobs, check_sample=300_000,
                        plain_reps=(1, 0))
    out["join_ms"] = round(ms_j, 3)
    out["pipeline"] = pipeline_phase(A, B, cap, out.pop("_tune"), out, dev)
    valu = recorded("clip_fp64_valu.json", "grid_batched", "clip")
    if valu:
        out["fp64_valu_counters"] = valu
    return out


def pipeline_phase(A, B, cap, tune, ov, dev):
    """The whole overlay pipeline (polygon MBRs -> bottom-up tree over B ->
    join with A -> batched clip -> total area) as one hipGraph, replayed;
    against the eager launch sequence and the separately timed stages."""
    import torch

    from parlllelrtree_amd import rtree
    from parlllelrtree_amd.pipeline import OverlayPipeline

    dA, dB = rtree.to_device_polys(A, dev), rtree.to_device_polys(B, dev)
    pipe = OverlayPipeline(dA, dB, cap=cap + 1024, tree="bu", tune=tune)
    ms_e = timed(pipe.run, 3, 1)
    e_pairs, e_area, e_total, e_nonempty = pipe.result()
    e_pairs, e_area = e_pairs.clone(), e_area.clone()
    pipe.capture()
    ms_g = timed(pipe.replay, 5, 1)
    pairs, area, total, nonempty = pipe.result()
    same = bool(torch.equal(pairs, e_pairs) and torch.equal(area.view(torch.int64),
                                                              e_area.view(torch.int64))
                and total == e_total and nonempty == e_nonempty)
    n = int(pairs.shape[0])
    pipe.close()
    log(f"pipeline graph {ms_g:.2f} ms, eager {ms_e:.2f} ms, {n} pairs, replay==eager {same}")
    return {"graph_ms": round(ms_g, 3), "eager_ms": round(ms_e, 3), "pairs": n,
            "pairs_s": round(n / ms_g * 1e3, 1), "total_area": total, "nonempty": nonempty,
            "graph_equals_eager": same, "pairs_match_join": n == cap,
            "nonempty_matches_stage_run": nonempty == ov["nonempty"],
            "total_rel_diff_vs_stage_run": abs(total - ov["total_area"]) / max(abs(total), 1e-300)}


def grid_batched_phase(dev, args, checks):
    """Batched FP64 overlay alone (default team shapes, no plain run, no
    autotune): every candidate pair of the 2 x 500,000-star grid.  The target
    of the FP64 VALU counter runs."""
    import numpy as np
    import torch

    from parlllelrtree_amd import rtree, workloads
    from parlllelrtree_amd._lib import lib
    from parlllelrtree_amd.pipeline import OverlayPipeline

    A, B = workloads.w2()
    dA, dB = rtree.to_device_polys(A, dev), rtree.to_device_polys(B, dev)
    pipe = OverlayPipeline(dA, dB, cap=12_000_000, tree="bu")
    pipe.run()
    pairs, _, _, _ = pipe.result()
    pd = pairs.contiguous()
    npairs = int(pd.shape[0])
    mv = dA.max_verts + dB.max_verts
    area = torch.empty(npairs, dtype=torch.float64, device=dev)
    ws = torch.empty(lib().prt_overlay_batched_workspace_size(npairs, mv), dtype=torch.uint8,
                     device=dev)
    ms = timed(lambda: rtree.overlay_batched(pd, dA, dB, None, area, ws), 3, 1)
    sel = np.sort(np.random.default_rng(0).choice(npairs, 50_000, replace=False))
    cp = pd.cpu().numpy().view(np.uint32)
    got = area.cpu().numpy()[sel]
    pipe.close()
    log(f"batched overlay {npairs} pairs {ms:.2f} ms")
    out = {"pairs": npairs, "batched_ms": round(ms, 3)}

    def check(O):
        ref, flops = O.overlay(cp[sel], A, B, nthreads=host_cores(), with_flops=True)
        fpp = flops / len(sel)
        out.update({"bit_identical_sample": bool(np.array_equal(got, ref)),
                    "oracle_fp64_flops_per_pair": round(fpp, 1),
                    "oracle_fp64_tflops": round(fpp * npairs / (ms * 1e-3) / 1e12, 4)})
        log(f"batched overlay: {fpp:.0f} oracle FP64 flops/pair, sample bit-identical "
            f"{out['bit_identical_sample']}")
    checks.append(check)
    return out


def grad_phase(dev, args, checks, cpu_jobs):
    """262,144 independent star pairs (16..128 vertices): FP64 forward area
    (batched, with the default and the autotuned shapes), FP64 adjoint, FP32 adjoint (+ Green's area), errors against
    central finite differences (h = 1e-6) of the FP64 area, the FP64 CPU
    oracle gradient, and FP32 against FP64."""
    import numpy as np
    import torch

    from parlllelrtree_amd import autograd, rtree, workloads
    from parlllelrtree_amd._lib import lib
    from parlllelrtree_amd.autotune import autotune, tune_to_json

    pairs, S = workloads.w3()
    npairs = len(pairs)
    pd = torch.from_numpy(pairs.view(np.int32)).to(dev)
    S64 = rtree.to_device_polys(S, dev)
    S32 = rtree.to_device_polys(S, dev, dtype=torch.float32)
    area = torch.empty(npairs, dtype=torch.float64, device=dev)
    mv = 2 * S64.max_verts
    ws = torch.empty(lib().prt_overlay_batched_workspace_size(npairs, mv), dtype=torch.uint8,
                     device=dev)
    # the forward with the default shapes, then with the autotuner's for
    # these pairs (as the other overlay phases); the same areas either way
    ms_fd = timed(lambda: rtree.overlay_batched(pd, S64, S64, None, area, ws), 3, 1)
    area_d = area.clone()
    tune, tstats = autotune(pairs, S, S, S64, S64)
    ms_f = timed(lambda: rtree.overlay_batched(pd, S64, S64, tune, area, ws), 3, 1)
    same_tuned = bool(torch.equal(area.view(torch.int64), area_d.view(torch.int64)))
    del area_d
    gA, gB = torch.zeros_like(S64.xy), torch.zeros_like(S64.xy)

    def bwd64():
        gA.zero_()
        gB.zero_()
        rtree.area_grad(pd, S64, S64, gA=gA, gB=gB)
    ms_b = timed(bwd64, 3, 1)
    hA, hB = torch.zeros_like(S32.xy), torch.zeros_like(S32.xy)
    gr32 = {}

    def fb32():
        hA.zero_()
        hB.zero_()
        gr32["r"] = rtree.area_grad(pd, S32, S32, gA=hA, gB=hB, green=True)
    ms_32 = timed(fb32, 3, 1)
    # FP64 forward and adjoint in one launch: the area by Green's theorem
    # over the boundary the adjoint walks (not the sweep's summation order)
    kA, kB = torch.zeros_like(S64.xy), torch.zeros_like(S64.xy)
    gr64 = {}

    def fb64():
        kA.zero_()
        kB.zero_()
        gr64["r"] = rtree.area_grad(pd, S64, S64, gA=kA, gB=kB, green=True)
    ms_fb64 = timed(fb64, 3, 1)
    g64 = (gA + gB).cpu().numpy()
    g32 = (hA + hB).double().cpu().numpy()
    rel32 = np.abs(g32 - g64).max() / np.abs(g64).max()
    a64 = area.cpu().numpy()
    # Green's area against the sweep's (bit-identical to the oracle): the
    # relative error with the overlay denominator max(|A|, 1e-12 min(|P|, |Q|))
    x, y = S.xy[:, 0], S.xy[:, 1]
    nxt = np.arange(len(x)) + 1
    nxt[S.off[1:] - 1] = S.off[:-1]
    shoe = 0.5 * np.abs(np.add.reduceat(x * y[nxt] - x[nxt] * y, S.off[:-1]))
    ag = gr64["r"][2].cpu().numpy()
    den = np.maximum(np.abs(a64), 1e-12 * np.minimum(shoe[pairs[:, 0]], shoe[pairs[:, 1]]))
    den = np.where(den == 0, 1.0, den)
    green_rel = float((np.abs(ag - a64) / den).max())
    fused_same_grad = bool(torch.equal(kA, gA) and torch.equal(kB, gB))
    a32 = gr32["r"][2].double().cpu().numpy()
    arel32 = float(np.median(np.abs(a32 - a64) / np.maximum(np.abs(a64), 1e-300)))
    # central finite differences (h = 1e-6) of the FP64 area on a sample of
    # pairs: every perturbed polygon is one more pair of one batched call
    fsel = np.sort(np.random.default_rng(7).choice(npairs, 128, replace=False))
    fd = autograd.finite_difference_grad(pairs[fsel], S, 1e-6, dev)
    fd_err = []
    for i, (a, b) in enumerate(pairs[fsel]):
        for F, idx in zip(fd[i], (a, b)):
            G = g64[S.off[idx]:S.off[idx + 1]]
            sc = max(np.abs(G).max(), 1e-300)
            fd_err.append(np.abs(F - G).max() / sc)
    out = {"pairs": npairs, "fp64_forward_ms": round(ms_f, 3),
           "fp64_forward_default_shapes_ms": round(ms_fd, 3),
           "forward_tuned_equals_default": same_tuned,
           "tune": tune_to_json(tune, tstats)["buckets"],
           "fp64_adjoint_ms": round(ms_b, 3),
           "fp64_pairs_s": round(npairs / (ms_f + ms_b) * 1e3, 1),
           "fp64_fused_forward_adjoint_ms": round(ms_fb64, 3),
           "fp64_fused_pairs_s": round(npairs / ms_fb64 * 1e3, 1),
           "fp64_fused_area_max_rel_err_vs_sweep": green_rel,
           "fp64_fused_grad_equals_adjoint": fused_same_grad,
           "fp32_forward_adjoint_ms": round(ms_32, 3),
           "fp32_pairs_s": round(npairs / ms_32 * 1e3, 1),
           "grad_rel_err_fp32_vs_fp64": float(rel32), "area_median_rel_err_fp32": arel32,
           "grad_rel_err_vs_central_fd_h1e-6": {"max": float(np.max(fd_err)),
                                                "median": float(np.median(fd_err)),
                                                "pairs": int(len(fsel)),
                                                "scale": "per polygon: max |g - fd| / max |g|"}}
    log(f"grad: fwd {ms_f:.2f} ms (default shapes {ms_fd:.2f} ms), adjoint {ms_b:.2f} ms (fp64), "
        f"fused fp64 {ms_fb64:.2f} ms (Green area rel err {green_rel:.1e}), fp32 {ms_32:.2f} ms; "
        f"fp32 vs fp64 {rel32:.1e}; "
        f"vs central FD max {np.max(fd_err):.1e} on {len(fsel)} pairs")

    csel = np.sort(np.random.default_rng(0).choice(npairs, 2000, replace=False))

    def check(O):
        # FP64 gradient against the CPU oracle on a sample of pairs (timed:
        # the CPU figure of the adjoint)
        worst, exact, tot = 0.0, 0, 0
        t0 = time.perf_counter()
        refs = [O.area_grad(S.poly(a), S.poly(b)) for a, b in pairs[csel]]
        t_cpu = time.perf_counter() - t0
        for (a, b), (gP, gQ) in zip(pairs[csel], refs):
            for G, idx in ((gP, a), (gQ, b)):
                mine = g64[S.off[idx]:S.off[idx + 1]]
                sc = max(np.abs(G).max(), 1e-300)
                worst = max(worst, np.abs(mine - G).max() / sc)
                exact += int((mine == G).sum())
                tot += G.size
        out.update({"grad_rel_err_vs_oracle_fp64": worst,
                    "grad_bit_identical_fraction": exact / tot, "checked_pairs": int(len(csel)),
                    "cpu_baseline": {"kind": "port", "unit": "Mpairs/s",
                                     "1_core": round(len(csel) / t_cpu / 1e6, 4),
                                     "sample": f"the oracle's FP64 adjoint (area + gradient) of "
                                               f"{len(csel)} random pairs, one run (the "
                                               f"comparison's own pass)"}})
        log(f"grad against the oracle on {len(csel)} pairs: rel err {worst:.1e}")
    checks.append(check)
    return out


def w4_overlay_phase(dev, args, checks, cpu_jobs):
    """Overlay of two 4,000,000-star mixture layers: join on the MBRs, then
    plain vs batched clip of every candidate pair."""
    import numpy as np
    import torch

    from parlllelrtree_amd import gen, rtree, workloads

    A, B = workloads.w4_polys()
    ma, mb = gen.poly_mbr(A), gen.poly_mbr(B)
    da, db = torch.from_numpy(ma).to(dev), torch.from_numpy(mb).to(dev)
    t = rtree.build_bottom_up(da, 16)
    cap = rtree.join(t, db).shape[0]
    j = rtree.Joiner(t, mb.shape[0], cap, dev)
    ms_j = timed(lambda: j.run(db), 3, 1)
    pqr = j.result().cpu().numpy().view(np.uint32)
    cp = np.ascontiguousarray(pqr[:, ::-1])
    del j, t
    log(f"w4 join {ms_j:.3f} ms ({cap} pairs)")
    ov = overlay_phase(cp, A, B, dev, args, checks, cpu_jobs, check_sample=200_000,
                       plain_reps=(1, 0))
    ov["join_ms"] = round(ms_j, 3)
    ov.pop("_tune")
    return ov


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU leg: the oracle's timings and the oracle comparisons "
                         "of the phases' results")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="median of this many runs for the phases' CPU figures")
    ap.add_argument("--cpu-reps-large", type=int, default=3,
                    help="median of this many runs for the serial 16.7M-rect builds")
    ap.add_argument("--no-phases", action="store_true")
    ap.add_argument("--phases-only", action="store_true",
                    help="skip the headline step (profiling runs of single phases); the line "
                         "then carries no value")
    ap.add_argument("--phases", default="bu,layers,grid,grad,w4",
                    help="comma list of the phases to run (bu,layers,grid,grad,w4; "
                         "also grid_batched: the batched overlay alone)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        src = torch.ones(args.copy_calib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): every box here has one "
              "GPU and this build has no multi-GPU path (its sharding is a design, not code)",
              file=sys.stderr)
        return 2

    import torch

    from parlllelrtree_amd import rtree

    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}")
    cpu_jobs, checks = [], []     # CPU figures and oracle comparisons, run after the GPU phases
    if args.phases_only:
        head = {"value": float("nan"), "ms": float("nan"), "roofline": None, "cpu": None}
        phases = {}
    else:
        head, w4_builds = headline(dev, args, cpu_jobs)
        phases = {"w4_builds": w4_builds}
    want = set() if args.no_phases else set(args.phases.split(","))
    if "bu" in want:
        phases.update(bu_phase(dev, args, cpu_jobs))
    if "layers" in want:
        phases.update(layers_phase(dev, args, cpu_jobs, checks))
    if "grid" in want:
        phases["overlay_grid_2x500k"] = grid_phase(dev, args, checks, cpu_jobs)
    if "grid_batched" in want:
        phases["overlay_grid_batched"] = grid_batched_phase(dev, args, checks)
    if "grad" in want:
        phases["overlay_grad_262144"] = grad_phase(dev, args, checks, cpu_jobs)
    if "w4" in want:
        phases["overlay_w4_2x4M"] = w4_overlay_phase(dev, args, checks, cpu_jobs)
    rtree.check_device_status()

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        log(f"CPU leg: host {cpu_model()}, nproc {nproc()}, using up to {host_cores()} cores")
        for job in cpu_jobs:
            job(O)
        cpu = head["cpu"] or None
        for chk in checks:
            chk(O)

    line = {
        "metric": METRIC, "value": round(head["value"], 1), "unit": "Mrect/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 boxes, u64 sort keys", "data": "synthetic",
        "config": {"workload": "bottom-up + top-down R-tree builds, fanout 16, of 16,777,216 fp32 "
                               "rects from a 64-component Gaussian mixture (sigma 0.01, sides "
                               "U[1e-5,1e-4]); value counts each rect once per tree",
                   "rects": W4_N, "rects_per_step": 2 * W4_N, "fanout": 16,
                   "parallelism": "1 GPU"},
        "roofline": head["roofline"], "cpu_baseline": cpu,
        "host": {"cpu": cpu_model(), "nproc": nproc(), "cores_used": host_cores()},
        "phases": phases,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
