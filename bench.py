#!/usr/bin/env python3
"""bench.py -- R-tree build throughput (and the join / overlay phases) on one
MI355X, printed as one JSON line.

Workload (the step): top-down R-tree build, fanout 16, of both layers of
1,048,576 fp32 rects (centres U[0,1]^2, sides U[1e-4, 1e-3]), inputs
resident in HBM, both layers in one shared launch sequence
(prt_td_build_layers); value = rects built per second (Mrect/s).  The same
line carries, measured separately: the same builds one layer per call, the
bottom-up build of 1,048,576 rects, the 65,536-window join, the layer-layer
join, and the plain FP64 overlay of the first 100,000 sorted layer pairs.
roofline: the dominant kernel of the step, its average duration from HIP
events that hipExtLaunchKernel stamps with each launch's own start and end.
cpu_baseline: the oracle's serial top-down build on the host; the same CPU
leg then compares the phases' GPU results (copied to the host when each
phase ends) with the oracle and adds the error fields to each phase.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "R-tree build Mrect/s; overlay pairs/s batched vs plain at ≤1e-12 area err, 1 GPU"


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


# algorithmic HBM bytes per processed item, per kernel (minimal traffic:
# every input read once, every output written once)
BYTES_PER_ITEM = {
    "k_onesweep": 16,       # one digit pass: 8-B item read + 8-B item written (+ GATHER_BYTES)
    "k_hist_all<Src>": 24,  # first radix level: 16-B rect read, 8-B item written (8-B read after)
    "k_td_gather": 48,      # 8-B item, 16-B rect + 4-B id read; 16 + 4 written
    "k_td_subtree<8>": 52,  # 8-B item, 16-B rect, 4-B id read; 16-B box + 4-B child + 4-B perm written
    "k_td_subtree<4>": 52,  # the same kernel when each block has a CU to itself
    "k_td_lds_sort": 40,    # (fanouts that are not powers of two) 16 + 4 read, 16 + 4 written
    "k_td_items": 24,       # (fanouts that are not powers of two) 16-B rect read, 8-B item written
    "k_td_leaves": 20,      # (fanouts that are not powers of two) 16-B rect read, 4-B child written
    "k_bu_leaves": 52,      # 8-B item, 16-B rect read; 4+4 key/id, 16+4 entry written
}


# the last digit pass of a top-down radix level that feeds another radix
# level also moves the rects and ids into the sorted order: 16-B rect + 4-B
# id read, 16 + 4 written per item
GATHER_BYTES = 40
TD_LDS_SLOTS = 4096  # segments up to this size are sorted in LDS by the subtree kernel


def td_gather_passes(n: int, f: int, levels: int) -> int:
    """Digit passes per top-down build that also gather (one per radix level
    followed by another radix level)."""
    radix_levels = 0
    for p in range(levels):
        a = 1
        for _ in range(levels - p):
            a *= f
            if a >= n:
                a = n
                break
        radix_levels += a > TD_LDS_SLOTS
    return max(0, radix_levels - 1)


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


def overlay_phase(cp, A, B, dev, args, checks, check_sample=None, plain_reps=(2, 1)):
    """Plain vs batched FP64 overlay of pairs cp over CSR sets A, B.  Appends
    to `checks` the comparison of the areas with the oracle (on
    `check_sample` pairs if given), run after every GPU phase."""
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
    log(f"overlay {npairs} pairs: plain {ms_p:.2f} ms, batched {ms_b:.2f} ms "
        f"({ms_p / ms_b:.2f}x)")

    def check(O):
        sel = np.arange(npairs) if check_sample is None or check_sample >= npairs else \
            np.sort(np.random.default_rng(0).choice(npairs, check_sample, replace=False))
        ref, flops = O.overlay(cp[sel], A, B, nthreads=min(os.cpu_count() or 1, 16),
                               with_flops=True)
        pa = np.abs([O.poly_area(A.poly(i)) for i in cp[sel, 0]])
        pb = np.abs([O.poly_area(B.poly(i)) for i in cp[sel, 1]])
        den = np.maximum(np.abs(ref), 1e-12 * np.minimum(pa, pb))
        den = np.where(den == 0, 1.0, den)
        rel_p = np.abs(got_p[sel] - ref) / den
        rel_b = np.abs(got_b[sel] - ref) / den
        flops_all = flops * npairs / len(sel)
        out.update({
            "checked_pairs": int(len(sel)),
            "max_rel_err_plain": float(rel_p.max()), "max_rel_err_batched": float(rel_b.max()),
            "bit_identical": bool(np.array_equal(got_p[sel], ref)
                                  and np.array_equal(got_b[sel], ref)),
            "oracle_fp64_flops_per_pair": round(flops / len(sel), 1),
            "batched_fp64_tflops": round(flops_all / (ms_b * 1e-3) / 1e12, 4)})
        log(f"overlay {npairs} pairs against the oracle: max rel err "
            f"{max(rel_p.max(), rel_b.max()):.1e} on {len(sel)} pairs")
    checks.append(check)
    return out


def contours_phase(cp, A, B, dev, checks, sample=2000):
    """Result contours (intersection) of the pairs cp: count pass, scan,
    write pass.  Appends to `checks` the vertex-for-vertex comparison of the
    first `sample` pairs with the oracle."""
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
        for t in range(k):
            ref, holes = O.clip_contours(A.poly(int(cp[t, 0])), B.poly(int(cp[t, 1])), "int")
            same &= len(mine[t]) == len(ref) and all(
                h == bool(rh) and np.array_equal(c, r)
                for (c, h), r, rh in zip(mine[t], ref, holes))
        out.update({"checked_pairs": int(k), "equal_to_oracle": bool(same)})
        log(f"contours against the oracle on {k} pairs: equal {same}")
    checks.append(check)
    return out


def grid_phase(dev, args, checks):
    """2 x 500,000 stars on the jittered 708x708 grid: bottom-up tree of
    layer A's MBRs, join with layer B's MBRs, overlay of every pair."""
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
    out = overlay_phase(cp, A, B, dev, args, checks, check_sample=300_000, plain_reps=(1, 0))
    out["join_ms"] = round(ms_j, 3)
    out["pipeline"] = pipeline_phase(A, B, cap, out.pop("_tune"), out, dev)
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
    of the FP64 VALU counter runs.  Appends to `checks` the comparison with
    the oracle and its FP64 flops per pair on a 50,000-pair sample."""
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
        ref, flops = O.overlay(cp[sel], A, B, nthreads=min(os.cpu_count() or 1, 16),
                               with_flops=True)
        fpp = flops / len(sel)
        out.update({"bit_identical_sample": bool(np.array_equal(got, ref)),
                    "oracle_fp64_flops_per_pair": round(fpp, 1),
                    "oracle_fp64_tflops": round(fpp * npairs / (ms * 1e-3) / 1e12, 4)})
        log(f"batched overlay: {fpp:.0f} oracle FP64 flops/pair, sample bit-identical "
            f"{out['bit_identical_sample']}")
    checks.append(check)
    return out


def grad_phase(dev, args, checks):
    """262,144 independent star pairs (16..128 vertices): FP64 forward area
    (batched), FP64 adjoint, FP32 adjoint (+ Green's area), errors against the
    FP64 CPU oracle gradient and FP32 against FP64."""
    import numpy as np
    import torch

    from parlllelrtree_amd import rtree, workloads
    from parlllelrtree_amd._lib import lib

    pairs, S = workloads.w3()
    npairs = len(pairs)
    pd = torch.from_numpy(pairs.view(np.int32)).to(dev)
    S64 = rtree.to_device_polys(S, dev)
    S32 = rtree.to_device_polys(S, dev, dtype=torch.float32)
    area = torch.empty(npairs, dtype=torch.float64, device=dev)
    mv = 2 * S64.max_verts
    ws = torch.empty(lib().prt_overlay_batched_workspace_size(npairs, mv), dtype=torch.uint8,
                     device=dev)
    ms_f = timed(lambda: rtree.overlay_batched(pd, S64, S64, None, area, ws), 3, 1)
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
    out = {"pairs": npairs, "fp64_forward_ms": round(ms_f, 3), "fp64_adjoint_ms": round(ms_b, 3),
           "fp64_pairs_s": round(npairs / (ms_f + ms_b) * 1e3, 1),
           "fp64_fused_forward_adjoint_ms": round(ms_fb64, 3),
           "fp64_fused_pairs_s": round(npairs / ms_fb64 * 1e3, 1),
           "fp64_fused_area_max_rel_err_vs_sweep": green_rel,
           "fp64_fused_grad_equals_adjoint": fused_same_grad,
           "fp32_forward_adjoint_ms": round(ms_32, 3),
           "fp32_pairs_s": round(npairs / ms_32 * 1e3, 1),
           "grad_rel_err_fp32_vs_fp64": float(rel32), "area_median_rel_err_fp32": arel32}
    log(f"grad: fwd {ms_f:.2f} ms, adjoint {ms_b:.2f} ms (fp64), fused fp64 {ms_fb64:.2f} ms "
        f"(Green area rel err {green_rel:.1e}), fp32 {ms_32:.2f} ms; fp32 vs fp64 {rel32:.1e}")

    def check(O):
        # FP64 gradient against the CPU oracle on a sample of pairs
        sel = np.sort(np.random.default_rng(0).choice(npairs, 2000, replace=False))
        worst, exact, tot = 0.0, 0, 0
        for t in sel:
            a, b = pairs[t]
            gP, gQ = O.area_grad(S.poly(a), S.poly(b))
            for G, idx in ((gP, a), (gQ, b)):
                mine = g64[S.off[idx]:S.off[idx + 1]]
                sc = max(np.abs(G).max(), 1e-300)
                worst = max(worst, np.abs(mine - G).max() / sc)
                exact += int((mine == G).sum())
                tot += G.size
        out.update({"grad_rel_err_vs_oracle_fp64": worst,
                    "grad_bit_identical_fraction": exact / tot, "checked_pairs": int(len(sel))})
        log(f"grad against the oracle on {len(sel)} pairs: rel err {worst:.1e}")
    checks.append(check)
    return out


def w4_phase(dev, args, checks):
    """16,777,216 Gaussian-mixture rects: bottom-up and top-down build
    throughput and the dominant build kernel against HBM (the working set
    exceeds the 256 MB Infinity Cache); overlay of two 4,000,000-star layers
    (join + plain vs batched)."""
    import numpy as np
    import torch

    from parlllelrtree_amd import gen, rtree, workloads
    from parlllelrtree_amd.profiling import KernelTimes

    n = 1 << 24
    r = workloads.w4_rects(n)
    d = torch.from_numpy(r).to(dev)
    tb = rtree.build_bottom_up(d, 16)
    ms_bu = timed(lambda: rtree.rebuild(tb, d), 5, 1)
    tt = rtree.build_top_down(d, 16)
    ms_td = timed(lambda: rtree.rebuild(tt, d), 3, 1)
    with KernelTimes() as kt:
        rtree.rebuild(tb, d)
        rtree.rebuild(tt, d)
        torch.cuda.synchronize()
    # k_scatter over all 16.7M items: 16 B per item algorithmic
    sc = [v for k, v in kt.stats.items() if k == "k_scatter"]
    out = {"rects": n, "bu_ms": round(ms_bu, 3), "bu_Mrect_s": round(n / ms_bu / 1e3, 1),
           "td_ms": round(ms_td, 3), "td_Mrect_s": round(n / ms_td / 1e3, 1),
           "kernels_ms": {k: round(v[1], 3) for k, v in
                          sorted(kt.stats.items(), key=lambda kv: -kv[1][1])[:8]}}
    if sc:
        cnt, tot = sc[0]
        # the unsegmented passes (BU and TD level 0) process all n items per launch
        out["k_scatter_avg_us"] = round(tot / cnt * 1e3, 2)
    log(f"w4 builds: bu {ms_bu:.3f} ms ({n / ms_bu / 1e3:.0f} Mrect/s), td {ms_td:.3f} ms")
    del tb, tt, d
    torch.cuda.empty_cache()
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
    ov = overlay_phase(cp, A, B, dev, args, checks, check_sample=200_000, plain_reps=(1, 0))
    ov["join_ms"] = round(ms_j, 3)
    ov.pop("_tune")
    out["overlay_2x4M"] = ov
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU leg: the oracle's serial build and the oracle "
                         "comparisons of the phases' results")
    ap.add_argument("--no-phases", action="store_true")
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
        print(f"bench.py measures one GPU only (asked for {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from parlllelrtree_amd import rtree, workloads
    from parlllelrtree_amd.profiling import KernelTimes

    dev = torch.device("cuda:0")
    log(f"device {torch.cuda.get_device_name(0)}")
    n = 1 << 20
    ra, rb = workloads.w1(n)
    da, db = torch.from_numpy(ra).to(dev), torch.from_numpy(rb).to(dev)
    # both layers in one shared launch sequence (every launch covers both)
    TL = rtree.TopDownLayers([da, db], 16)
    ta = TL.trees[0]
    torch.cuda.synchronize()

    def step():
        TL.rebuild([da, db])

    log(f"timing top-down build of 2 x {n} rects: {args.warmup} warmup + {args.steps} steps")
    ms = timed(step, args.steps, args.warmup)
    value = 2 * n / (ms * 1e-3) / 1e6
    log(f"step {ms:.3f} ms -> {value:.1f} Mrect/s")

    # dominant kernel of the step; durations from HIP events that
    # hipExtLaunchKernel stamps with each launch's own start and end
    with KernelTimes() as kt:
        for _ in range(3):
            step()
        torch.cuda.synchronize()
    dom = max((k for k in kt.stats if k in BYTES_PER_ITEM), key=lambda k: kt.stats[k][1])
    cnt, tot = kt.stats[dom]
    avg_ms = tot / cnt
    items = 2 * n                      # every launch of the step covers both layers
    alg_bytes = BYTES_PER_ITEM[dom] * items
    if dom == "k_onesweep":
        # mean over the step's launches, gathering passes included
        per_step = cnt / 3
        gathers = td_gather_passes(n, 16, ta.st.num_levels)
        alg_bytes = int(items * (16 * per_step + GATHER_BYTES * gathers) / per_step)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    # (scaled by a copy-kernel calibration), when recorded for this kernel
    traffic = None
    tpath = ROOT / "profiles" / "hbm_traffic.json"
    if tpath.exists():
        rec = json.loads(tpath.read_text()).get(dom)
        if rec and rec.get("items_per_launch") == items:
            traffic = rec["traffic_bytes"]
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "avg_us": round(avg_ms * 1e3, 2),
                "bytes_per_launch": alg_bytes,
                "kernels_ms_per_step": {k: round(v[1] / 3, 4) for k, v in
                                        sorted(kt.stats.items(), key=lambda kv: -kv[1][1])}}
    log(f"dominant kernel {dom}: {avg_ms * 1e3:.1f} us avg, {achieved:.0f} GB/s algorithmic")

    phases = {}
    checks = []     # oracle comparisons of the phases' results, run with the CPU baseline
    # the same two builds, one prt_td_build call per layer
    t1a, t1b = rtree.build_top_down(da, 16), rtree.build_top_down(db, 16)

    def step_per_layer():
        rtree.rebuild(t1a, da)
        rtree.rebuild(t1b, db)
    ms1 = timed(step_per_layer, args.steps, args.warmup)
    phases["td_build_one_layer_per_call"] = {"ms": round(ms1, 4),
                                             "Mrect_s": round(2 * n / ms1 / 1e3, 1)}
    log(f"one layer per call: {ms1:.3f} ms -> {2 * n / ms1 / 1e3:.1f} Mrect/s")
    del t1a, t1b
    want = set() if args.no_phases else set(args.phases.split(","))
    if "bu" in want:
        # bottom-up build of 1,048,576 rects + 65,536 windows
        r0, q0 = workloads.w0()
        d0, dq0 = torch.from_numpy(r0).to(dev), torch.from_numpy(q0).to(dev)
        t0 = rtree.build_bottom_up(d0, 16)
        ms_bu = timed(lambda: rtree.rebuild(t0, d0), args.steps, args.warmup)
        phases["bu_build_1M"] = {"ms": round(ms_bu, 4), "Mrect_s": round(n / ms_bu / 1e3, 1)}
        cap0 = rtree.join(t0, dq0).shape[0]
        j0 = rtree.Joiner(t0, q0.shape[0], cap0, dev)
        ms_j0 = timed(lambda: j0.run(dq0), max(args.steps // 2, 3), 1)
        phases["join_windows_65536"] = {"ms": round(ms_j0, 4), "pairs": cap0,
                                        "Mpairs_s": round(cap0 / ms_j0 / 1e3, 1)}
        log(f"bu build {ms_bu:.3f} ms, window join {ms_j0:.3f} ms ({cap0} pairs)")
    if "layers" in want:
        # layer-layer join (top-down tree of A, queries = B)
        cap1 = rtree.join(ta, db).shape[0]
        j1 = rtree.Joiner(ta, n, cap1, dev)
        ms_j1 = timed(lambda: j1.run(db), max(args.steps // 2, 3), 1)
        pairs_qr = j1.result().cpu().numpy().view(np.uint32)
        phases["join_layers_1M_x_1M"] = {"ms": round(ms_j1, 4), "pairs": cap1,
                                         "Mpairs_s": round(cap1 / ms_j1 / 1e3, 1)}
        log(f"layer join {ms_j1:.3f} ms ({cap1} pairs)")
        # plain and batched FP64 overlay of the first 100,000 sorted pairs
        cp, A, B, _, _ = workloads.clip_inputs_from_pairs(pairs_qr, ra, rb, 100_000)
        phases["overlay_layers_100k"] = overlay_phase(cp, A, B, dev, args, checks)
        phases["contours_layers_100k"] = contours_phase(cp, A, B, dev, checks)
        phases["overlay_layers_100k"].pop("_tune")
    if "grid" in want:
        # jittered-grid overlay: 2 x 500,000 lognormal stars, every candidate
        # pair, stage by stage and as one captured pipeline
        phases["overlay_grid_2x500k"] = grid_phase(dev, args, checks)
    if "grid_batched" in want:
        phases["overlay_grid_batched"] = grid_batched_phase(dev, args, checks)
    if "grad" in want:
        # differentiable overlay: 262,144 independent pairs, area + adjoint
        phases["overlay_grad_262144"] = grad_phase(dev, args, checks)
    if "w4" in want:
        # largest workload: 16.7M-rect builds, 2 x 4M-star overlay
        phases["w4_gmm"] = w4_phase(dev, args, checks)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.td_build(ra, 16)
            t.append(time.perf_counter() - t0)
        t_med = sorted(t)[1]
        cpu = {"value": round(n / t_med / 1e6, 3), "unit": "Mrect/s", "cores": 1, "kind": "port",
               "sample": f"serial top-down build of one 1,048,576-rect layer (oracle, C), "
                         f"median of 3 = {t_med:.3f} s; host {cpu_model()}"}
        log(f"cpu baseline {cpu['value']} Mrect/s")
        for chk in checks:
            chk(O)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mrect/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 boxes, u64 sort keys", "data": "synthetic",
        "config": {"workload": "top-down R-tree build, fanout 16, two layers of 1,048,576 "
                               "uniform fp32 rects (sides U[1e-4,1e-3]), both layers in one "
                               "launch sequence",
                   "rects_per_step": 2 * n, "fanout": 16, "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "phases": phases,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
