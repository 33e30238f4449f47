"""Autotuner of the batched overlay's launch shape per vertex bucket.

For every bucket of n1 + n2 (the bounds of prt_tune_default), the pairs of
a workload that fall into it are timed under each feasible team shape
(lanes per pair x waves per block) and the fastest is kept; then the
beam-list capacity (list_factor x bucket bound) is swept for that shape.
The buckets share the GPU in a call, so the search then continues on the
whole call, one bucket at a time (autotune's docstring).  The result is a
TuneStruct for prt_overlay_batched, and optionally a JSON cache:

    {"device": "...", "buckets": [{"max_verts": 32, "lanes": 8, "waves": 4,
                                   "list_factor": 1, "pairs": 1234,
                                   "ms": 0.12}, ...]}
"""
from __future__ import annotations

import ctypes
import json
from pathlib import Path

import numpy as np
import torch

from . import rtree
from ._lib import PRT_TUNE_BUCKETS, TuneStruct, lib

LANES = (8, 16, 32, 64, 128, 256)   # 128, 256: a block of whole waves per pair
WAVES = (1, 2, 4, 8)
FACTORS = (1, 2, 4)
LDS_BYTES = 160 * 1024


def _team_bytes(nv: int, factor: int) -> int:
    """LDS bytes of one team for pairs of up to nv vertices."""
    return int(lib().prt_overlay_team_lds(nv, factor))


def _fits(nv: int, lanes: int, waves: int, factor: int) -> bool:
    """Whether the teams of a block fit one CU's LDS (teams of whole waves,
    lanes > 64, are one per block)."""
    if lanes > 64:
        return waves == 1 and _team_bytes(nv, factor) <= LDS_BYTES
    return waves * (64 // lanes) * _team_bytes(nv, factor) <= LDS_BYTES


def _copy(t: TuneStruct) -> TuneStruct:
    c = TuneStruct()
    ctypes.memmove(ctypes.addressof(c), ctypes.addressof(t), ctypes.sizeof(TuneStruct))
    return c


def _time(fn, reps: int = 3) -> float:
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_to_json(t: TuneStruct, stats=None) -> dict:
    out = {"device": torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu",
           "buckets": []}
    for b in range(PRT_TUNE_BUCKETS):
        e = {"max_verts": t.bucket_max_verts[b], "lanes": t.lanes_per_pair[b],
             "waves": t.waves_per_block[b], "list_factor": t.list_factor[b]}
        if stats and b in stats:
            e.update(stats[b])
        out["buckets"].append(e)
    return out


def tune_from_json(d: dict) -> TuneStruct:
    t = TuneStruct()
    for b, e in enumerate(d["buckets"]):
        t.bucket_max_verts[b] = e["max_verts"]
        t.lanes_per_pair[b] = e["lanes"]
        t.waves_per_block[b] = e["waves"]
        t.list_factor[b] = e.get("list_factor", 1)
    return t


def autotune(pairs: np.ndarray, A, B, Ad, Bd, max_pairs_per_bucket: int = 40_000,
             cache: str | Path | None = None, reps: int = 3,
             whole_check: bool = True) -> tuple[TuneStruct, dict]:
    """pairs: (np, 2) u32 host array; A, B: host PolyCSR; Ad, Bd: device
    PolyDev of the same polygons.  Returns (tune, per-bucket stats).

    First every bucket is timed alone on a sample under each feasible
    shape.  In a call, though, the buckets share the GPU (the largest two
    run on helper streams beside the others), so with whole_check the
    search then continues on the whole call: starting from the faster of the
    sampled picks and the default shapes, each bucket in turn tries its
    neighbouring shapes (half, same and twice the lanes; list factors 1, 2,
    4), keeping a change that makes the whole call at least 1 % faster.
    stats["whole"] records the default shapes' and the final shapes' times
    on every pair and the number of whole calls timed."""
    tune = rtree.default_tune()
    n = np.diff(A.off)[pairs[:, 0]] + np.diff(B.off)[pairs[:, 1]]
    mv = Ad.max_verts + Bd.max_verts
    lo = 0
    stats = {}
    rng = np.random.default_rng(0)
    for b in range(PRT_TUNE_BUCKETS):
        hi = tune.bucket_max_verts[b]
        idx = np.nonzero((n > lo) & (n <= hi))[0]
        lo = hi
        if len(idx) == 0:
            continue
        if len(idx) > max_pairs_per_bucket:
            idx = np.sort(rng.choice(idx, max_pairs_per_bucket, replace=False))
        sub = torch.from_numpy(np.ascontiguousarray(pairs[idx]).view(np.int32)).cuda()
        area = torch.empty(len(idx), dtype=torch.float64, device="cuda")
        nv = min(hi, mv)

        def trial(lanes, waves, factor):
            if not _fits(nv, lanes, waves, factor):
                return None
            cand = rtree.default_tune()
            for k in range(PRT_TUNE_BUCKETS):
                cand.lanes_per_pair[k] = lanes
                cand.waves_per_block[k] = waves
                cand.list_factor[k] = factor
            return _time(lambda: rtree.overlay_batched(sub, Ad, Bd, cand, area), reps)

        best = None
        f0 = tune.list_factor[b]
        for lanes in LANES:
            for waves in WAVES:
                ms = trial(lanes, waves, f0)
                if ms is not None and (best is None or ms < best[0]):
                    best = (ms, lanes, waves, f0)
        if best is not None:
            for factor in FACTORS:
                if factor == f0:
                    continue
                ms = trial(best[1], best[2], factor)
                if ms is not None and ms < best[0]:
                    best = (ms, best[1], best[2], factor)
            tune.lanes_per_pair[b] = best[1]
            tune.waves_per_block[b] = best[2]
            tune.list_factor[b] = best[3]
            stats[b] = {"pairs": int(len(idx)), "ms": round(best[0], 4)}
    if whole_check and len(pairs):
        allp = torch.from_numpy(np.ascontiguousarray(pairs).view(np.int32)).cuda()
        area = torch.empty(len(pairs), dtype=torch.float64, device="cuda")

        def whole(t):
            return _time(lambda: rtree.overlay_batched(allp, Ad, Bd, t, area), 2)

        default = rtree.default_tune()
        t_def = whole(default)
        t_best = whole(tune)
        calls = 2
        if t_def < t_best:
            tune, t_best = _copy(default), t_def
        # buckets with pairs, the most pairs first
        lo, order = 0, []
        for b in range(PRT_TUNE_BUCKETS):
            hi = tune.bucket_max_verts[b]
            cnt = int(((n > lo) & (n <= hi)).sum())
            lo = hi
            if cnt:
                order.append((cnt, b))
        for _, b in sorted(order, reverse=True):
            nv = min(tune.bucket_max_verts[b], mv)
            l0 = tune.lanes_per_pair[b]
            for lanes in sorted({max(8, l0 // 2), l0, min(256, l0 * 2)}):
                waves = 1 if lanes >= 64 else 2
                for factor in FACTORS:
                    if (lanes, factor) == (l0, tune.list_factor[b]):
                        continue
                    if not _fits(nv, lanes, waves, factor):
                        continue
                    cand = _copy(tune)
                    cand.lanes_per_pair[b], cand.waves_per_block[b] = lanes, waves
                    cand.list_factor[b] = factor
                    t = whole(cand)
                    calls += 1
                    if t < 0.99 * t_best:
                        tune, t_best = cand, t
        stats["whole"] = {"tuned_ms": round(t_best, 3), "default_ms": round(t_def, 3),
                          "calls_timed": calls}
        del allp, area
    if cache is not None:
        Path(cache).write_text(json.dumps(tune_to_json(tune, stats), indent=1) + "\n")
    return tune, stats
