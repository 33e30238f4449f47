"""Host-side mirror of the reference's R-tree / overlay operators on torch
device tensors, calling libprt through its C-ABI.

Reference operators (BU = BotomUp-rtree.cu, TD = TopDown-rtree.cu,
OV = vector_map_overlay_gpc_kernel.cu):
  build_bottom_up   <- the GPU build of BU main (:836-869)
  build_top_down    <- the GPU build of TD main (:629-679)
  join              <- the query stage (BU:914-987, TD:694-758) with the hit
                       semantics of RtreeSearch (BU:439-483, TD:56-76)
  overlay_plain     <- run/Execute (OV:61-113)
Errors raise PrtError (the reference prints and exits, OV:25-29).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import torch

from ._lib import BUTreeStruct, PrtError, TDTreeStruct, TuneStruct, check, lib

__all__ = ["BottomUpTree", "TopDownTree", "build_bottom_up", "build_top_down", "join",
           "overlay_plain", "PolyDev", "to_device_polys", "morton_center30", "sort_pairs_u32"]


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def device_status(clear: bool = True) -> int:
    """Waits for the current stream and returns this device's PRT_STATUS_*
    flags (0: no kernel raised one since the last clear); clears them when
    `clear`.  Raises nothing: callers decide (see check_device_status)."""
    flags = ctypes.c_int32(0)
    lib().prt_device_status(ctypes.byref(flags), 1 if clear else 0, _stream())
    return int(flags.value)


def check_device_status() -> None:
    """Raises PrtError(PRT_E_DEVICE) when a kernel raised a status flag (a
    radix look-back that gave up, int64 coordinates fp32 cannot hold); the
    flags are cleared either way."""
    check(lib().prt_device_status(None, 1, _stream()))


def td_levels(n: int, fanout: int = 16) -> int:
    """Levels of the top-down slot layout of n rects (TopDown-rtree.cu:504-524)."""
    st = TDTreeStruct()
    check(lib().prt_td_plan(n, fanout, ctypes.byref(st)))
    return int(st.num_levels)


def _ws(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def _cuda_f32_rects(rects: torch.Tensor) -> torch.Tensor:
    if not (rects.is_cuda and rects.dtype == torch.float32 and rects.dim() == 2
            and rects.shape[1] == 4):
        raise ValueError("rects must be a (n, 4) float32 CUDA tensor (xmin, ymin, xmax, ymax)")
    return rects.contiguous()


@dataclass
class BottomUpTree:
    st: BUTreeStruct
    keys: torch.Tensor        # int32 view of u32 keys, sorted
    perm: torch.Tensor        # int32 (u32 rect ids) in leaf order
    node_count: torch.Tensor
    node_start: torch.Tensor
    node_end: torch.Tensor
    entry_mbr: torch.Tensor   # (num_nodes * fanout, 4) float32
    entry_ref: torch.Tensor
    workspace: torch.Tensor | None = field(default=None, repr=False)

    @property
    def num_nodes(self) -> int:
        return self.st.num_nodes

    @property
    def root(self) -> int:
        return self.st.num_nodes - 1

    @property
    def levels(self) -> list[tuple[int, int]]:
        return [(self.st.level_start[i], self.st.level_count[i]) for i in range(self.st.num_levels)]


@dataclass
class TopDownTree:
    st: TDTreeStruct
    slot_child: torch.Tensor
    slot_mbr: torch.Tensor    # (num_slots, 4) float32
    perm: torch.Tensor
    workspace: torch.Tensor | None = field(default=None, repr=False)

    @property
    def level_offset(self) -> list[int]:
        return [self.st.level_offset[i] for i in range(self.st.num_levels + 1)]


def morton_center30(rects: torch.Tensor) -> torch.Tensor:
    rects = _cuda_f32_rects(rects)
    keys = torch.empty(rects.shape[0], dtype=torch.int32, device=rects.device)
    check(lib().prt_morton_center30(_ptr(rects), rects.shape[0], _ptr(keys), _stream()))
    return keys


def sort_pairs_u32(keys: torch.Tensor, key_bits: int = 32) -> tuple[torch.Tensor, torch.Tensor]:
    n = keys.shape[0]
    out_k = torch.empty_like(keys)
    out_i = torch.empty(n, dtype=torch.int32, device=keys.device)
    ws = _ws(lib().prt_sort_workspace_size(n), keys.device)
    check(lib().prt_sort_pairs_u32(_ptr(keys), n, key_bits, _ptr(out_k), _ptr(out_i), _ptr(ws),
                                   ws.numel(), _stream()))
    return out_k, out_i


def build_bottom_up(rects: torch.Tensor, fanout: int = 16, workspace: torch.Tensor | None = None
                    ) -> BottomUpTree:
    rects = _cuda_f32_rects(rects)
    n, dev = rects.shape[0], rects.device
    st = BUTreeStruct()
    check(lib().prt_bu_plan(n, fanout, ctypes.byref(st)))
    nn = st.num_nodes
    i32 = dict(dtype=torch.int32, device=dev)
    t = BottomUpTree(st, torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(nn, **i32),
                     torch.empty(nn, **i32), torch.empty(nn, **i32),
                     torch.empty((nn * fanout, 4), dtype=torch.float32, device=dev),
                     torch.empty(nn * fanout, **i32))
    for name in ("keys", "perm", "node_count", "node_start", "node_end", "entry_mbr", "entry_ref"):
        setattr(st, name, getattr(t, name).data_ptr())
    need = lib().prt_bu_workspace_size(n, fanout)
    ws = workspace if workspace is not None and workspace.numel() >= need else _ws(need, dev)
    check(lib().prt_bu_build(_ptr(rects), n, fanout, ctypes.byref(st), _ptr(ws), ws.numel(),
                             _stream()))
    t.workspace = ws
    return t


def build_bottom_up_ref4d(rects_i64: torch.Tensor, scale: int, fanout: int = 16
                          ) -> tuple[BottomUpTree, torch.Tensor]:
    """Bottom-up tree in the reference's own key order (REF4D64 keys of
    BotomUp-rtree.cu:73-123 on rect // scale, stable by (key, id)).
    rects_i64: (n, 4) int64 CUDA tensor (x1, y1, x2, y2), normalised as the
    reference does (refio.normalize).  Returns (tree, sorted 64-bit keys)."""
    if not (rects_i64.is_cuda and rects_i64.dtype == torch.int64 and rects_i64.dim() == 2
            and rects_i64.shape[1] == 4):
        raise ValueError("rects_i64 must be a (n, 4) int64 CUDA tensor (x1, y1, x2, y2)")
    rects_i64 = rects_i64.contiguous()
    n, dev = rects_i64.shape[0], rects_i64.device
    st = BUTreeStruct()
    check(lib().prt_bu_plan(n, fanout, ctypes.byref(st)))
    nn = st.num_nodes
    i32 = dict(dtype=torch.int32, device=dev)
    t = BottomUpTree(st, torch.empty(n, **i32), torch.empty(n, **i32), torch.empty(nn, **i32),
                     torch.empty(nn, **i32), torch.empty(nn, **i32),
                     torch.empty((nn * fanout, 4), dtype=torch.float32, device=dev),
                     torch.empty(nn * fanout, **i32))
    for name in ("keys", "perm", "node_count", "node_start", "node_end", "entry_mbr", "entry_ref"):
        setattr(st, name, getattr(t, name).data_ptr())
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    ws = _ws(lib().prt_bu_ref4d_workspace_size(n, fanout), dev)
    check(lib().prt_bu_build_ref4d(_ptr(rects_i64), n, fanout, int(scale), ctypes.byref(st),
                                   _ptr(keys), _ptr(ws), ws.numel(), _stream()))
    t.workspace = None
    return t, keys


def rebuild(tree: BottomUpTree | TopDownTree, rects: torch.Tensor) -> None:
    """Rebuild `tree` in place from rects of the same count (reuses every
    output array and the workspace; launches only)."""
    if isinstance(tree, BottomUpTree):
        fn = lib().prt_bu_build
    else:
        fn = lib().prt_td_build
    check(fn(_ptr(rects), tree.st.num_rects, tree.st.fanout, ctypes.byref(tree.st),
             _ptr(tree.workspace), tree.workspace.numel(), _stream()))


def build_top_down(rects: torch.Tensor, fanout: int = 16, workspace: torch.Tensor | None = None
                   ) -> TopDownTree:
    rects = _cuda_f32_rects(rects)
    n, dev = rects.shape[0], rects.device
    st = TDTreeStruct()
    check(lib().prt_td_plan(n, fanout, ctypes.byref(st)))
    ns = st.num_slots
    t = TopDownTree(st, torch.empty(ns, dtype=torch.int32, device=dev),
                    torch.empty((ns, 4), dtype=torch.float32, device=dev),
                    torch.empty(n, dtype=torch.int32, device=dev))
    for name in ("slot_child", "slot_mbr", "perm"):
        setattr(st, name, getattr(t, name).data_ptr())
    need = lib().prt_td_workspace_size(n, fanout)
    ws = workspace if workspace is not None and workspace.numel() >= need else _ws(need, dev)
    check(lib().prt_td_build(_ptr(rects), n, fanout, ctypes.byref(st), _ptr(ws), ws.numel(),
                             _stream()))
    t.workspace = ws
    return t


class TopDownLayers:
    """Top-down trees of several layers of equal size, built by one shared
    launch sequence (prt_td_build_layers): every launch covers all layers.
    trees[l] equals build_top_down(layers[l])."""

    def __init__(self, layers: list[torch.Tensor], fanout: int = 16):
        from ._lib import PRT_MAX_LAYERS
        if not 1 <= len(layers) <= PRT_MAX_LAYERS:
            raise ValueError(f"1..{PRT_MAX_LAYERS} layers, got {len(layers)}")
        layers = [_cuda_f32_rects(r) for r in layers]
        n, dev = layers[0].shape[0], layers[0].device
        if any(r.shape[0] != n for r in layers):
            raise ValueError("layers of unequal size")
        nl = len(layers)
        self.fanout = fanout
        self.n = n
        self.sts = (TDTreeStruct * nl)()
        self.trees = []
        for l in range(nl):
            check(lib().prt_td_plan(n, fanout, ctypes.byref(self.sts[l])))
            ns = self.sts[l].num_slots
            t = TopDownTree(self.sts[l], torch.empty(ns, dtype=torch.int32, device=dev),
                            torch.empty((ns, 4), dtype=torch.float32, device=dev),
                            torch.empty(n, dtype=torch.int32, device=dev))
            for name in ("slot_child", "slot_mbr", "perm"):
                setattr(self.sts[l], name, getattr(t, name).data_ptr())
            self.trees.append(t)
        self.workspace = _ws(lib().prt_td_layers_workspace_size(nl, n, fanout), dev)
        self.rebuild(layers)

    def rebuild(self, layers: list[torch.Tensor]) -> None:
        """Rebuild every tree in place from layers of the same sizes."""
        if len(layers) != len(self.trees) or any(r.shape[0] != self.n for r in layers):
            raise ValueError("layers differ from the ones the trees were planned for")
        ptrs = (ctypes.c_void_p * len(layers))(*[r.data_ptr() for r in layers])
        check(lib().prt_td_build_layers(len(layers), ptrs, self.n, self.fanout, self.sts,
                                        _ptr(self.workspace), self.workspace.numel(), _stream()))


class Joiner:
    """Preallocated join of a fixed query count: run() only launches."""

    def __init__(self, tree: BottomUpTree | TopDownTree, nq: int, cap: int, device):
        self.tree, self.nq, self.cap = tree, nq, cap
        self.fn = lib().prt_join_bu if isinstance(tree, BottomUpTree) else lib().prt_join_td
        self.count = torch.zeros(2, dtype=torch.int64, device=device)
        self.ws = _ws(lib().prt_join_workspace_size(nq), device)
        self.pairs = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=device)

    def run(self, queries: torch.Tensor) -> None:
        check(self.fn(ctypes.byref(self.tree.st), _ptr(queries), self.nq, _ptr(self.pairs),
                      self.cap, _ptr(self.count), _ptr(self.ws), self.ws.numel(), _stream()))

    def result(self) -> torch.Tensor:
        total, status = (int(v) for v in self.count.tolist())
        if total > self.cap or status:
            raise RuntimeError(f"join: total {total} cap {self.cap} status {status}")
        return self.pairs[:total]


def join(tree: BottomUpTree | TopDownTree, queries: torch.Tensor, cap: int | None = None
         ) -> torch.Tensor:
    """All (q, r) with overlap(queries[q], rect r), sorted by (q, r): an
    (np, 2) int32 tensor holding u32 ids.  With cap=None the pair count is
    read back first (one device sync) and the buffer sized exactly."""
    queries = _cuda_f32_rects(queries)
    nq, dev = queries.shape[0], queries.device
    fn = lib().prt_join_bu if isinstance(tree, BottomUpTree) else lib().prt_join_td
    count = torch.zeros(2, dtype=torch.int64, device=dev)
    ws = _ws(lib().prt_join_workspace_size(nq), dev)
    if cap is None:
        check(fn(ctypes.byref(tree.st), _ptr(queries), nq, None, 0, _ptr(count), _ptr(ws),
                 ws.numel(), _stream()))
        cap = int(count[0].item())
    pairs = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=dev)
    check(fn(ctypes.byref(tree.st), _ptr(queries), nq, _ptr(pairs), cap, _ptr(count), _ptr(ws),
             ws.numel(), _stream()))
    total, status = (int(v) for v in count.tolist())
    if total > cap:
        raise RuntimeError(f"join produced {total} pairs > cap {cap}")
    if status:
        raise RuntimeError("join: a query has more hits than the per-query sort supports")
    return pairs[:total]


def search_stats(tree: BottomUpTree | TopDownTree, queries: torch.Tensor) -> dict:
    """The counters the reference prints after its CPU search over every
    query (BotomUp-rtree.cu:990-1027, TopDown-rtree.cu:760-798): sum / ave /
    max of `num` (entries tested), of the search's return value (sum1) and
    of leaf hits (sum2 = total hits // 2, as printed)."""
    queries = _cuda_f32_rects(queries)
    nq = queries.shape[0]
    out = torch.zeros(6, dtype=torch.int64, device=queries.device)
    fn = lib().prt_search_stats_bu if isinstance(tree, BottomUpTree) else \
        lib().prt_search_stats_td
    check(fn(ctypes.byref(tree.st), _ptr(queries), nq, _ptr(out), _stream()))
    v = [int(x) for x in out.tolist()]
    d = max(nq, 1)
    return {"sum": v[0], "ave": v[0] / d, "max": v[1], "sum1": v[2], "ave1": v[2] / d,
            "max1": v[3], "hits": v[4], "sum2": v[4] // 2, "ave2": (v[4] // 2) / d, "max2": v[5]}


def sibling_overlap(tree: BottomUpTree | TopDownTree) -> dict:
    """sum3: pairs of overlapping entries that share a node (RtreeSearch1,
    BotomUp-rtree.cu:574-590, TopDown-rtree.cu:78-93); ave3 as the reference
    prints it: bottom-up sum3 / num_nodes / m with the reference's node
    budget num_nodes = 3 * n // m (BU:674, 1026), top-down sum3 / all slots
    (TD:797).  A zero budget gives inf (or nan for sum3 = 0), as the C
    division does."""
    out = torch.zeros(1, dtype=torch.int64, device=tree.perm.device)
    if isinstance(tree, BottomUpTree):
        check(lib().prt_sibling_overlap_bu(ctypes.byref(tree.st), _ptr(out), _stream()))
        f = tree.st.fanout
        den, m = 3 * tree.st.num_rects // f, f
    else:
        check(lib().prt_sibling_overlap_td(ctypes.byref(tree.st), _ptr(out), _stream()))
        den, m = tree.st.num_slots, 1
    s3 = int(out.item())
    ave3 = s3 / den / m if den else (math.inf if s3 else math.nan)
    return {"sum3": s3, "ave3": ave3}


@dataclass
class PolyDev:
    xy: torch.Tensor    # (V, 2) float64 CUDA
    off: torch.Tensor   # (n + 1,) int64 CUDA
    max_verts: int


def to_device_polys(p, device="cuda", dtype=torch.float64) -> PolyDev:
    return PolyDev(torch.from_numpy(p.xy).to(device=device, dtype=dtype),
                   torch.from_numpy(p.off).to(device), p.max_verts())


def area_grad(pairs: torch.Tensor, A: PolyDev, B: PolyDev, weights: torch.Tensor | None = None,
              gA: torch.Tensor | None = None, gB: torch.Tensor | None = None,
              green: bool = False):
    """Adjoint of the pair areas: returns (gA, gB[, green_area]) with
    gA[v] = sum_t weights[t] * dArea_t/dA[v] (topology fixed).  FP64 or FP32
    follows the dtype of A.xy / B.xy."""
    pairs = pairs.contiguous()
    dt = A.xy.dtype
    if B.xy.dtype != dt or dt not in (torch.float64, torch.float32):
        raise ValueError("A and B must both be float64 or both float32")
    if gA is None:
        gA = torch.zeros_like(A.xy)
    if gB is None:
        gB = torch.zeros_like(B.xy)
    ga = torch.empty(pairs.shape[0], dtype=dt, device=pairs.device) if green else None
    w = None if weights is None else weights.to(dt).contiguous()
    fn = lib().prt_overlay_area_grad if dt == torch.float64 else lib().prt_overlay_area_grad_f32
    mv = max(A.max_verts, B.max_verts, 3)
    check(fn(_ptr(pairs), pairs.shape[0], _ptr(A.xy), _ptr(A.off), _ptr(B.xy), _ptr(B.off), mv,
             _ptr(w), _ptr(gA), _ptr(gB), _ptr(ga), _stream()))
    return (gA, gB, ga) if green else (gA, gB)


def default_tune() -> TuneStruct:
    t = TuneStruct()
    check(lib().prt_tune_default(ctypes.byref(t)))
    return t


OPS = {"diff": 0, "int": 1, "xor": 2, "union": 3}  # PRT_OP_*, gpc_op order


def _op(op) -> int:
    v = OPS.get(op, op)
    if v not in (0, 1, 2, 3):
        raise ValueError(f"op must be one of {list(OPS)}")
    return v


def overlay_batched(pairs: torch.Tensor, A: PolyDev, B: PolyDev, tune: TuneStruct | None = None,
                    area: torch.Tensor | None = None, workspace: torch.Tensor | None = None,
                    stats: dict | None = None, op="int") -> torch.Tensor:
    """Same areas as overlay_plain (bit-identical): bucketed, persistent,
    work-stealing launches with the team shapes of `tune`.  With `stats`
    given, stats["overflow"] = pairs that took the plain path (syncs).
    op: "int" (A & B), "diff" (A - B), "xor", "union"."""
    pairs = pairs.contiguous()
    npairs = pairs.shape[0]
    if area is None:
        area = torch.empty(npairs, dtype=torch.float64, device=pairs.device)
    mv = A.max_verts + B.max_verts
    need = lib().prt_overlay_batched_workspace_size(npairs, mv)
    ws = workspace if workspace is not None and workspace.numel() >= need else \
        _ws(need, pairs.device)
    tp = ctypes.byref(tune) if tune is not None else None
    check(lib().prt_overlay_batched_op(_ptr(pairs), npairs, _ptr(A.xy), _ptr(A.off), _ptr(B.xy),
                                       _ptr(B.off), mv, _op(op), tp, _ptr(area), _ptr(ws),
                                       ws.numel(), _stream()))
    if stats is not None:
        ov = ctypes.c_int64(0)
        check(lib().prt_overlay_batched_overflow(_ptr(ws), ws.numel(), npairs, mv,
                                                 ctypes.byref(ov), _stream()))
        stats["overflow"] = ov.value
    return area


def overlay_plain(pairs: torch.Tensor, A: PolyDev, B: PolyDev, area: torch.Tensor | None = None,
                  workspace: torch.Tensor | None = None, op="int") -> torch.Tensor:
    """FP64 area of A[pairs[t,0]] op B[pairs[t,1]] (default: intersection),
    one lane per pair."""
    pairs = pairs.contiguous()
    npairs = pairs.shape[0]
    if area is None:
        area = torch.empty(npairs, dtype=torch.float64, device=pairs.device)
    mv = A.max_verts + B.max_verts
    need = lib().prt_overlay_plain_workspace_size(npairs, mv)
    ws = workspace if workspace is not None and workspace.numel() >= need else \
        _ws(need, pairs.device)
    check(lib().prt_overlay_plain_op(_ptr(pairs), npairs, _ptr(A.xy), _ptr(A.off), _ptr(B.xy),
                                     _ptr(B.off), mv, _op(op), _ptr(area), _ptr(ws), ws.numel(),
                                     _stream()))
    return area


@dataclass
class Contours:
    """Result contours of a pair list: pair t owns contours
    cont_off[t]..cont_off[t+1] and vertices vert_off[t]..vert_off[t+1];
    contour c starts at vertex cont_start[c] (and ends where the next one of
    its pair starts, or at its pair's end); hole[c] = 1 for a hole.  Outer
    contours run counter-clockwise, holes clockwise."""
    cont_off: torch.Tensor   # (np + 1,) int64
    vert_off: torch.Tensor   # (np + 1,) int64
    cont_start: torch.Tensor  # (nc,) int64
    hole: torch.Tensor       # (nc,) int32
    xy: torch.Tensor         # (nv, 2) float64

    _host: tuple | None = field(default=None, repr=False)

    def pair(self, t: int) -> list[tuple]:
        """[(vertices (k, 2) numpy, is_hole)] of pair t, on the host (the
        arrays are copied to the host once)."""
        if self._host is None:
            self._host = tuple(a.cpu().numpy() for a in
                               (self.cont_off, self.vert_off, self.cont_start, self.hole, self.xy))
        co, vo, cs, hole, xy = self._host
        c0, c1 = int(co[t]), int(co[t + 1])
        starts = list(cs[c0:c1]) + [int(vo[t + 1])]
        return [(xy[starts[i]:starts[i + 1]], bool(hole[c0 + i])) for i in range(c1 - c0)]


def overlay_contours(pairs: torch.Tensor, A: PolyDev, B: PolyDev, op="int",
                     pool_per_pair: int = 32, stage_limit: int = 1 << 31) -> Contours:
    """Contours of A[pairs[t,0]] op B[pairs[t,1]]: the count pass (a team of
    lanes per pair), a scan of the counts, then the write pass.  The count
    pass keeps the vertices in a pool of pool_per_pair * npairs slots (halved
    until the stage is at most stage_limit bytes; 0 keeps none), so the
    write pass sweeps again (one lane per pair) only the pairs that did not
    fit."""
    pairs = pairs.contiguous()
    npairs, dev = pairs.shape[0], pairs.device
    mv = A.max_verts + B.max_verts
    ncont = torch.zeros(npairs, dtype=torch.int64, device=dev)
    nvert = torch.zeros(npairs, dtype=torch.int64, device=dev)
    pool, stage_bytes = min(pool_per_pair * npairs, (1 << 31) - 1), 0
    while npairs and pool >= 8:
        stage_bytes = lib().prt_overlay_contours_stage_size(npairs, pool)
        if stage_bytes <= stage_limit:
            break
        pool //= 2
    stage = _ws(stage_bytes, dev) if npairs and pool >= 8 else None
    sp, sb = (_ptr(stage), stage.numel()) if stage is not None else (None, 0)
    ws = _ws(lib().prt_overlay_contours_workspace_size(npairs, mv, 0), dev)
    check(lib().prt_overlay_contours_count(_ptr(pairs), npairs, _ptr(A.xy), _ptr(A.off),
                                           _ptr(B.xy), _ptr(B.off), mv, _op(op), _ptr(ncont),
                                           _ptr(nvert), sp, sb, _ptr(ws), ws.numel(), _stream()))
    if npairs and int(torch.minimum(ncont.min(), nvert.min())) < 0:
        raise PrtError(-5, "overlay_contours: a pair exceeds max_verts")
    zero = torch.zeros(1, dtype=torch.int64, device=dev)
    cont_off = torch.cat([zero, torch.cumsum(ncont, 0)])
    vert_off = torch.cat([zero, torch.cumsum(nvert, 0)])
    nc, nv = int(cont_off[-1]), int(vert_off[-1])
    cont_start = torch.empty(nc, dtype=torch.int64, device=dev)
    hole = torch.empty(nc, dtype=torch.int32, device=dev)
    xy = torch.empty((nv, 2), dtype=torch.float64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = _ws(lib().prt_overlay_contours_workspace_size(npairs, mv, nv), dev)
    check(lib().prt_overlay_contours_write(_ptr(pairs), npairs, _ptr(A.xy), _ptr(A.off),
                                           _ptr(B.xy), _ptr(B.off), mv, _op(op), _ptr(cont_off),
                                           _ptr(vert_off), nv, _ptr(cont_start), _ptr(hole),
                                           _ptr(xy), _ptr(status), sp, sb, _ptr(ws),
                                           ws.numel(), _stream()))
    if int(status.item()) != 0:
        raise PrtError(-4, "overlay_contours: write pass disagrees with the counts")
    return Contours(cont_off, vert_off, cont_start, hole, xy)


def run_like_reference(records: torch.Tensor, offset: int, num_base: torch.Tensor,
                       prefix_base: torch.Tensor, vertex_base: torch.Tensor,
                       num_overlay: torch.Tensor, prefix_overlay: torch.Tensor,
                       vertex_overlay: torch.Tensor, zorder_index: torch.Tensor,
                       hole_base: torch.Tensor | None = None,
                       hole_overlay: torch.Tensor | None = None) -> tuple[torch.Tensor, int]:
    """prt_run: the clip stage with the reference's run() arguments
    (vector_map_overlay_gpc_kernel.cu:61).  records: (np, 2) int32
    (nodeid, queryid); num_*: int32 counts; prefix_*: int64 first-vertex
    offsets; vertex_*: (V, 2) float64; zorder_index: int32 leaf position ->
    base polygon id.  Returns (area (np,) float64, count1)."""
    records = records.contiguous()
    np_ = records.shape[0]
    area = torch.empty(np_, dtype=torch.float64, device=records.device)
    c1 = ctypes.c_int64(0)
    check(lib().prt_run(_ptr(records), offset, _ptr(num_base), _ptr(prefix_base),
                        _ptr(vertex_base), _ptr(num_overlay), _ptr(prefix_overlay),
                        _ptr(vertex_overlay), np_, _ptr(hole_base), _ptr(hole_overlay),
                        _ptr(zorder_index), _ptr(area), ctypes.byref(c1)))
    return area, c1.value
