"""numpy wrapper of the CPU oracle (liboracle.so, built from prt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_L = None

_SIG = {
    "orc_quant15": (c_uint32, [c_float, c_float]),
    "orc_spread15": (c_uint32, [c_uint32]),
    "orc_morton_center30": (None, [c_void_p, c_int64, c_void_p]),
    "orc_interleave_ref4": (c_uint64, [c_uint64, c_uint64, c_uint64, c_uint64]),
    "orc_morton_ref4d": (None, [c_void_p, c_int64, c_int64, c_void_p]),
    "orc_ref_scale": (c_int64, [c_int64]),
    "orc_stable_order_u64": (None, [c_void_p, c_int64, c_void_p]),
    "orc_stable_order_u32": (None, [c_void_p, c_int64, c_void_p]),
    "orc_bu_levels": (c_int, [c_int64, c_int, c_void_p, c_void_p]),
    "orc_bu_build": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p]),
    "orc_bu_build_keys64": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_td_levels": (c_int, [c_int64, c_int, c_void_p]),
    "orc_td_build": (c_int, [c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p]),
    "orc_overlap": (c_int, [c_void_p, c_void_p]),
    "orc_join_bu": (c_int64, [c_int64, c_int, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_int64, c_void_p, c_int64, POINTER(c_int64)]),
    "orc_join_td": (c_int64, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                              c_int64, POINTER(c_int64)]),
    "orc_join_brute": (c_int64, [c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64]),
    "orc_search_stats_bu": (None, [c_int64, c_int, c_int64, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "orc_search_stats_td": (None, [c_int, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                   c_void_p, c_void_p]),
    "orc_sibling_overlap_bu": (c_int64, [c_int64, c_int, c_int64, c_void_p, c_void_p, c_void_p]),
    "orc_sibling_overlap_td": (c_int64, [c_int, c_void_p, c_void_p]),
    "orc_clip_area": (c_double, [c_void_p, c_int, c_void_p, c_int, POINTER(c_uint64)]),
    "orc_clip_area_op": (c_double, [c_void_p, c_int, c_void_p, c_int, c_int, POINTER(c_uint64)]),
    "orc_overlay_op": (None, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                              c_void_p, POINTER(c_uint64), c_int]),
    "orc_clip_contours": (None, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_int64,
                                 c_void_p, c_void_p, c_int64, POINTER(c_int64), POINTER(c_int64)]),
    "orc_poly_area": (c_double, [c_void_p, c_int]),
    "orc_green_area": (c_double, [c_void_p, c_int, c_void_p, c_int]),
    "orc_sh_convex_area": (c_double, [c_void_p, c_int, c_void_p, c_int]),
    "orc_area_grad": (None, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "orc_overlay": (None, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                           POINTER(c_uint64), c_int]),
}


def lib():
    global _L
    if _L is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(str(LIB_PATH))
        for k, (r, a) in _SIG.items():
            f = getattr(h, k)
            f.restype, f.argtypes = r, a
        _L = h
    return _L


def _p(a):
    return None if a is None else c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- keys
def morton_center30(r4):
    r4 = _c(r4, np.float32)
    k = np.empty(len(r4), np.uint32)
    lib().orc_morton_center30(_p(r4), len(r4), _p(k))
    return k


def interleave_ref4(a1, a2, a3, a4):
    return int(lib().orc_interleave_ref4(a1, a2, a3, a4))


def morton_ref4d(r4_i64, scale):
    r = _c(r4_i64, np.int64)
    k = np.empty(len(r), np.uint64)
    lib().orc_morton_ref4d(_p(r), len(r), scale, _p(k))
    return k


def ref_scale(rng):
    return int(lib().orc_ref_scale(rng))


def stable_order(keys):
    keys = np.ascontiguousarray(keys)
    perm = np.empty(len(keys), np.uint32)
    if keys.dtype == np.uint64:
        lib().orc_stable_order_u64(_p(keys), len(keys), _p(perm))
    else:
        k = _c(keys, np.uint32)
        lib().orc_stable_order_u32(_p(k), len(k), _p(perm))
    return perm


# ---------------------------------------------------------------- builds
def bu_levels(n, f):
    ls = np.zeros(64, np.int64)
    lc = np.zeros(64, np.int64)
    L = lib().orc_bu_levels(n, f, _p(ls), _p(lc))
    return [(int(ls[i]), int(lc[i])) for i in range(L)]


def bu_build(r4, f, keys64=None):
    """Returns dict: perm, keys (sorted), node_count/start/end, entry_mbr
    (nodes*f, 4), entry_ref, levels.  Leaf order: CENTER30 keys, or the
    given 64-bit keys (the reference's REF4D64 order)."""
    r4 = _c(r4, np.float32)
    n = len(r4)
    lv = bu_levels(n, f)
    nn = lv[-1][0] + lv[-1][1]
    keys = morton_center30(r4) if keys64 is None else _c(keys64, np.uint64)
    perm = np.empty(n, np.uint32)
    cnt = np.empty(nn, np.int32)
    st = np.empty(nn, np.int32)
    en = np.empty(nn, np.int32)
    mbr = np.empty((nn * f, 4), np.float32)
    ref = np.empty(nn * f, np.int32)
    fn = lib().orc_bu_build if keys64 is None else lib().orc_bu_build_keys64
    got = fn(_p(r4), n, f, _p(keys), _p(perm), _p(cnt), _p(st), _p(en), _p(mbr), _p(ref))
    assert got == nn
    return dict(perm=perm, keys=keys[perm], node_count=cnt, node_start=st, node_end=en,
                entry_mbr=mbr, entry_ref=ref, levels=lv, num_nodes=nn, fanout=f)


def td_levels(n, f):
    off = np.zeros(65, np.int64)
    L = lib().orc_td_levels(n, f, _p(off))
    return [int(v) for v in off[:L + 1]]


def td_build(r4, f):
    r4 = _c(r4, np.float32)
    n = len(r4)
    off = td_levels(n, f)
    ns = off[-1]
    child = np.empty(ns, np.int32)
    mbr = np.empty((ns, 4), np.float32)
    perm = np.empty(n, np.uint32)
    got = lib().orc_td_build(_p(r4), n, f, _p(child), _p(mbr), _p(perm))
    assert got == ns
    return dict(slot_child=child, slot_mbr=mbr, perm=perm, level_offset=off, fanout=f)


# ---------------------------------------------------------------- joins
def _pairs_out(fn, cap_guess):
    cap = max(cap_guess, 16)
    while True:
        pairs = np.empty((cap, 2), np.uint32)
        total, extra = fn(pairs, cap)
        if total <= cap:
            return pairs[:total], extra
        cap = total


def join_bu(tree, q4):
    q4 = _c(q4, np.float32)
    f = tree["fanout"]

    def run(pairs, cap):
        t = c_int64(0)
        tot = lib().orc_join_bu(tree["num_nodes"], f, tree["levels"][0][1], _p(tree["node_count"]),
                                _p(tree["entry_mbr"]), _p(tree["entry_ref"]), _p(q4), len(q4),
                                _p(pairs), cap, ctypes.byref(t))
        return tot, t.value
    return _pairs_out(run, 64 * len(q4))


def join_td(tree, q4):
    q4 = _c(q4, np.float32)

    def run(pairs, cap):
        t = c_int64(0)
        tot = lib().orc_join_td(tree["fanout"], _p(tree["slot_child"]), _p(tree["slot_mbr"]),
                                _p(tree["perm"]), _p(q4), len(q4), _p(pairs), cap, ctypes.byref(t))
        return tot, t.value
    return _pairs_out(run, 64 * len(q4))


def join_brute(r4, q4):
    r4 = _c(r4, np.float32)
    q4 = _c(q4, np.float32)

    def run(pairs, cap):
        return lib().orc_join_brute(_p(r4), len(r4), _p(q4), len(q4), _p(pairs), cap), 0
    return _pairs_out(run, 64 * len(q4))[0]


def search_stats(tree, q4):
    """Per-query (num, count, num1) of the reference's CPU search over a
    bu_build or td_build tree dict."""
    q4 = _c(q4, np.float32)
    nq = len(q4)
    num, cnt, num1 = (np.zeros(nq, np.int64) for _ in range(3))
    if "slot_child" in tree:
        lib().orc_search_stats_td(tree["fanout"], _p(tree["slot_child"]), _p(tree["slot_mbr"]),
                                  _p(q4), nq, _p(num), _p(cnt), _p(num1))
    else:
        lib().orc_search_stats_bu(tree["num_nodes"], tree["fanout"], tree["levels"][0][1],
                                  _p(tree["node_count"]), _p(tree["entry_mbr"]),
                                  _p(tree["entry_ref"]), _p(q4), nq, _p(num), _p(cnt), _p(num1))
    return num, cnt, num1


def sibling_overlap(tree):
    if "slot_child" in tree:
        return int(lib().orc_sibling_overlap_td(tree["fanout"], _p(tree["slot_child"]),
                                                _p(tree["slot_mbr"])))
    return int(lib().orc_sibling_overlap_bu(tree["num_nodes"], tree["fanout"],
                                            tree["levels"][0][1], _p(tree["node_count"]),
                                            _p(tree["entry_mbr"]), _p(tree["entry_ref"])))


# ---------------------------------------------------------------- polygons
def clip_area(P, Q, with_flops=False):
    P = _c(P, np.float64)
    Q = _c(Q, np.float64)
    fl = c_uint64(0)
    a = lib().orc_clip_area(_p(P), len(P), _p(Q), len(Q), ctypes.byref(fl))
    return (a, fl.value) if with_flops else a


OPS = {"diff": 0, "int": 1, "xor": 2, "union": 3}


def clip_area_op(P, Q, op):
    """Area of P op Q, op in OPS (names or gpc_op numbers)."""
    P = _c(P, np.float64)
    Q = _c(Q, np.float64)
    return lib().orc_clip_area_op(_p(P), len(P), _p(Q), len(Q), OPS.get(op, op), None)


def clip_contours(P, Q, op="int"):
    """Contours of P op Q: list of (k, 2) vertex arrays (outer counter-
    clockwise, holes clockwise) and the hole flags, in closing order."""
    P = _c(P, np.float64)
    Q = _c(Q, np.float64)
    nv, nc = c_int64(0), c_int64(0)
    o = OPS.get(op, op)
    lib().orc_clip_contours(_p(P), len(P), _p(Q), len(Q), o, None, 0, None, None, 0,
                            ctypes.byref(nv), ctypes.byref(nc))
    xy = np.empty((nv.value, 2), np.float64)
    cnt = np.empty(nc.value, np.int32)
    hole = np.empty(nc.value, np.int32)
    lib().orc_clip_contours(_p(P), len(P), _p(Q), len(Q), o, _p(xy), nv.value, _p(cnt), _p(hole),
                            nc.value, ctypes.byref(nv), ctypes.byref(nc))
    cuts = np.cumsum(cnt)[:-1] if len(cnt) else []
    return [c for c in np.split(xy, cuts)] if len(cnt) else [], hole.astype(bool)


def overlay_op(pairs, A, B, op, nthreads=1):
    pairs = _c(pairs, np.uint32)
    area = np.empty(len(pairs), np.float64)
    lib().orc_overlay_op(_p(pairs), len(pairs), _p(_c(A.xy, np.float64)), _p(_c(A.off, np.int64)),
                         _p(_c(B.xy, np.float64)), _p(_c(B.off, np.int64)), OPS.get(op, op),
                         _p(area), None, nthreads)
    return area


def poly_area(P):
    P = _c(P, np.float64)
    return lib().orc_poly_area(_p(P), len(P))


def green_area(P, Q):
    P = _c(P, np.float64)
    Q = _c(Q, np.float64)
    return lib().orc_green_area(_p(P), len(P), _p(Q), len(Q))


def sh_convex_area(P, Q):
    P = _c(P, np.float64)
    Q = _c(Q, np.float64)
    return lib().orc_sh_convex_area(_p(P), len(P), _p(Q), len(Q))


def area_grad(P, Q):
    P = _c(P, np.float64)
    Q = _c(Q, np.float64)
    gP = np.empty_like(P)
    gQ = np.empty_like(Q)
    lib().orc_area_grad(_p(P), len(P), _p(Q), len(Q), _p(gP), _p(gQ))
    return gP, gQ


def overlay(pairs, A, B, nthreads=1, with_flops=False):
    """pairs (np, 2) u32 into CSR polygon sets A (base) and B (overlay)."""
    pairs = _c(pairs, np.uint32)
    area = np.empty(len(pairs), np.float64)
    fl = c_uint64(0)
    lib().orc_overlay(_p(pairs), len(pairs), _p(_c(A.xy, np.float64)), _p(_c(A.off, np.int64)),
                      _p(_c(B.xy, np.float64)), _p(_c(B.off, np.int64)), _p(area),
                      ctypes.byref(fl), nthreads)
    return (area, fl.value) if with_flops else area
