"""The reference's text input formats (host side, numpy arrays), backed by
csrc/io.cpp of libprt.

Rect files (BotomUp-rtree.cu:668-691, TopDown-rtree.cu:486-497): n, then n
records "x1 x2 y1 y2 key" of integers.  Polygon files
(vector_map_overlay_gpc_kernel.cu:766-818): n, then per polygon "num hole",
a bounding box of four doubles and num vertices "x y".
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .gen import PolyCSR


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def read_rects_text(path: str) -> tuple[np.ndarray, np.ndarray]:
    """(rects (n, 4) int64 as (x1, y1, x2, y2), keys (n,) int32)."""
    n = ctypes.c_int64(0)
    check(lib().prt_read_rects_text(str(path).encode(), ctypes.byref(n), None, None))
    r = np.empty((n.value, 4), np.int64)
    k = np.empty(n.value, np.int32)
    check(lib().prt_read_rects_text(str(path).encode(), ctypes.byref(n), _p(r), _p(k)))
    return r, k


def write_rects_text(path: str, rects: np.ndarray, keys: np.ndarray | None = None) -> None:
    r = np.ascontiguousarray(rects, np.int64)
    k = None if keys is None else np.ascontiguousarray(keys, np.int32)
    check(lib().prt_write_rects_text(str(path).encode(), len(r), _p(r),
                                     None if k is None else _p(k)))


def normalize(rects: np.ndarray) -> tuple[np.ndarray, int, int, int]:
    """The bottom-up driver's normalisation (BotomUp-rtree.cu:682-721):
    (shifted copy, min_x, min_y, scale)."""
    r = np.array(rects, np.int64, copy=True, order="C")
    mx, my, sc = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().prt_rects_normalize(_p(r), len(r), ctypes.byref(mx), ctypes.byref(my),
                                    ctypes.byref(sc)))
    return r, mx.value, my.value, sc.value


def rects_to_f32(rects: np.ndarray) -> np.ndarray:
    """int64 (x1, y1, x2, y2) -> float32 boxes for the builds; raises unless
    every coordinate is exact in fp32 (|v| <= 2^24), so the overlap tests on
    the fp32 boxes are those of the reference's integers."""
    r = np.asarray(rects, np.int64)
    if r.size and np.abs(r).max() > (1 << 24):
        raise ValueError("coordinates beyond 2^24 are not exact in fp32: scale them first")
    return r.astype(np.float32)


@dataclass
class PolyFile:
    polys: PolyCSR
    hole: np.ndarray      # (n,) int32
    box: np.ndarray       # (n, 4) int32, (int)(v * 100000) of the file's box
    min_xy: tuple[int, int]


def read_polys_text(path: str) -> PolyFile:
    n, nv = ctypes.c_int64(0), ctypes.c_int64(0)
    b = str(path).encode()
    check(lib().prt_read_polys_text(b, ctypes.byref(n), ctypes.byref(nv), None, None, None, None,
                                    None))
    off = np.empty(n.value + 1, np.int64)
    xy = np.empty((nv.value, 2), np.float64)
    hole = np.empty(n.value, np.int32)
    box = np.empty((n.value, 4), np.int32)
    mn = np.empty(2, np.int32)
    check(lib().prt_read_polys_text(b, ctypes.byref(n), ctypes.byref(nv), _p(off), _p(xy),
                                    _p(hole), _p(box), _p(mn)))
    return PolyFile(PolyCSR(xy, off), hole, box, (int(mn[0]), int(mn[1])))


def write_polys_text(path: str, polys: PolyCSR, hole: np.ndarray | None = None) -> None:
    h = None if hole is None else np.ascontiguousarray(hole, np.int32)
    check(lib().prt_write_polys_text(str(path).encode(), polys.n,
                                     _p(np.ascontiguousarray(polys.off, np.int64)),
                                     _p(np.ascontiguousarray(polys.xy, np.float64)),
                                     None if h is None else _p(h)))
