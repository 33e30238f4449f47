"""Seeded synthetic inputs of the five workloads (host side, numpy arrays).

All values come from csrc/gen.cpp (counter-based SplitMix64); the seeds are
0x5EED0000 + workload number, layer B adds 0x100.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import lib

SEED_BASE = 0x5EED0000
LAYER_B = 0x100
QUERY_STREAM = 0x200
STAR_STREAM = 0x300


def seed(workload: int, layer_b: bool = False) -> int:
    return SEED_BASE + workload + (LAYER_B if layer_b else 0)


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class PolyCSR:
    """Polygons as CSR: xy[off[i]:off[i+1]] are the (x, y) vertices of i."""
    xy: np.ndarray   # (V, 2) float64
    off: np.ndarray  # (n + 1,) int64

    @property
    def n(self) -> int:
        return len(self.off) - 1

    def poly(self, i: int) -> np.ndarray:
        return self.xy[self.off[i]:self.off[i + 1]]

    def max_verts(self) -> int:
        return int(np.diff(self.off).max()) if self.n else 0


def rects_uniform(s: int, n: int, wmin: float = 1e-4, wmax: float = 1e-3) -> np.ndarray:
    r = np.empty((n, 4), np.float32)
    lib().prt_gen_rects_uniform(s, n, wmin, wmax, _p(r))
    return r


def windows(s: int, n: int, side: float = 1e-2) -> np.ndarray:
    r = np.empty((n, 4), np.float32)
    lib().prt_gen_windows(s, n, side, _p(r))
    return r


def rects_gmm(s: int, n: int, ncomp: int = 64, sigma: float = 0.01, smin: float = 1e-5,
              smax: float = 1e-4) -> np.ndarray:
    r = np.empty((n, 4), np.float32)
    lib().prt_gen_rects_gmm(s, n, ncomp, sigma, smin, smax, _p(r))
    return r


def _two_phase(fill, n: int) -> PolyCSR:
    off = np.empty(n + 1, np.int64)
    fill(off, None)
    xy = np.empty((int(off[-1]), 2), np.float64)
    fill(off, xy)
    return PolyCSR(xy, off)


def stars_in_rects(s: int, rects: np.ndarray, ids: np.ndarray, kmin: int = 8,
                   kmax: int = 64) -> PolyCSR:
    rects = np.ascontiguousarray(rects, np.float32)
    ids = np.ascontiguousarray(ids, np.uint32)
    n = len(ids)
    return _two_phase(lambda off, xy: lib().prt_gen_stars_in_rects(
        s, _p(rects), _p(ids), n, kmin, kmax, _p(off), None if xy is None else _p(xy)), n)


def stars_grid(s: int, n: int = 500_000, grid: int = 708, jitter: float = 0.25,
               rfac: float = 1.2, mu: float = 3.0, sigma: float = 0.8, kmin: int = 4,
               kmax: int = 512) -> PolyCSR:
    return _two_phase(lambda off, xy: lib().prt_gen_stars_grid(
        s, n, grid, jitter, rfac, mu, sigma, kmin, kmax, _p(off),
        None if xy is None else _p(xy)), n)


def stars_gmm(s: int, n: int, ncomp: int = 64, sigma: float = 0.01, rmin: float = 1e-5,
              rmax: float = 1e-4, mu: float = 3.0, lsigma: float = 0.8, kmin: int = 4,
              kmax: int = 512) -> PolyCSR:
    return _two_phase(lambda off, xy: lib().prt_gen_stars_gmm(
        s, n, ncomp, sigma, rmin, rmax, mu, lsigma, kmin, kmax, _p(off),
        None if xy is None else _p(xy)), n)


def star_pairs(s: int, npairs: int, kmin: int = 16, kmax: int = 128,
               off_frac: float = 0.3) -> PolyCSR:
    """Polygon 2i is P of pair i (centred at 0), 2i+1 is Q (offset centre)."""
    return _two_phase(lambda off, xy: lib().prt_gen_star_pairs(
        s, npairs, kmin, kmax, off_frac, _p(off), None if xy is None else _p(xy)), 2 * npairs)


def poly_mbr(p: PolyCSR) -> np.ndarray:
    r = np.empty((p.n, 4), np.float32)
    lib().prt_poly_mbr(_p(np.ascontiguousarray(p.xy)), _p(p.off), p.n, _p(r))
    return r


def subset(p: PolyCSR, ids: np.ndarray) -> PolyCSR:
    """Polygons ids[0], ids[1], ... of p as a new CSR."""
    cnt = np.diff(p.off)[ids]
    off = np.zeros(len(ids) + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    idx = np.repeat(p.off[ids] - off[:-1], cnt) + np.arange(off[-1])
    return PolyCSR(np.ascontiguousarray(p.xy[idx]), off)
