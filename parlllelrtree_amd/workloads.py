"""The five synthetic workloads (inputs only), shared by tests and bench.py.

  w0: bottom-up build of 1,048,576 uniform rects (centres U[0,1]^2, sides
      U[1e-4, 1e-3]), fanout 16, then 65,536 windows of side 1e-2.
  w1: two such layers (independent seeds), top-down build fanout 16,
      layer-layer join, then the first 100,000 sorted pairs clipped as star
      polygons (8..64 vertices) inscribed in the two MBRs.
  w2: two layers of 500,000 stars on a jittered 708x708 grid (jitter
      +-0.25 cell, radius 1.2 cell, lognormal(3, 0.8) vertex counts clipped
      to [4, 512]); join on the polygons' MBRs, clip every candidate pair.
  w3: 262,144 independent star pairs (16..128 vertices, unit radius, centre
      offset U[-0.3, 0.3]), area and gradient.
  w4: 16,777,216 rects from a 64-component Gaussian mixture (sigma 0.01,
      sides U[1e-5, 1e-4]); two layers of 4,000,000 mixture stars (radius
      U[1e-4, 3e-4], lognormal vertex counts; a few neighbours per polygon
      in the cluster cores, none in the tails).
"""
from __future__ import annotations

import numpy as np

from . import gen

N1M = 1 << 20


def w0(n: int = N1M, nq: int = 65536):
    return (gen.rects_uniform(gen.seed(0), n), gen.windows(gen.seed(0) + gen.QUERY_STREAM, nq))


def w1(n: int = N1M):
    return gen.rects_uniform(gen.seed(1), n), gen.rects_uniform(gen.seed(1, True), n)


def clip_inputs_from_pairs(pairs_qr: np.ndarray, ra: np.ndarray, rb: np.ndarray,
                           npairs: int = 100_000, kmin: int = 8, kmax: int = 64):
    """Star polygons for the first `npairs` (q = B id, r = A id) join pairs.
    Returns (pairs (np, 2) u32 of (A index, B index) into the two CSR sets,
    A polygons, B polygons, the A ids, the B ids)."""
    sel = np.ascontiguousarray(pairs_qr[:npairs]).astype(np.uint32)
    a_ids = np.unique(sel[:, 1])
    b_ids = np.unique(sel[:, 0])
    A = gen.stars_in_rects(gen.seed(1) + gen.STAR_STREAM, ra[a_ids], a_ids, kmin, kmax)
    B = gen.stars_in_rects(gen.seed(1, True) + gen.STAR_STREAM, rb[b_ids], b_ids, kmin, kmax)
    cp = np.stack([np.searchsorted(a_ids, sel[:, 1]), np.searchsorted(b_ids, sel[:, 0])],
                  1).astype(np.uint32)
    return cp, A, B, a_ids, b_ids


def w2(n: int = 500_000):
    A = gen.stars_grid(gen.seed(2), n)
    B = gen.stars_grid(gen.seed(2, True), n)
    return A, B


def w3(npairs: int = 262_144):
    S = gen.star_pairs(gen.seed(3), npairs)
    pairs = np.stack([np.arange(0, 2 * npairs, 2), np.arange(1, 2 * npairs, 2)],
                     1).astype(np.uint32)
    return pairs, S


def w4_rects(n: int = 1 << 24):
    return gen.rects_gmm(gen.seed(4), n)


def w4_polys(n: int = 4_000_000):
    return (gen.stars_gmm(gen.seed(4), n, rmin=1e-4, rmax=3e-4),
            gen.stars_gmm(gen.seed(4, True), n, rmin=1e-4, rmax=3e-4))
