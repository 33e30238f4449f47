"""GPU parity on degenerate geometry: shared and overlapping edges, touching
corners, repeated and collinear vertices, horizontal runs, self-crossing
polygons, a vertex on the other polygon's edge, a polygon of zero area, and
random polygons on a coarse lattice (many equal heights, coincident edges
and crossings at vertices).  For the four operations: plain and batched
areas bit-identical to the oracle, contours equal to the oracle's vertex for
vertex."""
import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree

pytestmark = pytest.mark.gpu


def sq(x0, y0, x1, y1):
    return [[x0, y0], [x1, y0], [x1, y1], [x0, y1]]


HAND = [
    (sq(0, 0, 1, 1), sq(1, 0, 2, 1)),                          # shared edge
    (sq(0, 0, 1, 1), sq(0, 0, 1, 1)),                          # identical
    (sq(0, 0, 1, 1), sq(1, 1, 2, 2)),                          # touching corners
    ([[0, 0], [0, 0], [1, 0], [1, 1], [0, 1]], sq(0.5, 0.5, 1.5, 1.5)),  # repeated vertex
    ([[0, 0], [0.5, 0], [1, 0], [1, 1], [0, 1]], sq(0, 0.5, 1, 1.5)),    # collinear vertex
    ([[0, 0], [1, 1], [1, 0], [0, 1]], sq(0, 0, 1, 1)),        # bow-tie
    ([[0, 0], [2, 0], [1, 1]], [[1, 0], [2, 1], [0, 1]]),      # vertex on an edge
    (sq(0, 0, 2, 2), sq(0.5, 0, 1.5, 2)),                      # overlapping top and bottom
    ([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], sq(0.5, 0.5, 1.5, 1.5)),  # staircase
    ([[0, 0], [1, 1], [2, 2]], sq(0, 0, 2, 2)),                # zero area
    ([[0, 0], [3, 0], [3, 3], [0, 3], [0, 0], [1, 1], [1, 2], [2, 2], [2, 1], [1, 1]],
     sq(0.5, 0.5, 2.5, 2.5)),                                  # hole through a bridge
]


def lattice_polys(seed, n, kmin=3, kmax=10, step=0.5, size=4):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(kmin, kmax + 1))
        out.append((rng.integers(0, int(size / step) + 1, size=(k, 2)) * step).tolist())
    return out


def csr(polys):
    off = np.zeros(len(polys) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in polys])
    xy = np.concatenate([np.asarray(p, float).reshape(-1, 2) for p in polys])
    return gen.PolyCSR(np.ascontiguousarray(xy), off)


def inputs():
    A = [p for p, _ in HAND] + lattice_polys(1, 600)
    B = [q for _, q in HAND] + lattice_polys(2, 600)
    rng = np.random.default_rng(3)
    h = np.arange(len(HAND))
    rp = rng.integers(len(HAND), len(A), size=(3000, 2))
    pairs = np.concatenate([np.stack([h, h], 1), rp]).astype(np.uint32)
    return pairs, csr(A), csr(B)


@pytest.mark.parametrize("op", ["int", "union", "diff", "xor"])
def test_degenerate_areas_plain_and_batched(dev, op):
    pairs, A, B = inputs()
    ref = O.overlay_op(pairs, A, B, op, nthreads=8)
    assert np.isfinite(ref).all()
    # closed forms: shared edge, identical squares, touching corners
    want = {"int": (0, 1, 0), "union": (2, 1, 2), "diff": (1, 0, 1), "xor": (2, 0, 2)}[op]
    assert tuple(ref[:3]) == want
    d = torch.from_numpy(pairs.view(np.int32)).cuda()
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    got_p = rtree.overlay_plain(d, dA, dB, op=op).cpu().numpy()
    got_b = rtree.overlay_batched(d, dA, dB, None, op=op).cpu().numpy()
    assert np.array_equal(got_p, ref), f"plain: {(got_p != ref).sum()} of {len(ref)} differ"
    assert np.array_equal(got_b, ref), f"batched: {(got_b != ref).sum()} of {len(ref)} differ"


@pytest.mark.parametrize("op", ["int", "union", "diff", "xor"])
def test_degenerate_contours(dev, op):
    pairs, A, B = inputs()
    pairs = pairs[:1000]
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    got = rtree.overlay_contours(torch.from_numpy(pairs.view(np.int32)).cuda(), dA, dB, op)
    for t, (a, b) in enumerate(pairs):
        ref, holes = O.clip_contours(A.poly(int(a)), B.poly(int(b)), op)
        mine = got.pair(t)
        assert len(mine) == len(ref), (t, op)
        for (c, hl), r, rh in zip(mine, ref, holes):
            assert hl == bool(rh) and np.array_equal(c, r), (t, op)
