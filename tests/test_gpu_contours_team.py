"""GPU: the contour count pass by teams (a team of lanes per pair, the
pair's tables, interval lists and open contours in LDS) against the
one-lane-per-pair kernel and the oracle: the same CSR output (offsets,
contour starts, hole flags, vertices) for the four operations, for every
team shape, with and without a stage, on grid pairs, uneven stars,
crossing-dense sawtooths (beams the team sweeps, and beams beyond the team's
scratch that one lane sweeps) and combs whose inside intervals outnumber the
team's LDS arrays (those pairs are passed to the one-lane kernel)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree
from parlllelrtree_amd._lib import TuneStruct, lib

pytestmark = pytest.mark.gpu
FIELDS = ("cont_off", "vert_off", "cont_start", "hole", "xy")


@pytest.fixture
def hook(dev):
    """prt_debug_contours, reset to the default team path afterwards."""
    def set_(one_lane=False, tune=None):
        ref = ctypes.byref(tune) if tune is not None else None
        assert lib().prt_debug_contours(1 if one_lane else 0, ref) == 0
    yield set_
    set_()


def csr(polys):
    off = np.zeros(len(polys) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in polys])
    xy = np.concatenate([np.asarray(p, float).reshape(-1, 2) for p in polys])
    return gen.PolyCSR(np.ascontiguousarray(xy), off)


def contours(pairs, A, B, op, **kw):
    return rtree.overlay_contours(torch.from_numpy(pairs.view(np.int32)).cuda(),
                                  rtree.to_device_polys(A), rtree.to_device_polys(B), op, **kw)


def assert_same(got, ref, what):
    for f in FIELDS:
        assert torch.equal(getattr(got, f), getattr(ref, f)), (what, f)


def assert_oracle(got, pairs, A, B, op, sample):
    for t in sample:
        a, b = pairs[t]
        want, holes = O.clip_contours(A.poly(int(a)), B.poly(int(b)), op)
        mine = got.pair(int(t))
        assert len(mine) == len(want), (t, op)
        for (c, h), r, rh in zip(mine, want, holes):
            assert h == bool(rh) and np.array_equal(c, r), (t, op)


def grid_pairs(n=3000, side=55, count=3000):
    A = gen.stars_grid(gen.seed(2), n, side)
    B = gen.stars_grid(gen.seed(2, True), n, side)
    return np.ascontiguousarray(O.join_brute(gen.poly_mbr(B), gen.poly_mbr(A))[:count]), A, B


@pytest.mark.parametrize("op", ["int", "union", "diff", "xor"])
def test_team_equals_one_lane_and_oracle(hook, op):
    pairs, A, B = grid_pairs()
    S = gen.star_pairs(gen.seed(3), 400, kmin=3, kmax=300)  # up to 600 vertices a pair
    sp = np.stack([np.arange(0, 800, 2), np.arange(1, 800, 2)], 1).astype(np.uint32)
    for (pr, P, Q) in ((pairs, A, B), (sp, S, S)):
        for kw in ({"pool_per_pair": 0}, {}):
            hook(one_lane=True)
            ref = contours(pr, P, Q, op, **kw)
            hook()
            got = contours(pr, P, Q, op, **kw)
            assert_same(got, ref, (op, kw))
        assert_oracle(got, pr, P, Q, op, range(0, len(pr), 7))


@pytest.mark.parametrize("lanes,waves,factor", [(8, 4, 1), (16, 2, 2), (32, 2, 1), (64, 1, 8),
                                                (64, 4, 2)])
def test_every_team_shape(hook, lanes, waves, factor):
    pairs, A, B = grid_pairs(count=1500)
    hook(one_lane=True)
    ref = contours(pairs, A, B, "xor")
    t = TuneStruct()
    assert lib().prt_tune_default(ctypes.byref(t)) == 0
    for b in range(6):
        t.lanes_per_pair[b], t.waves_per_block[b], t.list_factor[b] = lanes, waves, factor
    hook(tune=t)
    assert_same(contours(pairs, A, B, "xor"), ref, (lanes, waves, factor))
    assert_same(contours(pairs, A, B, "xor", pool_per_pair=0), ref, (lanes, waves, factor))


def zigzag(k, lean, y0=0.0, y1=1.0, x0=0.0):
    """Sawtooth of k teeth whose edges all span the band (y0, y1), leaning by
    `lean`, closed below: the band is one beam holding 2k long edges."""
    v = []
    for i in range(k):
        v.append((x0 + i, y0))
        v.append((x0 + i + 0.5 + lean, y1))
    v += [(x0 + k, y0), (x0 + k, y0 - 1.0), (x0, y0 - 1.0)]
    return v


@pytest.mark.parametrize("k", [6, 12, 20, 40, 90])
def test_crossing_dense_sawtooths(hook, k):
    """Beams of 2k crossing edges: swept by the team (up to 64 entries and
    crossing heights), else by lane 0 alone; equal to the one-lane kernel
    and the oracle."""
    polys_a, polys_b = [], []
    for j, lean in enumerate((0.3, 1.5, 3.7)):
        polys_a.append(zigzag(k, lean, x0=0.1 * j))
        polys_b.append(zigzag(k, -lean, y0=0.05, y1=1.1, x0=0.3))
    A, B = csr(polys_a), csr(polys_b)
    pairs = np.array([(i, j) for i in range(3) for j in range(3)], np.uint32)
    for op in ("int", "xor"):
        hook(one_lane=True)
        ref = contours(pairs, A, B, op)
        hook()
        got = contours(pairs, A, B, op)
        assert_same(got, ref, (k, op))
        assert_oracle(got, pairs, A, B, op, range(len(pairs)))


def comb(k):
    """A base bar y in [-1, 0] with k teeth x in [2i, 2i+1], y in [0, 3]."""
    v = [(0.0, -1.0), (2.0 * k - 1, -1.0)]
    for i in reversed(range(k)):
        v += [(2.0 * i + 1, 3.0), (2.0 * i, 3.0)]
        if i > 0:
            v += [(2.0 * i, 0.0), (2.0 * i - 1, 0.0)]
    return v


@pytest.mark.parametrize("k", [30, 80, 140])
def test_combs_beyond_the_team_arrays(hook, k):
    """A comb against a band across its teeth: k inside intervals in one
    sub-beam.  Up to 33 the team keeps them; beyond, the pair takes the
    one-lane kernel.  Both give the oracle's contours."""
    band = [(-1.0, 1.0), (2.0 * k, 1.0), (2.0 * k, 2.0), (-1.0, 2.0)]
    A, B = csr([comb(k), comb(k)]), csr([band, [(0.5, -0.5), (2.0 * k, -0.5), (2.0 * k, 2.5)]])
    pairs = np.array([(0, 0), (1, 1), (0, 1)], np.uint32)
    for op in ("int", "union", "diff", "xor"):
        got = contours(pairs, A, B, op)
        assert_oracle(got, pairs, A, B, op, range(len(pairs)))
        hook(one_lane=True)
        assert_same(contours(pairs, A, B, op), got, (k, op))
        hook()
    assert len(contours(pairs[:1], A, B, "int").pair(0)) == k
