"""GPU: result contours (count pass by teams, scan, write pass)
equal to the oracle's, vertex for vertex, with the same hole flags and the
same contour order, for the four operations on grid-overlay pairs, random
star pairs of uneven sizes, and closed-form cases (rects, nested holes,
disjoint and empty operands).  The oracle's contours are pinned in
test_oracle_contours.py."""
import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree
from parlllelrtree_amd._lib import PrtError

pytestmark = pytest.mark.gpu


def check_all(pairs, A, B, op, **kw):
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    got = rtree.overlay_contours(torch.from_numpy(pairs.view(np.int32)).cuda(), dA, dB, op,
                                 **kw)
    assert got.cont_off.shape[0] == len(pairs) + 1
    for t, (a, b) in enumerate(pairs):
        ref, holes = O.clip_contours(A.poly(int(a)), B.poly(int(b)), op)
        mine = got.pair(t)
        assert len(mine) == len(ref), (t, op)
        for (c, h), r, rh in zip(mine, ref, holes):
            assert h == bool(rh), (t, op)
            assert np.array_equal(c, r), (t, op)
    return got


@pytest.mark.parametrize("op", ["int", "union", "diff", "xor"])
def test_grid_pairs_match_oracle(dev, op):
    A = gen.stars_grid(gen.seed(2), 3000, 55)
    B = gen.stars_grid(gen.seed(2, True), 3000, 55)
    pairs = O.join_brute(gen.poly_mbr(B), gen.poly_mbr(A))[:2500]
    got = check_all(np.ascontiguousarray(pairs), A, B, op)
    assert int(got.cont_off[-1]) > 1000


def test_uneven_star_pairs(dev):
    S = gen.star_pairs(gen.seed(3), 600, kmin=3, kmax=300)   # pair i = (2i, 2i + 1)
    pairs = np.stack([np.arange(0, 1200, 2), np.arange(1, 1200, 2)], 1).astype(np.uint32)
    for op in ("int", "xor"):
        check_all(pairs, S, S, op)


def csr(polys):
    off = np.zeros(len(polys) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in polys])
    xy = np.concatenate([np.asarray(p, float).reshape(-1, 2) for p in polys]) if len(polys) else \
        np.zeros((0, 2))
    return gen.PolyCSR(np.ascontiguousarray(xy), off)


def rect(x0, y0, x1, y1):
    return [[x0, y0], [x1, y0], [x1, y1], [x0, y1]]


def test_closed_forms_and_empty(dev):
    A = csr([rect(0, 0, 4, 4), rect(0, 0, 2, 2), rect(0, 0, 1, 1), [[0, 0], [1, 1]]])
    B = csr([rect(1, 1, 2, 2), rect(1, 1, 3, 3), rect(5, 5, 6, 6), rect(0, 0, 1, 1)])
    pairs = np.array([[0, 0], [1, 1], [2, 2], [3, 3], [0, 2]], np.uint32)
    for op in ("int", "union", "diff", "xor"):
        got = check_all(pairs, A, B, op)
    got = rtree.overlay_contours(torch.from_numpy(pairs.view(np.int32)).cuda(),
                                 rtree.to_device_polys(A), rtree.to_device_polys(B), "diff")
    (outer, h0), (hole, h1) = sorted(got.pair(0), key=lambda c: c[1])
    assert not h0 and h1
    np.testing.assert_array_equal(outer, np.array(rect(0, 0, 4, 4), float))
    assert hole.shape == (4, 2)
    empty = rtree.overlay_contours(torch.zeros((0, 2), dtype=torch.int32, device="cuda"),
                                   rtree.to_device_polys(A), rtree.to_device_polys(B))
    assert empty.xy.shape == (0, 2) and empty.cont_off.tolist() == [0]


def test_rejects_bad_arguments(dev):
    from parlllelrtree_amd._lib import lib
    A = rtree.to_device_polys(csr([rect(0, 0, 1, 1)]))
    pairs = torch.zeros((1, 2), dtype=torch.int32, device="cuda")
    n = torch.zeros(1, dtype=torch.int64, device="cuda")
    ws = torch.empty(lib().prt_overlay_contours_workspace_size(1, 8, 0), dtype=torch.uint8,
                     device="cuda")
    args = (pairs.data_ptr(), 1, A.xy.data_ptr(), A.off.data_ptr(), A.xy.data_ptr(),
            A.off.data_ptr())
    s = torch.cuda.current_stream().cuda_stream
    assert lib().prt_overlay_contours_count(*args, 8, 7, n.data_ptr(), n.data_ptr(), None, 0,
                                            ws.data_ptr(), ws.numel(), s) == -1
    assert lib().prt_overlay_contours_count(*args, 8, 1, n.data_ptr(), n.data_ptr(), None, 0,
                                            ws.data_ptr(), 16, s) == -2
    # a stage too small for one vertex per pair is refused
    assert lib().prt_overlay_contours_count(*args, 8, 1, n.data_ptr(), n.data_ptr(),
                                            ws.data_ptr(), 300, ws.data_ptr(), ws.numel(),
                                            s) == -2
    # a pair larger than max_verts is reported, not clipped
    assert lib().prt_overlay_contours_count(*args, 3, 1, n.data_ptr(), n.data_ptr(), None, 0,
                                            ws.data_ptr(), ws.numel(), s) == 0
    assert int(n.item()) == -1


def test_stage_sizes_give_the_same_contours(dev):
    """No stage; a vertex pool of 4 slots per pair (most pairs spill and are
    swept again, a few fit); half the vertices needed (some fit); 32 per pair
    (all fit); a stage limit that halves the pool: the same CSR output,
    equal to the oracle."""
    A = gen.stars_grid(gen.seed(2), 3000, 55)
    B = gen.stars_grid(gen.seed(2, True), 3000, 55)
    pairs = np.ascontiguousarray(O.join_brute(gen.poly_mbr(B), gen.poly_mbr(A))[:3000])
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    for op in ("int", "xor"):
        ref = check_all(pairs, A, B, op, pool_per_pair=0)  # no stage, against the oracle
        total = int(ref.vert_off[-1])
        for kw in ({"pool_per_pair": 4}, {"pool_per_pair": total // (2 * len(pairs))},
                   {"pool_per_pair": 32}, {"stage_limit": 300_000}):
            got = rtree.overlay_contours(pd, dA, dB, op, **kw)
            for f in ("cont_off", "vert_off", "cont_start", "hole", "xy"):
                assert torch.equal(getattr(got, f), getattr(ref, f)), (op, kw, f)


def test_write_with_a_stage_the_count_pass_did_not_fill(dev):
    """The write pass trusts a stage only when it carries the tag of the
    count pass with the same arguments, and checks every index it reads from
    it: a zeroed stage, or the count pass's stage given to a write pass with
    another op, sets the status instead of writing stale or out-of-bounds
    output."""
    from parlllelrtree_amd._lib import lib
    A = gen.stars_grid(gen.seed(2), 500, 23)
    B = gen.stars_grid(gen.seed(2, True), 500, 23)
    pairs = np.ascontiguousarray(O.join_brute(gen.poly_mbr(B), gen.poly_mbr(A))[:400])
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    np_, mv = len(pairs), dA.max_verts + dB.max_verts
    s = torch.cuda.current_stream().cuda_stream
    sb = lib().prt_overlay_contours_stage_size(np_, 64 * np_)
    assert sb > 0
    stage = torch.empty(sb, dtype=torch.uint8, device="cuda")
    ncont = torch.empty(np_, dtype=torch.int64, device="cuda")
    nvert = torch.empty(np_, dtype=torch.int64, device="cuda")
    ws = torch.empty(lib().prt_overlay_contours_workspace_size(np_, mv, 0), dtype=torch.uint8,
                     device="cuda")
    common = (pd.data_ptr(), np_, dA.xy.data_ptr(), dA.off.data_ptr(), dB.xy.data_ptr(),
              dB.off.data_ptr(), mv, 1)
    assert lib().prt_overlay_contours_count(*common, ncont.data_ptr(), nvert.data_ptr(),
                                            stage.data_ptr(), sb, ws.data_ptr(), ws.numel(),
                                            s) == 0
    zero = torch.zeros(1, dtype=torch.int64, device="cuda")
    cont_off = torch.cat([zero, torch.cumsum(ncont, 0)])
    vert_off = torch.cat([zero, torch.cumsum(nvert, 0)])
    nc, nv = int(cont_off[-1]), int(vert_off[-1])
    assert nc > 0 and nv > 0
    cs = torch.empty(nc, dtype=torch.int64, device="cuda")
    hole = torch.empty(nc, dtype=torch.int32, device="cuda")
    xy = torch.empty((nv, 2), dtype=torch.float64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws2 = torch.empty(lib().prt_overlay_contours_workspace_size(np_, mv, nv),
                      dtype=torch.uint8, device="cuda")
    # the count pass's stage; a zeroed one; the count pass's stage given to a
    # write pass with another op (its tag differs)
    for st, op, want in ((stage, 1, 0), (torch.zeros_like(stage), 1, 1), (stage, 2, 1)):
        args = common[:-1] + (op,)
        assert lib().prt_overlay_contours_write(*args, cont_off.data_ptr(),
                                                vert_off.data_ptr(), nv, cs.data_ptr(),
                                                hole.data_ptr(), xy.data_ptr(),
                                                status.data_ptr(), st.data_ptr(), sb,
                                                ws2.data_ptr(), ws2.numel(), s) == 0
        assert int(status.item()) == want, (op, want)


def test_more_pairs_than_lanes_with_and_without_stage(dev):
    """300,000 pairs: lanes take two pairs each (262,144 lanes at most), so
    the stage's per-pair records and the spill list span several pairs per
    lane.  No stage, the default pool and a pool of 2 slots per pair (most
    pairs spill) give the same output; a sample across the lanes' first and
    second pairs equals the oracle."""
    n = 300_000
    S = gen.star_pairs(gen.seed(3) + 7, n, kmin=3, kmax=12)
    pairs = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2)], 1).astype(np.uint32)
    dS = rtree.to_device_polys(S)
    pd = torch.from_numpy(pairs.view(np.int32)).cuda()
    ref = rtree.overlay_contours(pd, dS, dS, "int", pool_per_pair=0)
    assert int(ref.cont_off[-1]) > n // 2
    for kw in ({}, {"pool_per_pair": 2}):
        got = rtree.overlay_contours(pd, dS, dS, "int", **kw)
        for f in ("cont_off", "vert_off", "cont_start", "hole", "xy"):
            assert torch.equal(getattr(got, f), getattr(ref, f)), (kw, f)
    for t in list(range(0, 100)) + list(range(262_100, 262_200)) + list(range(n - 100, n)):
        want, holes = O.clip_contours(S.poly(2 * t), S.poly(2 * t + 1), "int")
        mine = ref.pair(t)
        assert len(mine) == len(want), t
        for (c, h), r, rh in zip(mine, want, holes):
            assert h == bool(rh) and np.array_equal(c, r), t
