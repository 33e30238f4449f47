"""GPU: the overlay pipeline (polygon MBRs -> R-tree over layer B -> join
with layer A -> batched FP64 clip -> total area / non-empty count), run
eagerly and replayed from a captured hipGraph.

Bars: MBRs bit-exact and provably enclosing (each fp32 bound within one ulp
outside the FP64 extreme); the pair list equal to the brute-force pair set of
the MBRs in (A id, B id) order; areas bit-identical to the CPU oracle; every
graph replay identical to the eager run, with unrelated eager work between
the replays and the outputs overwritten before each one."""
import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree
from parlllelrtree_amd._lib import PrtError, lib
from parlllelrtree_amd.pipeline import OverlayPipeline

pytestmark = pytest.mark.gpu


def grid_layers(n, grid, shift=0):
    return (gen.stars_grid(gen.seed(2) + shift, n, grid),
            gen.stars_grid(gen.seed(2, True) + shift, n, grid))


def device_mbr(P):
    d = rtree.to_device_polys(P)
    out = torch.empty((P.n, 4), dtype=torch.float32, device="cuda")
    assert lib().prt_poly_mbr_dev(d.xy.data_ptr(), d.off.data_ptr(), P.n, out.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream) == 0
    return out.cpu().numpy()


def test_poly_mbr_bit_exact_and_enclosing(dev):
    A, _ = grid_layers(5000, 71)
    got = device_mbr(A)
    np.testing.assert_array_equal(got.view(np.uint32), gen.poly_mbr(A).view(np.uint32))
    seg = A.off[:-1]
    lo = np.minimum.reduceat(A.xy, seg, axis=0)
    hi = np.maximum.reduceat(A.xy, seg, axis=0)
    g = got.astype(np.float64)
    assert (g[:, 0] <= lo[:, 0]).all() and (g[:, 1] <= lo[:, 1]).all()
    assert (g[:, 2] >= hi[:, 0]).all() and (g[:, 3] >= hi[:, 1]).all()
    # tight: one ulp inward would cut the polygon
    assert (np.nextafter(got[:, 0], np.float32(np.inf)).astype(np.float64) > lo[:, 0]).all()
    assert (np.nextafter(got[:, 2], np.float32(-np.inf)).astype(np.float64) < hi[:, 0]).all()


def expected(A, B):
    pairs = O.join_brute(gen.poly_mbr(B), gen.poly_mbr(A))  # (A id, B id)
    return pairs, O.overlay(pairs, A, B, nthreads=8)


@pytest.mark.parametrize("tree", ["bu", "td"])
def test_pipeline_eager_matches_oracle(dev, tree):
    A, B = grid_layers(4000, 64)
    ref_pairs, ref_area = expected(A, B)
    p = OverlayPipeline(rtree.to_device_polys(A), rtree.to_device_polys(B),
                        cap=len(ref_pairs) + 100, tree=tree)
    p.area.fill_(float("nan"))
    p.run()
    pairs, area, total, nonempty = p.result()
    np.testing.assert_array_equal(pairs.cpu().numpy().view(np.uint32), ref_pairs)
    np.testing.assert_array_equal(area.cpu().numpy(), ref_area)
    assert nonempty == int((ref_area > 0).sum())
    assert total == pytest.approx(float(ref_area.sum()), rel=1e-12)


def snapshot(p):
    pairs, area, total, nonempty = p.result()
    return pairs.clone(), area.clone(), total, nonempty


def assert_same(a, b):
    assert torch.equal(a[0], b[0])
    assert torch.equal(a[1].view(torch.int64), b[1].view(torch.int64))  # bitwise
    assert a[2] == b[2] and a[3] == b[3]


def test_graph_replays_match_eager(dev):
    A, B = grid_layers(20000, 142)
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    p = OverlayPipeline(dA, dB, cap=40 * 20000)
    p.run()
    eager = snapshot(p)
    assert eager[0].shape[0] > 100000
    p.capture()
    # unrelated eager work on the same device between replays
    C, D = grid_layers(3000, 55, shift=7)
    cp = O.join_brute(gen.poly_mbr(D), gen.poly_mbr(C))[:5000]
    cpd = torch.from_numpy(cp.view(np.int32)).cuda()
    dC, dD = rtree.to_device_polys(C), rtree.to_device_polys(D)
    for rep in range(3):
        side = rtree.overlay_batched(cpd, dC, dD)
        x = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
        _ = float((x * x).sum())
        p.pairs.fill_(-1)
        p.area.fill_(float("nan"))
        p.count.fill_(-7)
        p.total.fill_(float("nan"))
        p.nonempty.fill_(-7)
        p.replay()
        got = snapshot(p)
        assert_same(got, eager)
        torch.cuda.synchronize()
        assert torch.isfinite(side).all(), rep
    # the graph reads its inputs at replay time: move layer A, replay, and
    # compare with a fresh eager pipeline on the moved polygons
    dA.xy.add_(torch.tensor([0.3 / 142, -0.2 / 142], dtype=torch.float64, device="cuda"))
    p.replay()
    moved = snapshot(p)
    q = OverlayPipeline(dA, dB, cap=40 * 20000)
    q.run()
    assert_same(moved, snapshot(q))
    assert not torch.equal(moved[1][:1000], eager[1][:1000])
    # oracle on a sample of the moved result
    A2 = gen.PolyCSR(dA.xy.cpu().numpy(), A.off)
    sel = np.random.default_rng(5).choice(moved[0].shape[0], 20000, replace=False)
    sp = moved[0].cpu().numpy().view(np.uint32)[sel]
    np.testing.assert_array_equal(moved[1].cpu().numpy()[sel], O.overlay(sp, A2, B, nthreads=8))
    p.close()


def test_capacity_overflow_reported(dev):
    A, B = grid_layers(2000, 45)
    ref_pairs, ref_area = expected(A, B)
    p = OverlayPipeline(rtree.to_device_polys(A), rtree.to_device_polys(B), cap=100)
    p.capture()
    p.replay()
    with pytest.raises(PrtError):
        p.result()
    assert int(p.count[0]) == len(ref_pairs)
    kept = p.pairs.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(p.area.cpu().numpy(), O.overlay(kept, A, B))
    p.close()


@pytest.mark.parametrize("which", ["A", "B"])
def test_empty_layer(dev, which):
    A, B = grid_layers(500, 23)
    E = gen.PolyCSR(np.zeros((0, 2)), np.zeros(1, np.int64))
    dA = rtree.to_device_polys(E if which == "A" else A)
    dB = rtree.to_device_polys(E if which == "B" else B)
    p = OverlayPipeline(dA, dB, cap=1000)
    p.count.fill_(5)
    p.capture()
    p.replay()
    pairs, area, total, nonempty = p.result()
    assert pairs.shape[0] == 0 and total == 0.0 and nonempty == 0
    p.close()


@pytest.mark.parametrize("tree", ["bu", "td"])
def test_graph_on_lognormal_grid_subset(dev, tree):
    """The captured pipeline on the grid overlay's own polygons: the first
    60,000 stars of both 500,000-star layers (rows 0-84 of the jittered
    708x708 grid, lognormal vertex counts up to 512, the same polygons as the
    full workload's).  Three replays with eager work between them, each equal
    to the eager run; pairs equal to the oracle's join of the MBRs, every
    area bit-identical to the oracle."""
    A, B = grid_layers(60000, 708)
    assert A.max_verts() > 128 and B.max_verts() > 128
    ot = O.td_build(gen.poly_mbr(B), 16)
    ref_pairs, _ = O.join_td(ot, gen.poly_mbr(A))           # (A id, B id)
    ref_area = O.overlay(ref_pairs, A, B, nthreads=16)
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    p = OverlayPipeline(dA, dB, cap=len(ref_pairs) + 4096, tree=tree)
    p.run()
    eager = snapshot(p)
    np.testing.assert_array_equal(eager[0].cpu().numpy().view(np.uint32), ref_pairs)
    np.testing.assert_array_equal(eager[1].cpu().numpy(), ref_area)
    assert eager[3] == int((ref_area > 0).sum())
    p.capture()
    for rep in range(3):
        x = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
        _ = float((x * x).sum())
        p.pairs.fill_(-1)
        p.area.fill_(float("nan"))
        p.count.fill_(-7)
        p.total.fill_(float("nan"))
        p.nonempty.fill_(-7)
        p.replay()
        assert_same(snapshot(p), eager)
    p.close()
