"""GPU: prt_run, the clip stage behind the reference's run() argument list
(vector_map_overlay_gpc_kernel.cu:61-113): rtree_search records of leaf
positions mapped through the bottom-up tree's Z-order permutation, polygons
as num / prefix / vertex arrays.  Areas bit-identical to the oracle for the
mapped pairs, count1 = non-empty intersections."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree
from parlllelrtree_amd._lib import PrtError

pytestmark = pytest.mark.gpu


def ref_layout(P, dev):
    num = torch.from_numpy(np.diff(P.off).astype(np.int32)).to(dev)
    prefix = torch.from_numpy(P.off[:-1].copy()).to(dev)
    return num, prefix, torch.from_numpy(P.xy).to(dev)


def test_run_matches_oracle(dev):
    A = gen.stars_grid(gen.seed(2), 6000, 78)        # base layer (tree over its MBRs)
    B = gen.stars_grid(gen.seed(2, True), 6000, 78)  # overlay layer (queries)
    tree = rtree.build_bottom_up(torch.from_numpy(gen.poly_mbr(A)).to(dev), 16)
    qr = rtree.join(tree, torch.from_numpy(gen.poly_mbr(B)).to(dev)).cpu().numpy().view(np.uint32)
    perm = tree.perm.cpu().numpy().view(np.uint32)
    pos_of = np.empty_like(perm)
    pos_of[perm] = np.arange(len(perm), dtype=np.uint32)
    offset = 7777                                   # the tree's node-id offset of leaf 0
    rec = np.stack([pos_of[qr[:, 1]].astype(np.int64) + offset, qr[:, 0]], 1).astype(np.int32)
    nA, pA, vA = ref_layout(A, dev)
    nB, pB, vB = ref_layout(B, dev)
    area, count1 = rtree.run_like_reference(torch.from_numpy(rec).to(dev), offset, nA, pA, vA,
                                            nB, pB, vB, tree.perm)
    ab = np.ascontiguousarray(np.stack([qr[:, 1], qr[:, 0]], 1))
    ref = O.overlay(ab, A, B, nthreads=8)
    got = area.cpu().numpy()
    assert np.array_equal(got, ref), f"{(got != ref).sum()} of {len(ref)} differ"
    assert count1 == int((ref > 0).sum())


def test_run_errors_and_empty(dev):
    A = gen.stars_grid(gen.seed(2), 200, 15)
    nA, pA, vA = ref_layout(A, dev)
    z = torch.arange(A.n, dtype=torch.int32, device=dev)
    rec = torch.tensor([[5, 1], [3, 2]], dtype=torch.int32, device=dev)
    # nodeid below the offset
    with pytest.raises(PrtError):
        rtree.run_like_reference(rec, 4, nA, pA, vA, nA, pA, vA, z)
    # prefix that is not the running sum of num
    bad = pA.clone()
    bad[2] += 1
    with pytest.raises(PrtError):
        rtree.run_like_reference(rec, 0, nA, bad, vA, nA, pA, vA, z)
    area, c1 = rtree.run_like_reference(rec[:0], 0, nA, pA, vA, nA, pA, vA, z)
    assert area.numel() == 0 and c1 == 0
    area, c1 = rtree.run_like_reference(rec, 0, nA, pA, vA, nA, pA, vA, z)
    ref = O.overlay(np.array([[5, 1], [3, 2]], np.uint32), A, A)
    assert np.array_equal(area.cpu().numpy(), ref) and c1 == int((ref > 0).sum())


def test_read_data_matches_the_readers(dev, tmp_path):
    """prt_read_data (readData's argument list, OV:752-888): host arrays equal
    to the text readers' output, vertices in device memory, minX / minY only
    lowered by the base layer."""
    from parlllelrtree_amd import refio
    from parlllelrtree_amd._lib import lib
    A = gen.stars_grid(gen.seed(5), 400, 21)
    B = gen.stars_grid(gen.seed(5, True), 300, 18)
    pa, pb = tmp_path / "base.txt", tmp_path / "overlay.txt"
    refio.write_polys_text(pa, A, np.arange(A.n, dtype=np.int32) % 2)
    refio.write_polys_text(pb, B)
    ra, rb = refio.read_polys_text(pa), refio.read_polys_text(pb)
    i32p = lambda: ctypes.POINTER(ctypes.c_int32)()
    r = [i32p() for _ in range(4)]
    rq = [i32p() for _ in range(4)]
    nb, no, hb, ho = i32p(), i32p(), i32p(), i32p()
    pfb, pfo = ctypes.POINTER(ctypes.c_int64)(), ctypes.POINTER(ctypes.c_int64)()
    vb, vo = ctypes.c_void_p(), ctypes.c_void_p()
    nc, nq = ctypes.c_int32(0), ctypes.c_int32(0)
    mx, my = ctypes.c_int32(10**9), ctypes.c_int32(-10**9)
    by = ctypes.byref
    rc = lib().prt_read_data(str(pa).encode(), str(pb).encode(), *[by(x) for x in r], by(nc),
                             *[by(x) for x in rq], by(nq), by(nb), by(pfb), by(vb), by(no), by(pfo),
                             by(vo), by(mx), by(my), by(hb), by(ho))
    assert rc == 0
    hip = ctypes.CDLL("libamdhip64.so")
    try:
        assert (nc.value, nq.value) == (A.n, B.n)
        for L, f, rr, num, pf, v, hole in ((A, ra, r, nb, pfb, vb, hb), (B, rb, rq, no, pfo, vo, ho)):
            for j in range(4):
                assert np.array_equal(np.ctypeslib.as_array(rr[j], (L.n,)), f.box[:, j])
            assert np.array_equal(np.ctypeslib.as_array(num, (L.n,)), np.diff(L.off))
            assert np.array_equal(np.ctypeslib.as_array(pf, (L.n,)), L.off[:-1])
            assert np.array_equal(np.ctypeslib.as_array(hole, (L.n,)), f.hole)
            host = np.empty_like(L.xy)
            assert hip.hipMemcpy(ctypes.c_void_p(host.ctypes.data), v, host.nbytes, 2) == 0
            assert np.array_equal(host, L.xy)
        assert mx.value == ra.min_xy[0]      # lowered
        assert my.value == -10**9            # already smaller: kept
    finally:
        libc = ctypes.CDLL(None)
        libc.free.argtypes = [ctypes.c_void_p]
        for x in r + rq + [nb, no, hb, ho, pfb, pfo]:
            libc.free(ctypes.cast(x, ctypes.c_void_p))
        hip.hipFree(vb)
        hip.hipFree(vo)
