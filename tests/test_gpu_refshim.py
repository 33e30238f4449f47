"""GPU: prt_run, the clip stage behind the reference's run() argument list
(vector_map_overlay_gpc_kernel.cu:61-113): rtree_search records of leaf
positions mapped through the bottom-up tree's Z-order permutation, polygons
as num / prefix / vertex arrays.  Areas bit-identical to the oracle for the
mapped pairs, count1 = non-empty intersections."""
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
