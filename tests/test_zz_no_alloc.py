"""The join with no allocation (prt_join_bu_noalloc: workspace in the
library's fixed device array) gives exactly what prt_join_bu gives with a
caller workspace, on two calls in a row (the second smaller, so stale
contents of the shared array from the first call are in place), and rejects
query counts beyond its maximum before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from parlllelrtree_amd import gen, rtree
from parlllelrtree_amd._lib import PrtError, check, lib

pytestmark = pytest.mark.gpu


def _noalloc(tree, q, cap):
    pairs = torch.full((max(cap, 1), 2), -1, dtype=torch.int32, device=q.device)
    cnt = torch.full((2,), -7, dtype=torch.int64, device=q.device)
    check(lib().prt_join_bu_noalloc(ctypes.byref(tree.st), q.data_ptr(), q.shape[0],
                                    pairs.data_ptr(), cap, cnt.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream))
    return pairs, cnt


def _with_ws(tree, q, cap):
    pairs = torch.full((max(cap, 1), 2), -1, dtype=torch.int32, device=q.device)
    cnt = torch.full((2,), -7, dtype=torch.int64, device=q.device)
    ws = torch.empty(lib().prt_join_workspace_size(q.shape[0]), dtype=torch.uint8,
                     device=q.device)
    check(lib().prt_join_bu(ctypes.byref(tree.st), q.data_ptr(), q.shape[0], pairs.data_ptr(),
                            cap, cnt.data_ptr(), ws.data_ptr(), ws.numel(),
                            torch.cuda.current_stream().cuda_stream))
    return pairs, cnt


def test_join_bu_noalloc_matches_join_bu(dev):
    r = gen.rects_uniform(gen.seed(0), 1 << 20)
    tree = rtree.build_bottom_up(torch.from_numpy(r).to(dev), 16)
    for nq, side in ((65536, 1e-2), (3000, 3e-2)):
        q = torch.from_numpy(gen.windows(gen.seed(0) + gen.QUERY_STREAM + nq, nq, side)).to(dev)
        want_pairs, want_cnt = _with_ws(tree, q, 8_000_000)
        got_pairs, got_cnt = _noalloc(tree, q, 8_000_000)
        torch.cuda.synchronize()
        total = int(want_cnt[0])
        assert total > nq and int(want_cnt[1]) == 0
        assert torch.equal(got_cnt, want_cnt)
        assert torch.equal(got_pairs[:total], want_pairs[:total])


def test_join_bu_noalloc_rejects_beyond_its_maximum(dev):
    r = gen.rects_uniform(gen.seed(0), 4096)
    tree = rtree.build_bottom_up(torch.from_numpy(r).to(dev), 16)
    q = torch.zeros((262145, 4), dtype=torch.float32, device=dev)
    with pytest.raises(PrtError) as e:
        _noalloc(tree, q, 16)
    assert "at most 262144" in str(e.value)
