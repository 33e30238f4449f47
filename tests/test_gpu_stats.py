"""GPU: the reference's search counters and sibling-overlap count, computed
on the device over the GPU-built trees, equal to the oracle's (exact
integers)."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())


def expected(num, cnt, num1):
    n = max(len(num), 1)
    return {"sum": int(num.sum()), "max": int(num.max(initial=0)), "sum1": int(cnt.sum()),
            "max1": int(cnt.max(initial=0)), "hits": int(num1.sum()),
            "sum2": int(num1.sum()) // 2, "max2": int(num1.max(initial=0)),
            "ave": num.sum() / n}


def build(kind, r, f):
    d = torch.from_numpy(r).cuda()
    return (rtree.build_bottom_up(d, f), O.bu_build(r, f)) if kind == "bu" else \
        (rtree.build_top_down(d, f), O.td_build(r, f))


@pytest.mark.parametrize("kind,n,f", [("bu", 20000, 16), ("bu", 20000, 10), ("td", 20000, 16),
                                      ("td", 30001, 10), ("bu", 5, 2), ("td", 17, 16)])
def test_search_stats_self_join(dev, kind, n, f):
    # the reference searches every data rect against its own tree
    r = gen.rects_uniform(gen.seed(0) + n, n, 1e-3, 1e-2)
    t, ot = build(kind, r, f)
    got = rtree.search_stats(t, torch.from_numpy(r).cuda())
    want = expected(*O.search_stats(ot, r))
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)
    assert rtree.sibling_overlap(t)["sum3"] == O.sibling_overlap(ot)


@pytest.mark.parametrize("kind", ["bu", "td"])
def test_search_stats_windows(dev, kind):
    r = gen.rects_uniform(gen.seed(0), 100000)
    q = gen.windows(gen.seed(0) + 1, 4096)
    t, ot = build(kind, r, 16)
    got = rtree.search_stats(t, torch.from_numpy(q).cuda())
    want = expected(*O.search_stats(ot, q))
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)


def test_tiny_golden_tree(dev):
    g = GOLD["bu_tiny"]
    r = np.array(g["rects"], np.float32)
    t = rtree.build_bottom_up(torch.from_numpy(r).cuda(), g["fanout"])
    q = np.array([[0.3, 0.3, 0.5, 0.5], [0.0, 0.0, 1.0, 1.0], [0.96, 0.96, 0.99, 0.99]],
                 np.float32)
    got = rtree.search_stats(t, torch.from_numpy(q).cuda())
    assert (got["sum"], got["max"], got["sum1"], got["max1"], got["hits"]) == (20, 10, 5, 4, 7)
    assert rtree.sibling_overlap(t)["sum3"] == 0
    empty = rtree.search_stats(t, torch.zeros((0, 4), dtype=torch.float32, device=dev))
    assert empty["sum"] == 0 and empty["max"] == 0
