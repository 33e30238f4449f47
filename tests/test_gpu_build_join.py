"""GPU parity: Morton keys, radix sort, bottom-up / top-down builds and joins,
bit-exact against the CPU oracle, from tiny edge cases up to the full
1,048,576-rect workloads."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def dev_rects(r):
    return torch.from_numpy(np.ascontiguousarray(r, np.float32)).cuda()


def rects(n, s):
    return gen.rects_uniform(gen.seed(0) + s, n)


def test_morton_center30(dev):
    g = np.array([c["rect"] for c in GOLD["center30"]], np.float32)
    k = u32(rtree.morton_center30(dev_rects(g)))
    assert [int(v) for v in k] == [c["key"] for c in GOLD["center30"]]
    r = rects(1 << 20, 1)
    np.testing.assert_array_equal(u32(rtree.morton_center30(dev_rects(r))), O.morton_center30(r))


def test_morton_ref4d(dev):
    from parlllelrtree_amd._lib import lib, check
    cases = GOLD["morton_ref4"]
    inp = torch.tensor([c["in"] for c in cases], dtype=torch.int64, device=dev)
    out = torch.empty(len(cases), dtype=torch.int64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    check(lib().prt_morton_ref4d(inp.data_ptr(), len(cases), 1, out.data_ptr(), s))
    got = out.cpu().numpy().view(np.uint64)
    assert [int(v) for v in got] == [c["key"] for c in cases]
    rng = np.random.default_rng(1)
    r = rng.integers(0, 10**7, size=(100000, 4), dtype=np.int64)
    scale = O.ref_scale(int(r.max()))
    inp = torch.from_numpy(r).cuda()
    out = torch.empty(len(r), dtype=torch.int64, device=dev)
    check(lib().prt_morton_ref4d(inp.data_ptr(), len(r), scale, out.data_ptr(), s))
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint64), O.morton_ref4d(r, scale))


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 100003, 1 << 20, (1 << 22) + 4099])
@pytest.mark.parametrize("bits,hi", [(30, 1 << 30), (32, 1 << 32), (8, 7)])
def test_sort_pairs_stable(dev, n, bits, hi):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    kt = torch.from_numpy(keys.view(np.int32)).cuda()
    ko, io = rtree.sort_pairs_u32(kt, bits)
    perm = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(u32(io), perm)
    np.testing.assert_array_equal(u32(ko), keys[perm])


def check_bu(t, o):
    np.testing.assert_array_equal(u32(t.perm), o["perm"])
    np.testing.assert_array_equal(u32(t.keys), o["keys"])
    for k in ("node_count", "node_start", "node_end", "entry_ref"):
        np.testing.assert_array_equal(getattr(t, k).cpu().numpy(), o[k], err_msg=k)
    # bitwise: the sign of a zero counts
    np.testing.assert_array_equal(t.entry_mbr.cpu().numpy().view(np.uint32),
                                  np.asarray(o["entry_mbr"], np.float32).view(np.uint32))
    assert t.levels == o["levels"]


def test_bu_golden(dev):
    g = GOLD["bu_tiny"]
    t = rtree.build_bottom_up(dev_rects(g["rects"]), g["fanout"])
    assert list(u32(t.perm)) == g["perm"]
    for k in ("node_count", "node_start", "node_end", "entry_ref"):
        assert getattr(t, k).cpu().tolist() == g[k], k
    np.testing.assert_array_equal(t.entry_mbr.cpu().numpy(), np.array(g["entry_mbr"], np.float32))


@pytest.mark.parametrize("n,f", [(1, 16), (2, 2), (16, 16), (17, 16), (1000, 10), (4097, 16),
                                 (65537, 16), (3000, 64), (20000, 4), (1 << 20, 16)])
def test_bu_build_parity(dev, n, f):
    r = rects(n, n)
    t = rtree.build_bottom_up(dev_rects(r), f)
    check_bu(t, O.bu_build(r, f))


def test_bu_build_duplicate_keys(dev):
    # many identical centres: ties must keep input order
    r = np.tile(np.array([[0.25, 0.25, 0.25, 0.25]], np.float32), (50000, 1))
    r[::3] = [0.5, 0.5, 0.5, 0.5]
    t = rtree.build_bottom_up(dev_rects(r), 16)
    check_bu(t, O.bu_build(r, 16))


def check_td(t, o):
    assert t.level_offset == o["level_offset"]
    np.testing.assert_array_equal(u32(t.perm), o["perm"])
    np.testing.assert_array_equal(t.slot_child.cpu().numpy(), o["slot_child"])
    np.testing.assert_array_equal(t.slot_mbr.cpu().numpy().view(np.uint32),
                                  np.asarray(o["slot_mbr"], np.float32).view(np.uint32))


def test_td_golden(dev):
    g = GOLD["td_tiny"]
    t = rtree.build_top_down(dev_rects(g["rects"]), g["fanout"])
    assert list(u32(t.perm)) == g["perm"]
    assert t.slot_child.cpu().tolist() == g["slot_child"]
    np.testing.assert_array_equal(t.slot_mbr.cpu().numpy(), np.array(g["slot_mbr"], np.float32))


@pytest.mark.parametrize("n,f", [(1, 16), (2, 16), (16, 16), (17, 16), (300, 4), (1000, 10),
                                 (5000, 16), (70000, 16), (200000, 10), (300000, 4), (1 << 20, 16),
                                 ((1 << 22) + 17, 16)])
def test_td_build_parity(dev, n, f):
    r = rects(n, 7 * n)
    t = rtree.build_top_down(dev_rects(r), f)
    check_td(t, O.td_build(r, f))


@pytest.mark.parametrize("nl,n,f", [(1, 1, 16), (2, 1, 16), (2, 17, 16), (3, 300, 4), (2, 5000, 16),
                                    (4, 70000, 16), (2, 70001, 16), (3, 1000, 10), (2, 200000, 8),
                                    (2, 300000, 2), (2, 1 << 20, 16)])
def test_td_layers_parity(dev, nl, n, f):
    """Layers built by one shared launch sequence: each tree bit-exact
    against the oracle's build of that layer alone; then rebuilt in place
    from other layers."""
    layers = [rects(n, 11 * n + l) for l in range(nl)]
    TL = rtree.TopDownLayers([dev_rects(r) for r in layers], f)
    for t, r in zip(TL.trees, layers):
        check_td(t, O.td_build(r, f))
    other = [rects(n, 13 * n + 5 + l) for l in range(nl)]
    TL.rebuild([dev_rects(r) for r in other])
    for t, r in zip(TL.trees, other):
        check_td(t, O.td_build(r, f))


def test_td_layers_equal_single_builds_16M(dev):
    """Two layers of 16,777,216 clustered rects (three radix levels, the
    first two moving the rects): equal to the single-layer builds."""
    from parlllelrtree_amd import workloads
    a = torch.from_numpy(workloads.w4_rects(1 << 24)).cuda()
    b = a.flip(0).contiguous()
    TL = rtree.TopDownLayers([a, b], 16)
    for t, r in zip(TL.trees, (a, b)):
        s = rtree.build_top_down(r, 16)
        assert torch.equal(t.perm, s.perm)
        assert torch.equal(t.slot_child, s.slot_child)
        assert torch.equal(t.slot_mbr.view(torch.int32), s.slot_mbr.view(torch.int32))
        del s
    with pytest.raises(ValueError):
        rtree.TopDownLayers([a, b[:10]], 16)


def test_kernel_times_of_layered_build(dev):
    """Per-launch timing (hipExtLaunchKernel events): one launch of each
    kernel covers both layers; the profiled launches build the same trees."""
    from parlllelrtree_amd.profiling import KernelTimes
    layers = [rects(1 << 20, s) for s in (21, 22)]
    TL = rtree.TopDownLayers([dev_rects(r) for r in layers], 16)
    with KernelTimes() as kt:
        TL.rebuild([dev_rects(r) for r in layers])
        torch.cuda.synchronize()
    assert kt.stats["k_onesweep"][0] == 8          # 4 digit passes x 2 radix levels
    assert kt.stats["k_hist_all<Src>"][0] == 2
    assert kt.stats["k_td_subtree<8>"][0] == 1 and kt.stats["k_td_top"][0] == 1
    assert all(0 < ms < 50 for _, ms in kt.stats.values())
    for t, r in zip(TL.trees, layers):
        check_td(t, O.td_build(r, 16))


def signed_zero_rects(n, seed):
    """Every rect has x1, y1 = a zero of random sign and x2, y2 = a zero of
    random sign or a positive value, so that every union meets zeros of
    both signs (min and max take -0 below +0)."""
    rng = np.random.default_rng(seed)
    z = lambda: np.where(rng.random(n) < 0.5, np.float32(-0.0), np.float32(0.0))
    r = np.empty((n, 4), np.float32)
    r[:, 0], r[:, 1] = z(), z()
    r[:, 2] = np.where(rng.random(n) < 0.7, z(), rng.random(n).astype(np.float32))
    r[:, 3] = np.where(rng.random(n) < 0.7, z(), rng.random(n).astype(np.float32))
    return r


@pytest.mark.parametrize("n", [300, 5000, 70000])
def test_signed_zero_unions(dev, n):
    r = signed_zero_rects(n, n)
    check_td(rtree.build_top_down(dev_rects(r), 16), O.td_build(r, 16))
    check_bu(rtree.build_bottom_up(dev_rects(r), 16), O.bu_build(r, 16))
    TL = rtree.TopDownLayers([dev_rects(r), dev_rects(r[::-1].copy())], 16)
    check_td(TL.trees[0], O.td_build(r, 16))


def test_td_build_duplicate_coords(dev):
    r = rects(30000, 5)
    r[:, 0] = np.round(r[:, 0] * 50) / 50   # many equal x1
    r[:, 1] = np.round(r[:, 1] * 50) / 50
    t = rtree.build_top_down(dev_rects(r), 16)
    check_td(t, O.td_build(r, 16))


@pytest.mark.parametrize("n,nq,f", [(1, 5, 16), (3000, 2000, 16), (3000, 500, 10), (5000, 300, 4),
                                    (20000, 4000, 64)])
def test_join_small_parity(dev, n, nq, f):
    r = rects(n, 3)
    q = gen.windows(gen.seed(0) + 0x200, nq, 0.02)
    k = min(nq // 4, n)
    q[:k] = r[:k]  # self hits: exact touching boundaries
    brute = O.join_brute(r, q)
    bu = rtree.join(rtree.build_bottom_up(dev_rects(r), f), dev_rects(q))
    td = rtree.join(rtree.build_top_down(dev_rects(r), f), dev_rects(q))
    np.testing.assert_array_equal(u32(bu), brute)
    np.testing.assert_array_equal(u32(td), brute)


def test_join_touching_and_empty(dev):
    r = np.array([[0, 0, 1, 1], [1, 1, 2, 2], [3, 3, 4, 4]], np.float32)
    q = np.array([[1, 1, 1, 1], [5, 5, 6, 6], [-1, -1, 10, 10]], np.float32)
    for build in (rtree.build_bottom_up, rtree.build_top_down):
        pairs = rtree.join(build(dev_rects(r), 2), dev_rects(q))
        assert pairs.cpu().tolist() == [[0, 0], [0, 1], [2, 0], [2, 1], [2, 2]]
    none = rtree.join(rtree.build_bottom_up(dev_rects(r), 2), dev_rects(q[1:2]))
    assert none.shape[0] == 0


def test_join_segment_sizes(dev):
    """Queries with 0 .. ~1000 hits: every sort path of the hit segments
    (one thread, a wave with 1-16 registers per lane) against brute force."""
    r = rects(20000, 17)
    rng = np.random.default_rng(5)
    side = np.concatenate([np.full(50, 0.002), np.linspace(0.01, 0.22, 350)]).astype(np.float32)
    c = rng.random((len(side), 2)).astype(np.float32) * 0.7 + 0.15
    q = np.stack([c[:, 0], c[:, 1], c[:, 0] + side, c[:, 1] + side], 1).astype(np.float32)
    brute = O.join_brute(r, q)
    hits = np.bincount(brute[:, 0], minlength=len(q))
    assert hits.max() > 700 and (hits <= 1024).all() and (hits == 0).any()
    for build in (rtree.build_bottom_up, rtree.build_top_down):
        pairs = rtree.join(build(dev_rects(r), 16), dev_rects(q))
        np.testing.assert_array_equal(u32(pairs), brute)


def test_join_truncated_at_cap(dev):
    """A pair buffer smaller than the hit count: the total still comes back,
    join() refuses the result, and the queries whose hits all fit below cap
    hold their complete sorted hit lists (both the copied and the traversed
    write paths)."""
    r = rects(20000, 23)
    rng = np.random.default_rng(8)
    side = np.where(rng.random(600) < 0.5, 0.004, 0.05).astype(np.float32)
    c = rng.random((600, 2)).astype(np.float32) * 0.9
    q = np.stack([c[:, 0], c[:, 1], c[:, 0] + side, c[:, 1] + side], 1).astype(np.float32)
    brute = O.join_brute(r, q)
    cap = len(brute) // 2
    for build in (rtree.build_bottom_up, rtree.build_top_down):
        t = build(dev_rects(r), 16)
        with pytest.raises(RuntimeError):
            rtree.join(t, dev_rects(q), cap=cap)
        j = rtree.Joiner(t, len(q), cap, torch.device("cuda"))
        j.run(dev_rects(q))
        total = int(j.count[0])
        assert total == len(brute)
        got = u32(j.pairs[:cap])
        hits = np.bincount(brute[:, 0], minlength=len(q))
        ends = np.cumsum(hits)
        full = int(np.searchsorted(ends, cap, side="right"))   # queries 0..full-1 fit
        k = int(ends[full - 1]) if full else 0
        assert k > 1000 and (hits[:full] <= 8).any() and (hits[:full] > 8).any()
        np.testing.assert_array_equal(got[:k], brute[:k])


def test_join_big_segments(dev):
    # one window covering everything: 6000 hits for one query (block sort path)
    r = rects(6000, 9)
    q = np.array([[-1, -1, 2, 2], [0.4, 0.4, 0.6, 0.6]], np.float32)
    pairs = rtree.join(rtree.build_bottom_up(dev_rects(r), 16), dev_rects(q))
    np.testing.assert_array_equal(u32(pairs), O.join_brute(r, q))


@pytest.mark.parametrize("n", [8193, 20000, 100003])
def test_join_segments_beyond_the_block_sort(dev, n):
    """Queries with more hits than one block sorts in LDS (8192): sorted in
    place in the pair buffer (chunks in LDS, the merges over the buffer),
    for both trees, beside queries with a few hits; the join status stays 0."""
    r = rects(n, 31)
    q = np.array([[-1, -1, 2, 2], [0.4, 0.4, 0.45, 0.45], [0, 0, 0.7, 1.5], [5, 5, 6, 6]],
                 np.float32)
    brute = O.join_brute(r, q)
    assert np.bincount(brute[:, 0], minlength=4)[0] == n
    for build in (rtree.build_bottom_up, rtree.build_top_down):
        pairs = rtree.join(build(dev_rects(r), 16), dev_rects(q))
        np.testing.assert_array_equal(u32(pairs), brute)


def test_workload0_windows_full(dev):
    """1,048,576 rects, BU fanout 16, 65,536 windows of side 1e-2."""
    r = gen.rects_uniform(gen.seed(0), 1 << 20)
    q = gen.windows(gen.seed(0) + gen.QUERY_STREAM, 65536)
    t = rtree.build_bottom_up(dev_rects(r), 16)
    o = O.bu_build(r, 16)
    check_bu(t, o)
    pairs = u32(rtree.join(t, dev_rects(q)))
    opairs, _ = O.join_bu(o, q)
    assert len(pairs) == len(opairs) and len(pairs) > 7_000_000
    np.testing.assert_array_equal(pairs, opairs)


def test_workload1_layers_full(dev):
    """Two layers of 1,048,576 rects, TD fanout 16, layer-layer join."""
    ra = gen.rects_uniform(gen.seed(1), 1 << 20)
    rb = gen.rects_uniform(gen.seed(1, True), 1 << 20)
    t = rtree.build_top_down(dev_rects(ra), 16)
    o = O.td_build(ra, 16)
    check_td(t, o)
    pairs = u32(rtree.join(t, dev_rects(rb)))
    opairs, _ = O.join_td(o, rb)
    assert len(pairs) == len(opairs) and len(pairs) > 1_000_000
    np.testing.assert_array_equal(pairs, opairs)


@pytest.mark.parametrize("n,f", [(1, 10), (5, 2), (1000, 10), (20000, 16), (300001, 10)])
def test_bu_build_ref4d_bit_exact(dev, n, f):
    # integer rects in the reference's units, normalised as BotomUp-rtree.cu:682-721 does
    from parlllelrtree_amd import refio
    rng = np.random.default_rng(n)
    x1 = rng.integers(-3_000_000, 9_000_000, n)
    y1 = rng.integers(0, 12_000_000, n)
    r = np.stack([x1, y1, x1 + rng.integers(0, 5000, n), y1 + rng.integers(0, 5000, n)], 1)
    r[: n // 7] = r[0]                   # duplicate keys: ties must keep input order
    rn, _, _, scale = refio.normalize(r)
    keys = O.morton_ref4d(rn, scale)
    ot = O.bu_build(refio.rects_to_f32(rn), f, keys64=keys)
    t, k64 = rtree.build_bottom_up_ref4d(torch.from_numpy(rn).cuda(), scale, f)
    np.testing.assert_array_equal(t.perm.cpu().numpy().view(np.uint32), ot["perm"])
    np.testing.assert_array_equal(k64.cpu().numpy().view(np.uint64), keys[ot["perm"]])
    np.testing.assert_array_equal(t.keys.cpu().numpy().view(np.uint32),
                                  (keys[ot["perm"]] >> np.uint64(32)).astype(np.uint32))
    for k in ("node_count", "node_start", "node_end", "entry_ref"):
        np.testing.assert_array_equal(getattr(t, k).cpu().numpy(), ot[k], err_msg=k)
    np.testing.assert_array_equal(t.entry_mbr.cpu().numpy().view(np.uint32),
                                  ot["entry_mbr"].view(np.uint32))


def test_builds_and_joins_on_signed_points_and_wide_magnitudes(dev):
    """Negative coordinates, points (zero-size rects), subnormal and huge
    magnitudes: the top-down keys order them by ord32, the bottom-up
    CENTER30 key clamps centres outside [0, 1] (long runs of equal keys kept
    in input order); trees bit-exact, joins equal to brute force."""
    rng = np.random.default_rng(31)
    n = 40000
    lo = rng.choice([-1e3, -1.0, 0.0, 1e-40, 1e-3, 1.0, 1e30], size=(n, 2)).astype(np.float32)
    lo *= rng.random((n, 2)).astype(np.float32)
    w = np.where(rng.random((n, 2)) < 0.3, np.float32(0),
                 rng.random((n, 2)).astype(np.float32) * np.float32(0.01))
    r = np.concatenate([lo, lo + w], 1).astype(np.float32)
    check_td(rtree.build_top_down(dev_rects(r), 16), O.td_build(r, 16))
    obu = O.bu_build(r, 16)
    check_bu(rtree.build_bottom_up(dev_rects(r), 16), obu)
    TL = rtree.TopDownLayers([dev_rects(r), dev_rects(r[::-1].copy())], 16)
    check_td(TL.trees[0], O.td_build(r, 16))
    check_td(TL.trees[1], O.td_build(r[::-1].copy(), 16))
    q = r[rng.integers(0, n, 3000)].copy()
    q[:, 2:] += np.float32(0.005)
    brute = O.join_brute(r, q)
    assert len(brute) > 3000
    np.testing.assert_array_equal(u32(rtree.join(rtree.build_bottom_up(dev_rects(r), 16),
                                                 dev_rects(q))), brute)
    np.testing.assert_array_equal(u32(rtree.join(TL.trees[0], dev_rects(q))), brute)


def test_lookback_giveup_is_reported(dev):
    """The radix look-back's safety net, forced from the host: tile 1 of
    every pass publishes nothing and the others poll it 16 times before they
    give up.  Every kernel still ends; the sort's order is then wrong and
    prt_device_status reports PRT_STATUS_LOOKBACK as PRT_E_DEVICE.  With the
    defaults restored the same sort is exact and raises nothing."""
    from parlllelrtree_amd._lib import PRT_E_DEVICE, PRT_STATUS_LOOKBACK, PrtError, check, lib
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 32, size=100_000, dtype=np.uint64).astype(np.uint32)
    kd = torch.from_numpy(keys.view(np.int32)).to(dev)
    assert rtree.device_status() == 0
    check(lib().prt_debug_lookback(16, 1))
    try:
        rtree.sort_pairs_u32(kd, 32)
        assert rtree.device_status(clear=False) & PRT_STATUS_LOOKBACK
        with pytest.raises(PrtError) as e:
            rtree.check_device_status()
        assert e.value.code == PRT_E_DEVICE and "look-back" in str(e.value)
        assert rtree.device_status() == 0      # check_device_status cleared them
    finally:
        check(lib().prt_debug_lookback(-1, -1))
    sk, si = rtree.sort_pairs_u32(kd, 32)
    order = O.stable_order(keys)
    np.testing.assert_array_equal(u32(si), order)
    np.testing.assert_array_equal(u32(sk), keys[order])
    assert rtree.device_status() == 0


def test_bu_build_ref4d_inexact_coordinates_flagged(dev):
    """int64 coordinates just above 2^24 that fp32 cannot hold: the boxes
    are rounded outward (they contain the rects) and PRT_STATUS_INEXACT is
    raised; coordinates above 2^24 that fp32 holds exactly raise nothing."""
    from parlllelrtree_amd._lib import PRT_STATUS_INEXACT
    base = 1 << 24
    exact = np.array([[base, 0, base + 2, 6], [base + 4, 2, base + 8, 10],
                      [0, 0, 1 << 30, 1 << 40]], np.int64)
    rtree.build_bottom_up_ref4d(torch.from_numpy(exact).cuda(), 1, 2)
    assert rtree.device_status() == 0
    odd = np.array([[base + 1, 0, base + 3, 5], [base + 5, 1, base + 9, 7],
                    [-(base + 1), -3, 3, base + 3]], np.int64)
    t, _ = rtree.build_bottom_up_ref4d(torch.from_numpy(odd).cuda(), 1, 2)
    assert rtree.device_status() & PRT_STATUS_INEXACT
    mbr = t.entry_mbr.cpu().numpy().reshape(-1, 4)
    ref = t.entry_ref.cpu().numpy()
    for slot in range(t.st.level_count[0] * t.st.fanout):
        if ref[slot] < 0:
            continue
        lo_x, lo_y, hi_x, hi_y = (int(v) for v in mbr[slot])
        r = odd[ref[slot]]
        assert lo_x <= r[0] and lo_y <= r[1] and hi_x >= r[2] and hi_y >= r[3]
        assert (lo_x, lo_y, hi_x, hi_y) != tuple(int(v) for v in r)
