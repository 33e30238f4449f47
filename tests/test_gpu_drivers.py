"""GPU: the command-line drivers (host C++ over the C-ABI only) run on files
in the reference's text formats and print the reference's counters.

prt_bu_rtree / prt_td_rtree: every line but the timings equals the text the
reference's printf formats give for the oracle's counters on the same tree
(BotomUp-rtree.cu:703-713, 1021-1026; TopDown-rtree.cu:793-797).
prt_overlay: pair count and count1 equal the oracle's, total area within
1e-12 relative of the oracle's FP64 areas summed."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from parlllelrtree_amd import gen, refio
from parlllelrtree_amd.build import BIN

pytestmark = pytest.mark.gpu


def run(name, *args):
    exe = BIN / name
    assert exe.exists(), f"{exe} is not built (parlllelrtree_amd.build.build_drivers)"
    res = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    return [ln for ln in res.stdout.splitlines() if not ln.startswith("GPU ")]


def int_rects(n, seed, lo=-3_000_000, span=9_000_000, wmax=40_000):
    rng = np.random.default_rng(seed)
    x1 = rng.integers(lo, lo + span, n)
    y1 = rng.integers(0, span, n)
    return np.stack([x1, y1, x1 + rng.integers(0, wmax, n), y1 + rng.integers(0, wmax, n)], 1)


def stat_lines(num, cnt, num1, sum3, ave3, td):
    n = max(len(num), 1)
    s, s1, s2 = int(num.sum()), int(cnt.sum()), int(num1.sum()) // 2
    m, m1, m2 = (int(a.max(initial=0)) for a in (num, cnt, num1))
    if td:
        out = ["sum = %d  ave = %0.2f max = %d" % (s, s / n, m),
               "sum1 = %d ave1 = %0.2f max1 = %d " % (s1, s1 / n, m1),
               "sum2 = %d ave2 = %0.2f max2 = %d " % (s2, s2 / n, m2)]
    else:
        out = ["sum=%d  ave = %0.2f max = %d" % (s, s / n, m),
               "sum1=%d  ave = %0.2f max1 = %d" % (s1, s1 / n, m1),
               "sum2=%d  ave = %0.2f max2 = %d" % (s2, s2 / n, m2)]
    return out + ["sum3 = %d  ave3 = %0.2f" % (sum3, ave3)]


def same_stat_lines(got, want):
    """Line-for-line equality; an ave3 of a zero node budget (inf or nan,
    whose sign the C library prints its own way) only has to be one."""
    if len(got) != len(want):
        return False
    for g, w in zip(got, want):
        if g != w and not (w.endswith(("nan", "inf")) and g.split("ave3 = ")[0] ==
                           w.split("ave3 = ")[0] and g.endswith(w[-3:])):
            return False
    return True


@pytest.mark.parametrize("n,f", [(1, 10), (37, 4), (20000, 10), (100003, 16)])
def test_bu_driver(dev, tmp_path, n, f):
    r = int_rects(n, n)
    r[: n // 9] = r[0]
    path = tmp_path / "rects.txt"
    refio.write_rects_text(path, r, np.arange(n, dtype=np.int32))
    got = run("prt_bu_rtree", path, f)
    rn, mnx, mny, scale = refio.normalize(r)
    head = ["%d %d %d %d" % (mnx, max(0, r[:, 2].max()), mny, max(0, r[:, 3].max())), "%d" % scale]
    t = O.bu_build(refio.rects_to_f32(rn), f, keys64=O.morton_ref4d(rn, scale))
    q = refio.rects_to_f32(rn)
    pairs, _ = O.join_bu(t, q)
    s3 = O.sibling_overlap(t)
    # ave3 = sum3 / num_nodes / m with the reference's node budget
    # num_nodes = 3 * num_contours / m (integer division, BU:674, 1026)
    budget = 3 * n // f
    ave3 = s3 / budget / f if budget else (float("inf") if s3 else float("nan"))
    want = head + ["pairs=%d" % len(pairs)] + \
        stat_lines(*O.search_stats(t, q), s3, ave3, td=False)
    assert same_stat_lines(got, want), (got, want)


@pytest.mark.parametrize("n,f", [(1, 10), (37, 4), (20000, 10), (100003, 16)])
def test_td_driver(dev, tmp_path, n, f):
    r = int_rects(n, n + 1, lo=0, span=8_000_000)
    path = tmp_path / "rects.txt"
    refio.write_rects_text(path, r)
    got = run("prt_td_rtree", path, f)
    q = refio.rects_to_f32(r)
    t = O.td_build(q, f)
    pairs, _ = O.join_td(t, q)
    s3 = O.sibling_overlap(t)
    want = ["pairs=%d" % len(pairs)] + \
        stat_lines(*O.search_stats(t, q), s3, s3 / len(t["slot_child"]), td=True)
    assert got == want


def test_overlay_driver(dev, tmp_path):
    base = gen.stars_grid(gen.seed(2), 3000, 55)
    over = gen.stars_grid(gen.seed(2, True), 3000, 55)
    pb, po = tmp_path / "base.txt", tmp_path / "overlay.txt"
    refio.write_polys_text(pb, base)
    refio.write_polys_text(po, over)
    got = run("prt_overlay", pb, po, 16)
    pairs = O.join_brute(gen.poly_mbr(base), gen.poly_mbr(over))  # (overlay id, base id)
    area = O.overlay(pairs, over, base, nthreads=8)
    assert got[0] == "base 3000 polygons, overlay 3000 polygons"
    assert got[1] == "pairs=%d" % len(pairs)
    assert got[2] == "count1=%d" % int((area > 0).sum())
    assert got[3].startswith("total_area=")
    assert float(got[3].split("=")[1]) == pytest.approx(float(area.sum()), rel=1e-12)
