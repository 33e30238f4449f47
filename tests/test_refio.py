"""The reference's text formats: hand-written fixtures in the layouts its
readers scan (rects: BotomUp-rtree.cu:668-691; polygons:
vector_map_overlay_gpc_kernel.cu:766-818), the bottom-up normalisation
(BU:682-721) against the oracle's scale rule, and write/read round trips of
the synthetic workloads (host only)."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from parlllelrtree_amd import gen, refio
from parlllelrtree_amd._lib import PrtError

GOLD = Path(__file__).parent / "golden"


def test_read_rects_fixture():
    r, k = refio.read_rects_text(GOLD / "rects_ref_format.txt")
    # file order per record is x1 x2 y1 y2 key; arrays hold (x1, y1, x2, y2)
    np.testing.assert_array_equal(r, [[100, 300, 220, 410], [-50, -20, 10, 35],
                                      [1000000, 2000000, 1000100, 2000050], [0, 0, 0, 0],
                                      [5, 7, 6, 8]])
    np.testing.assert_array_equal(k, [7, 8, 9, 10, -1])


def test_normalize_fixture_and_oracle_scale():
    r, _ = refio.read_rects_text(GOLD / "rects_ref_format.txt")
    s, mx, my, sc = refio.normalize(r)
    assert (mx, my) == (-50, -20)
    # ranges 1000150 (x) and 2000070 (y): 2000070 -> 200007 -> 20000 -> 2000
    assert sc == 1000
    assert sc == O.ref_scale(max(r[:, 2].max() - mx, r[:, 3].max() - my))
    np.testing.assert_array_equal(s[0], [150, 320, 270, 430])
    np.testing.assert_array_equal(s, r - np.array([mx, my, mx, my]))


def test_normalize_random_against_oracle():
    rng = np.random.default_rng(1)
    for scale_exp in (2, 5, 9, 12):
        x1 = rng.integers(-10 ** scale_exp, 10 ** scale_exp, 500)
        y1 = rng.integers(-10 ** scale_exp, 10 ** scale_exp, 500)
        r = np.stack([x1, y1, x1 + rng.integers(0, 1000, 500), y1 + rng.integers(0, 1000, 500)], 1)
        s, mx, my, sc = refio.normalize(r)
        assert (mx, my) == (r[:, 0].min(), r[:, 1].min())
        # the reference starts its maxima at 0 (BU:664, 666)
        rng_ = max(max(r[:, 2].max(), 0) - mx, max(r[:, 3].max(), 0) - my)
        assert sc == O.ref_scale(int(rng_))


def test_read_polys_fixture():
    pf = refio.read_polys_text(GOLD / "polys_ref_format.txt")
    np.testing.assert_array_equal(pf.polys.off, [0, 4, 7])
    np.testing.assert_array_equal(pf.polys.xy[:4], [[0.1, 0.2], [0.3, 0.2], [0.3, 0.4], [0.1, 0.4]])
    np.testing.assert_array_equal(pf.polys.xy[4:], [[-0.000015, 0.5], [1.25, 0.5], [0.0, 2.0]])
    np.testing.assert_array_equal(pf.hole, [0, 1])
    # (int)(v * 100000) truncates toward zero: -1.5 -> -1
    np.testing.assert_array_equal(pf.box, [[10000, 20000, 30000, 40000],
                                           [-1, 50000, 125000, 200000]])
    assert pf.min_xy == (-1, 20000)


def test_rects_round_trip(tmp_path):
    r = np.random.default_rng(2).integers(-2 ** 40, 2 ** 40, (1000, 4))
    k = np.arange(1000, dtype=np.int32) * 3 - 7
    refio.write_rects_text(tmp_path / "r.txt", r, k)
    r2, k2 = refio.read_rects_text(tmp_path / "r.txt")
    np.testing.assert_array_equal(r2, r)
    np.testing.assert_array_equal(k2, k)


def test_polys_round_trip_bit_exact(tmp_path):
    S = gen.stars_grid(gen.seed(2), 2000, 45)
    hole = (np.arange(S.n) % 3 == 0).astype(np.int32)
    refio.write_polys_text(tmp_path / "p.txt", S, hole)
    pf = refio.read_polys_text(tmp_path / "p.txt")
    np.testing.assert_array_equal(pf.polys.off, S.off)
    assert np.array_equal(pf.polys.xy.view(np.int64), S.xy.view(np.int64))  # %.17g round trip
    np.testing.assert_array_equal(pf.hole, hole)
    lo = np.minimum.reduceat(S.xy, S.off[:-1], axis=0)
    np.testing.assert_array_equal(pf.box[:, 0], (lo[:, 0] * 100000).astype(np.int64))


def test_malformed_files_raise(tmp_path):
    (tmp_path / "bad.txt").write_text("3\n1 2 3 4 5\n1 2 x 4 5\n")
    with pytest.raises(PrtError):
        refio.read_rects_text(tmp_path / "bad.txt")
    (tmp_path / "short.txt").write_text("2\n3 0\n0 0 1 1\n0 0\n1 0\n")
    with pytest.raises(PrtError):
        refio.read_polys_text(tmp_path / "short.txt")
    with pytest.raises(PrtError):
        refio.read_rects_text(tmp_path / "missing.txt")


def test_rects_to_f32_exactness():
    ok = np.array([[0, 0, 1 << 24, 5]], np.int64)
    assert refio.rects_to_f32(ok).dtype == np.float32
    with pytest.raises(ValueError):
        refio.rects_to_f32(np.array([[0, 0, (1 << 24) + 1, 5]], np.int64))
