"""GPU parity: FP64 intersection areas of the plain per-pair kernel against
the CPU oracle.  Bar: relative error <= 1e-12 with the denominator
max(|A_oracle|, 1e-12 * min(area P, area Q)); the kernel evaluates the
oracle's expressions in the oracle's order, so the areas are expected to be
bit-identical and the exact-match count is asserted as well."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle as O
from parlllelrtree_amd import gen, rtree, workloads
from parlllelrtree_amd.gen import PolyCSR

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())
REL_TOL = 1e-12


def rel_err(got, ref, A, B, pairs):
    pa = np.array([abs(O.poly_area(A.poly(i))) for i in pairs[:, 0]])
    pb = np.array([abs(O.poly_area(B.poly(i))) for i in pairs[:, 1]])
    den = np.maximum(np.abs(ref), 1e-12 * np.minimum(pa, pb))
    den = np.where(den == 0, 1.0, den)
    return np.abs(got - ref) / den


def csr(polys):
    off = np.zeros(len(polys) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in polys])
    xy = np.concatenate([np.asarray(p, float).reshape(-1, 2) for p in polys]) if polys else \
        np.zeros((0, 2))
    return PolyCSR(np.ascontiguousarray(xy), off)


def nan_area(n):
    # every output slot starts as NaN: a slot the kernel does not write fails
    return torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")


def run_plain(pairs, A, B):
    area = rtree.overlay_plain(torch.from_numpy(pairs.view(np.int32)).cuda(),
                               rtree.to_device_polys(A), rtree.to_device_polys(B),
                               nan_area(len(pairs)))
    return area.cpu().numpy()


def test_golden_cases(dev):
    cases = GOLD["clip"]
    A = csr([c["P"] if len(c["P"]) else [[0, 0]] for c in cases])
    B = csr([c["Q"] for c in cases])
    idx = np.arange(len(cases), dtype=np.uint32)
    pairs = np.stack([idx, idx], 1)
    got = run_plain(pairs, A, B)
    for c, g in zip(cases, got):
        assert g == pytest.approx(c["area"], rel=1e-15, abs=1e-15), c["name"]
    got_sw = run_plain(pairs, B, A)
    np.testing.assert_allclose(got_sw, [c["area"] for c in cases], rtol=1e-15, atol=1e-15)


def test_random_stars_bit_exact(dev):
    S = gen.stars_grid(gen.seed(9), n=4000, grid=64, jitter=0.25, rfac=1.2, kmax=300)
    T = gen.stars_grid(gen.seed(9, True), n=4000, grid=64, jitter=0.25, rfac=1.2, kmax=300)
    rng = np.random.default_rng(0)
    a = rng.integers(0, 4000, 20000)
    b = np.clip(a + rng.choice([-65, -64, -63, -1, 0, 1, 63, 64, 65], 20000), 0, 3999)
    pairs = np.stack([a, b], 1).astype(np.uint32)
    got = run_plain(pairs, S, T)
    ref = O.overlay(pairs, S, T, nthreads=8)
    err = rel_err(got, ref, S, T, pairs)
    assert (ref > 0).sum() > 2000
    assert err.max() <= REL_TOL
    assert np.array_equal(got, ref), f"{(got != ref).sum()} of {len(got)} areas not bit-identical"


def test_edge_sizes(dev):
    sq = [[0, 0], [1, 0], [1, 1], [0, 1]]
    A = csr([sq, [[0, 0], [1, 1]], [[0.5, 0.5], [0.5, 0.5], [0.5, 0.5]]])
    B = csr([sq])
    pairs = np.array([[0, 0], [1, 0], [2, 0]], np.uint32)
    np.testing.assert_array_equal(run_plain(pairs, A, B), [1.0, 0.0, 0.0])
    empty = rtree.overlay_plain(torch.zeros((0, 2), dtype=torch.int32, device=dev),
                                rtree.to_device_polys(A), rtree.to_device_polys(B))
    assert empty.numel() == 0


def test_workload1_first_100k_pairs(dev):
    """Join of the two 1,048,576-rect layers, first 100,000 sorted pairs as
    stars with 8..64 vertices, plain FP64 clipping."""
    ra, rb = workloads.w1()
    t = rtree.build_top_down(torch.from_numpy(ra).cuda(), 16)
    pairs_qr = rtree.join(t, torch.from_numpy(rb).cuda()).cpu().numpy().view(np.uint32)
    cp, A, B, _, _ = workloads.clip_inputs_from_pairs(pairs_qr, ra, rb, 100_000)
    got = run_plain(cp, A, B)
    ref = O.overlay(cp, A, B, nthreads=8)
    err = rel_err(got, ref, A, B, cp)
    assert len(cp) == 100_000
    assert err.max() <= REL_TOL
    assert np.array_equal(got, ref)


def _grid_pairs():
    S = gen.stars_grid(gen.seed(9), n=4000, grid=64, jitter=0.25, rfac=1.2, kmax=300)
    T = gen.stars_grid(gen.seed(9, True), n=4000, grid=64, jitter=0.25, rfac=1.2, kmax=300)
    rng = np.random.default_rng(0)
    a = rng.integers(0, 4000, 20000)
    b = np.clip(a + rng.choice([-65, -64, -63, -1, 0, 1, 63, 64, 65], 20000), 0, 3999)
    return np.stack([a, b], 1).astype(np.uint32), S, T


def run_batched(pairs, A, B, tune=None, max_overflow=0):
    st = {}
    area = rtree.overlay_batched(torch.from_numpy(pairs.view(np.int32)).cuda(),
                                 rtree.to_device_polys(A), rtree.to_device_polys(B), tune,
                                 nan_area(len(pairs)), stats=st)
    if max_overflow is not None:
        assert st["overflow"] <= max_overflow, f"{st['overflow']} pairs took the plain path"
    return area.cpu().numpy()


@pytest.mark.parametrize("lanes,waves,factor", [(8, 4, 1), (16, 4, 2), (32, 2, 4), (64, 1, 1),
                                                (64, 4, 2), (128, 1, 1), (128, 1, 4), (256, 1, 2)])
def test_batched_bit_exact_all_team_shapes(dev, lanes, waves, factor):
    pairs, S, T = _grid_pairs()
    tune = rtree.default_tune()
    for b in range(6):
        tune.lanes_per_pair[b] = lanes
        tune.waves_per_block[b] = waves
        tune.list_factor[b] = factor
    # shapes whose teams do not fit LDS for the larger buckets route those
    # buckets to the plain path: results must still be identical
    got = run_batched(pairs, S, T, tune, max_overflow=None)
    ref = O.overlay(pairs, S, T, nthreads=8)
    assert np.array_equal(got, ref), f"{(got != ref).sum()} areas differ"


def test_autotune_picks_valid_shapes(dev):
    # the autotuner's shapes (sampled per bucket, then refined on the whole
    # call) are valid, give the oracle's areas, and time no slower than the
    # defaults on the whole call (noise aside)
    from parlllelrtree_amd.autotune import autotune, tune_to_json
    pairs, S, T = _grid_pairs()
    dS, dT = rtree.to_device_polys(S), rtree.to_device_polys(T)
    tune, stats = autotune(pairs, S, T, dS, dT, max_pairs_per_bucket=2000, reps=1)
    for e in tune_to_json(tune, stats)["buckets"]:
        assert e["lanes"] in (8, 16, 32, 64, 128, 256)
        assert e["lanes"] <= 64 or e["waves"] == 1
    whole = stats["whole"]
    assert whole["calls_timed"] >= 2
    assert whole["tuned_ms"] <= 1.05 * whole["default_ms"]
    got = run_batched(pairs, S, T, tune, max_overflow=None)
    assert np.array_equal(got, O.overlay(pairs, S, T, nthreads=8))


def test_batched_routes_large_buckets_to_plain(dev):
    pairs, S, T = _grid_pairs()
    tune = rtree.default_tune()
    bounds = [8, 12, 16, 20, 24, 8192]   # almost every pair in the last (too large) bucket
    for b in range(6):
        tune.bucket_max_verts[b] = bounds[b]
        tune.lanes_per_pair[b] = 8
        tune.waves_per_block[b] = 8
    got = run_batched(pairs, S, T, tune, max_overflow=None)
    ref = O.overlay(pairs, S, T, nthreads=8)
    assert np.array_equal(got, ref)


def zigzag(k, lean, y0=0.0, y1=1.0, x0=0.0):
    """Sawtooth of k teeth whose edges all span the band (y0, y1), leaning by
    `lean`, closed below: the band is one beam holding 2k long edges."""
    v = []
    for i in range(k):
        v.append((x0 + i, y0))
        v.append((x0 + i + 0.5 + lean, y1))
    v += [(x0 + k, y0), (x0 + k, y0 - 1.0), (x0, y0 - 1.0)]
    return v


@pytest.mark.parametrize("k", [6, 9, 12, 20, 40, 90])
def test_batched_crossing_dense_beams(dev, k):
    # two sawtooths leaning opposite ways: every tooth edge of one crosses
    # many of the other inside the same beam (many crossing heights, from a
    # few up to more than the team's scratch holds)
    polys_a, polys_b = [], []
    for j, lean in enumerate((0.3, 1.5, 3.7)):
        polys_a.append(zigzag(k, lean, x0=0.1 * j))
        polys_b.append(zigzag(k, -lean, y0=0.05, y1=1.1, x0=0.3))
    A, B = csr(polys_a), csr(polys_b)
    pairs = np.array([(i, j) for i in range(3) for j in range(3)], np.uint32)
    got = run_batched(pairs, A, B, max_overflow=None)
    ref = O.overlay(pairs, A, B)
    assert np.array_equal(got, ref), (got, ref)
    assert np.array_equal(run_plain(pairs, A, B), ref)


def padded_zigzag(k, lean, pad, y0=0.0, y1=1.0, x0=0.0):
    """zigzag() with `pad` extra vertices on its bottom side (below the
    band), so the pair lands in a larger bucket while the band's beam keeps
    its 2k edges."""
    v = zigzag(k, lean, y0, y1, x0)
    bottom = [(x0 + k - (i + 1) * k / (pad + 1), y0 - 1.0 - 0.01 * ((i % 2) + 1)) for i in range(pad)]
    return v[:-1] + bottom + v[-1:]


@pytest.mark.parametrize("pad", [12, 40])
@pytest.mark.parametrize("lanes", [64, 128, 256])
def test_batched_sub_beams_in_several_passes(dev, pad, lanes):
    # beams of 8..16 entries with more crossing heights than a wave has
    # lane groups (64 / pow2(m) sub-beams per pass), in the 64- and
    # 128-vertex buckets, under teams of one, two and four waves
    polys_a, polys_b = [], []
    for k in (2, 3, 4):
        for lean in (0.7, 1.9):
            polys_a.append(padded_zigzag(k, lean, pad, x0=0.05 * k))
            polys_b.append(padded_zigzag(k, -lean, pad, y0=0.05, y1=1.1, x0=0.3))
    A, B = csr(polys_a), csr(polys_b)
    n = len(polys_a)
    pairs = np.array([(i, j) for i in range(n) for j in range(n)], np.uint32)
    tune = rtree.default_tune()
    for b in range(6):
        tune.lanes_per_pair[b] = lanes
        tune.waves_per_block[b] = 1
        tune.list_factor[b] = 1
    got = run_batched(pairs, A, B, tune)
    ref = O.overlay(pairs, A, B)
    assert np.array_equal(got, ref), (got, ref)
    assert (ref > 0).all()


@pytest.mark.parametrize("K", [250, 300])
@pytest.mark.parametrize("lanes", [64, 128, 256])
def test_batched_many_crossing_beams_in_one_window(dev, K, lanes):
    # two staircases whose zigzag sides cross once in every beam: K crossing
    # beams of 4 entries in one window of beams, more than a wave's 64 when
    # the beams are dealt over the waves of a team (two or four waves)
    P = [(1 + 0.3 * (i % 2), float(i)) for i in range(K + 1)] + [(0.0, float(K)), (0.0, 0.0)]
    Q = [(2.0, 0.0), (2.0, float(K))] + [(1.3 - 0.3 * (i % 2), float(i)) for i in reversed(range(K + 1))]
    A, B = csr([P, P]), csr([Q, Q[::-1]])
    pairs = np.array([(0, 0), (1, 1), (0, 1)], np.uint32)
    tune = rtree.default_tune()
    for b in range(6):
        tune.lanes_per_pair[b] = lanes
        tune.waves_per_block[b] = 1
        tune.list_factor[b] = 4
    got = run_batched(pairs, A, B, tune)
    ref = O.overlay(pairs, A, B)
    assert np.array_equal(got, ref), (got, ref)
    assert (ref > 0).all()


def test_batched_golden_and_edges(dev):
    cases = GOLD["clip"]
    A = csr([c["P"] if len(c["P"]) else [[0, 0]] for c in cases])
    B = csr([c["Q"] for c in cases])
    idx = np.arange(len(cases), dtype=np.uint32)
    got = run_batched(np.stack([idx, idx], 1), A, B)
    np.testing.assert_allclose(got, [c["area"] for c in cases], rtol=1e-15, atol=1e-15)
    empty = rtree.overlay_batched(torch.zeros((0, 2), dtype=torch.int32, device=dev),
                                  rtree.to_device_polys(A), rtree.to_device_polys(B))
    assert empty.numel() == 0


def test_batched_workload1_first_100k_pairs(dev):
    ra, rb = workloads.w1()
    t = rtree.build_top_down(torch.from_numpy(ra).cuda(), 16)
    pairs_qr = rtree.join(t, torch.from_numpy(rb).cuda()).cpu().numpy().view(np.uint32)
    cp, A, B, _, _ = workloads.clip_inputs_from_pairs(pairs_qr, ra, rb, 100_000)
    got = run_batched(cp, A, B)
    ref = O.overlay(cp, A, B, nthreads=8)
    assert np.array_equal(got, ref)


def test_batched_lognormal_grid_sample(dev):
    """Stars of the 708x708 grid layers (lognormal 4..512 vertices): a
    neighbourhood sample of candidate pairs, batched vs plain vs oracle."""
    A, B = workloads.w2(60_000)
    rng = np.random.default_rng(3)
    a = rng.integers(0, 60_000, 30_000)
    b = np.clip(a + rng.choice([-709, -708, -707, -1, 0, 1, 707, 708, 709], 30_000), 0, 59_999)
    pairs = np.stack([a, b], 1).astype(np.uint32)
    got = run_batched(pairs, A, B)
    ref = O.overlay(pairs, A, B, nthreads=8)
    assert (ref > 0).sum() > 10_000
    assert np.array_equal(got, ref)
    assert np.array_equal(run_plain(pairs, A, B), ref)


@pytest.mark.parametrize("op", ["diff", "xor", "union", "int"])
def test_set_operations_bit_exact(dev, op):
    # GPC's other operations: areas of A - B, A ^ B, A | B by the same sweep;
    # plain and batched (with windows and crossing-dense beams) against the
    # oracle, including empty operands and a sawtooth pair
    pairs, S, T = _grid_pairs()
    ref = O.overlay_op(pairs, S, T, op, nthreads=8)
    d = torch.from_numpy(pairs.view(np.int32)).cuda()
    dS, dT = rtree.to_device_polys(S), rtree.to_device_polys(T)
    got_p = rtree.overlay_plain(d, dS, dT, nan_area(len(pairs)), op=op).cpu().numpy()
    got_b = rtree.overlay_batched(d, dS, dT, None, nan_area(len(pairs)), op=op).cpu().numpy()
    assert np.array_equal(got_p, ref), f"plain: {(got_p != ref).sum()} differ"
    assert np.array_equal(got_b, ref), f"batched: {(got_b != ref).sum()} differ"
    tune = rtree.default_tune()
    for b in range(6):
        tune.lanes_per_pair[b], tune.waves_per_block[b], tune.list_factor[b] = 16, 2, 1
    got_t = rtree.overlay_batched(d, dS, dT, tune, nan_area(len(pairs)), op=op).cpu().numpy()
    assert np.array_equal(got_t, ref)
    A = csr([zigzag(20, 0.3), [[0, 0], [1, 1]], zigzag(9, 1.5)])
    B = csr([zigzag(20, -1.5, y0=0.05, y1=1.1, x0=0.3), zigzag(5, 0.5), [[2, 2]]])
    pp = np.array([(i, j) for i in range(3) for j in range(3)], np.uint32)
    ref2 = O.overlay_op(pp, A, B, op)
    dp = torch.from_numpy(pp.view(np.int32)).cuda()
    dA, dB = rtree.to_device_polys(A), rtree.to_device_polys(B)
    assert np.array_equal(rtree.overlay_batched(dp, dA, dB, op=op).cpu().numpy(), ref2)
    assert np.array_equal(rtree.overlay_plain(dp, dA, dB, op=op).cpu().numpy(), ref2)
