"""Oracle: Morton keys and the stable ordering, against the golden vectors,
per-bit closed forms and numpy's stable argsort."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLD = json.loads((Path(__file__).parent / "golden" / "golden.json").read_text())


def test_ref4_interleave_golden():
    for c in GOLD["morton_ref4"]:
        assert O.interleave_ref4(*c["in"]) == c["key"], c


def test_ref4d_keys_match_scaled_interleave():
    rng = np.random.default_rng(3)
    r = rng.integers(0, 10**6, size=(500, 4), dtype=np.int64)
    scale = O.ref_scale(int(r.max()))
    keys = O.morton_ref4d(r, scale)
    for i in range(0, 500, 7):
        a = [int(v) // scale for v in r[i]]
        assert int(keys[i]) == O.interleave_ref4(*a)


def test_ref_scale_golden():
    for rng, s in GOLD["ref_scale"]:
        assert O.ref_scale(rng) == s


def test_center30_golden():
    r = np.array([c["rect"] for c in GOLD["center30"]], np.float32)
    k = O.morton_center30(r)
    assert [int(v) for v in k] == [c["key"] for c in GOLD["center30"]]


def _morton_bits(qx, qy):
    k = np.zeros_like(qx, dtype=np.uint64)
    for b in range(15):
        k |= ((qx >> b) & 1).astype(np.uint64) << np.uint64(2 * b)
        k |= ((qy >> b) & 1).astype(np.uint64) << np.uint64(2 * b + 1)
    return k.astype(np.uint32)


def test_center30_random_matches_per_bit_formula():
    rng = np.random.default_rng(5)
    c = rng.uniform(-0.1, 1.1, size=(20000, 2))
    h = rng.uniform(0, 1e-3, size=(20000, 2))
    r = np.concatenate([c - h, c + h], 1).astype(np.float32)[:, [0, 1, 2, 3]]
    k = O.morton_center30(r)
    cx = (np.float32(0.5) * (r[:, 0] + r[:, 2])).astype(np.float32)
    cy = (np.float32(0.5) * (r[:, 1] + r[:, 3])).astype(np.float32)
    q = []
    for cc in (cx, cy):
        s = cc * np.float32(32768.0)
        q.append(np.where(s < 0, 0, np.where(s >= 32767, 32767, np.floor(s))).astype(np.int64))
    np.testing.assert_array_equal(k, _morton_bits(q[0], q[1]))


@pytest.mark.parametrize("dtype,hi", [(np.uint32, 50), (np.uint32, 2**30), (np.uint64, 7),
                                      (np.uint64, 2**62)])
def test_stable_order_matches_numpy(dtype, hi):
    rng = np.random.default_rng(11)
    for n in (0, 1, 2, 17, 1000, 65537):
        keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(dtype)
        np.testing.assert_array_equal(O.stable_order(keys), np.argsort(keys, kind="stable"))
