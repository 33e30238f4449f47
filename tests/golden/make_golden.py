"""Writes tests/golden/golden.json: known-answer vectors for the hot path.

The reference ships no tests, fixtures or data files, so these vectors are
derived here from closed forms and from brute force on tiny inputs, in plain
Python -- neither the oracle nor the HIP library is used to make them.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import struct
from pathlib import Path

HERE = Path(__file__).resolve().parent


def f32(v: float) -> float:
    return struct.unpack("f", struct.pack("f", v))[0]


def ref4_bit(w: int, k: int) -> int:
    """Key bit of input bit k of coordinate w in the reference 4-D interleave:
    each key byte j holds the 2-bit pairs (bits 2j+1:2j) of x1 y1 x2 y2 from
    MSB to LSB (BotomUp-rtree.cu:98-123)."""
    return 8 * (k // 2) + (6 - 2 * w) + (k & 1)


def morton_ref4_cases():
    cases = []
    for w in range(4):
        for k in range(16):
            a = [0, 0, 0, 0]
            a[w] = 1 << k
            cases.append({"in": a, "key": 1 << ref4_bit(w, k)})
    cases.append({"in": [0xFFFF, 0, 0, 0], "key": 0xC0C0C0C0C0C0C0C0})
    cases.append({"in": [0, 0, 0, 0xFFFF], "key": 0x0303030303030303})
    cases.append({"in": [0xFFFF] * 4, "key": 0xFFFFFFFFFFFFFFFF})
    cases.append({"in": [1 << 16, 1 << 17, 1 << 20, 1 << 31], "key": 0})  # bits >= 16 dropped
    cases.append({"in": [0x10001, 0, 0, 0], "key": 1 << 6})
    return cases


def center30_cases():
    """(rect, key) with centres on exact binary fractions."""
    def key(qx, qy):
        k = 0
        for b in range(15):
            k |= ((qx >> b) & 1) << (2 * b)
            k |= ((qy >> b) & 1) << (2 * b + 1)
        return k
    out = []
    for cx, cy in [(0.5, 0.5), (0.0, 0.0), (0.25, 0.75), (0.125, 0.5), (1.0, 1.0),
                   (-0.25, 0.5), (2.0, -3.0), (0.999969482421875, 0.0)]:
        h = 0.0625
        rect = [cx - h, cy - h, cx + h, cy + h]
        q = []
        for c in (cx, cy):
            s = c * 32768.0
            q.append(0 if s < 0 else (32767 if s >= 32767 else int(s)))
        out.append({"rect": rect, "key": key(q[0], q[1])})
    return out


def bu_tiny():
    """n = 5, fanout 2, centres on the diagonal: key order = centre order.
    Leaves {1,3} {2,4} {0}, level 1 {leaf0, leaf1} {leaf2}, root last."""
    centres = [0.9, 0.1, 0.5, 0.3, 0.7]
    rects = [[f32(c - 0.05), f32(c - 0.05), f32(c + 0.05), f32(c + 0.05)] for c in centres]
    perm = [1, 3, 2, 4, 0]

    def union(ids):
        return [min(rects[i][0] for i in ids), min(rects[i][1] for i in ids),
                max(rects[i][2] for i in ids), max(rects[i][3] for i in ids)]
    INF = float("inf")
    empty = [INF, INF, -INF, -INF]
    leaves = [[1, 3], [2, 4], [0]]
    entry_mbr, entry_ref = [], []
    for ids in leaves:
        for j in range(2):
            entry_mbr.append(rects[ids[j]] if j < len(ids) else empty)
            entry_ref.append(ids[j] if j < len(ids) else -1)
    # node 3 -> children 0, 1 ; node 4 -> child 2 ; node 5 (root) -> 3, 4
    entry_mbr += [union(leaves[0]), union(leaves[1]), union(leaves[2]), empty,
                  union(leaves[0] + leaves[1]), union(leaves[2])]
    entry_ref += [0, 1, 2, -1, 3, 4]
    return {"rects": rects, "fanout": 2, "perm": perm, "num_nodes": 6,
            "levels": [[0, 3], [3, 2], [5, 1]],
            "node_count": [2, 2, 1, 2, 1, 2], "node_start": [0, 2, 4, 0, 4, 0],
            "node_end": [1, 3, 4, 3, 4, 4], "entry_mbr": entry_mbr, "entry_ref": entry_ref}


def td_tiny():
    """n = 17, fanout 16: offsets [0, 16, 48]; level 0 sorts all by x1, level
    1 sorts the first 16 (one segment) and the 17th by y1."""
    n = 17
    rects = []
    for i in range(n):
        x1 = f32(((i * 7) % 17) / 17.0)
        y1 = f32(((i * 5) % 17) / 17.0)
        rects.append([x1, y1, f32(x1 + 0.01), f32(y1 + 0.01)])
    by_x = sorted(range(n), key=lambda i: (rects[i][0], i))
    seg0 = sorted(by_x[:16], key=lambda i: (rects[i][1], by_x.index(i)))
    perm = seg0 + by_x[16:]
    INF = float("inf")
    empty = [INF, INF, -INF, -INF]
    slot_child = [16, 32] + [0] * 14 + [-(i + 1) for i in range(17)] + [0] * 15

    def union(ids):
        return [min(rects[i][0] for i in ids), min(rects[i][1] for i in ids),
                max(rects[i][2] for i in ids), max(rects[i][3] for i in ids)]
    slot_mbr = [union(perm[:16]), union(perm[16:])] + [empty] * 14
    slot_mbr += [rects[perm[i]] for i in range(17)] + [empty] * 15
    return {"rects": rects, "fanout": 16, "level_offset": [0, 16, 48], "perm": perm,
            "slot_child": slot_child, "slot_mbr": slot_mbr}


def overlap_cases():
    return [
        {"a": [0, 0, 1, 1], "b": [1, 1, 2, 2], "hit": 1},      # corner touch counts
        {"a": [0, 0, 1, 1], "b": [1, 0, 2, 1], "hit": 1},      # edge touch
        {"a": [0, 0, 1, 1], "b": [f32(1.0000001), 0, 2, 1], "hit": 0},
        {"a": [0, 0, 1, 1], "b": [0.25, 0.25, 0.5, 0.5], "hit": 1},  # containment
        {"a": [0, 0, 1, 1], "b": [0.5, -1, 0.6, -0.001], "hit": 0},
    ]


def clip_cases():
    sq = lambda x0, y0, x1, y1: [[x0, y0], [x1, y0], [x1, y1], [x0, y1]]
    cases = [
        {"name": "rect_rect", "P": sq(0, 0, 2, 1), "Q": sq(1, 0.5, 3, 3), "area": 0.5},
        {"name": "rect_self", "P": sq(0, 0, 2, 1), "Q": sq(0, 0, 2, 1), "area": 2.0},
        {"name": "disjoint", "P": sq(0, 0, 1, 1), "Q": sq(2, 2, 3, 3), "area": 0.0},
        {"name": "touching_edge", "P": sq(0, 0, 1, 1), "Q": sq(1, 0, 2, 1), "area": 0.0},
        {"name": "nested", "P": sq(0, 0, 4, 4), "Q": sq(1, 1, 2, 3), "area": 2.0},
        {"name": "square_diamond", "P": sq(-1, -1, 1, 1),
         "Q": [[1.5, 0], [0, 1.5], [-1.5, 0], [0, -1.5]], "area": 3.5},
        {"name": "triangle_square", "P": [[0, 0], [2, 0], [0, 2]], "Q": sq(0, 0, 1, 1),
         "area": 1.0},
        {"name": "triangle_half", "P": [[0, 0], [1, 0], [0, 1]], "Q": sq(0, 0, 1, 1), "area": 0.5},
        {"name": "clockwise_rect", "P": sq(0, 0, 2, 1)[::-1], "Q": sq(1, 0.5, 3, 3), "area": 0.5},
        {"name": "two_vertices", "P": [[0, 0], [1, 1]], "Q": sq(0, 0, 1, 1), "area": 0.0},
        # cross: two 3x1 bars -> 1x1 centre
        {"name": "cross", "P": sq(0, 1, 3, 2), "Q": sq(1, 0, 2, 3), "area": 1.0},
        # self-intersecting bow-tie (even-odd) against a square: two triangles of
        # area 1/4 each inside [0,1]^2
        {"name": "bowtie", "P": [[0, 0], [1, 1], [1, 0], [0, 1]], "Q": sq(0, 0, 1, 1), "area": 0.5},
        # L-shape (non convex) against a square covering its corner
        {"name": "l_shape", "P": [[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]],
         "Q": sq(0.5, 0.5, 1.5, 1.5), "area": 0.75},
        # duplicate consecutive vertices and a collinear vertex
        {"name": "degenerate_vertices", "P": [[0, 0], [1, 0], [1, 0], [2, 0], [2, 2], [0, 2]],
         "Q": sq(1, 1, 3, 3), "area": 1.0},
    ]
    return cases


def main():
    data = {
        "morton_ref4": morton_ref4_cases(),
        "center30": center30_cases(),
        "bu_tiny": bu_tiny(),
        "td_tiny": td_tiny(),
        "overlap": overlap_cases(),
        "clip": clip_cases(),
        "ref_scale": [[0, 1], [10000, 1], [10001, 10], [99999, 10], [100000, 10],
                      [100001, 10], [1000000, 100], [123456789, 100000]],
    }
    out = HERE / "golden.json"
    out.write_text(json.dumps(data, indent=1, allow_nan=True) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
