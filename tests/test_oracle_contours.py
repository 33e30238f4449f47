"""Oracle: result contours of P op Q (GPC's output polygon,
vector_map_overlay_gpc_kernel.cu:527-750), pinned without GPC:

* closed forms with exact vertex lists: rect op rect, P n P = P, disjoint
  union = both polygons, nested difference = outer + reversed hole, every
  contour compared as a cyclic sequence;
* orientation: outer contours counter-clockwise, holes clockwise, matching
  the hole flags;
* areas: the signed shoelace sum of the contours equals the sweep's area
  (test_oracle_clip.py pins that one independently) for all four operations;
* convex n convex (generic position): the vertices are exactly the vertices
  of each polygon inside the other plus one per boundary crossing, every one
  on both closed polygons, none twice.
GPC's own vertex order and start vertex stay unpinned (GPC is absent)."""
import numpy as np
import pytest

import oracle as O

from test_oracle_clip import star


def rect(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], float)


def shoelace(c):
    x, y = c[:, 0], c[:, 1]
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def canon(c):
    """Cyclic rotation starting at the smallest (y, x) vertex."""
    k = min(range(len(c)), key=lambda i: (c[i, 1], c[i, 0]))
    return np.roll(c, -k, axis=0)


def same_cycle(a, b):
    return a.shape == b.shape and np.array_equal(canon(a), canon(b))


def test_rect_ops_exact():
    rng = np.random.default_rng(4)
    for _ in range(200):
        x0, y0 = rng.integers(0, 10, 2)
        w, h = rng.integers(1, 10, 2)
        u0, v0 = rng.integers(-5, 15, 2)
        p, q = rect(x0, y0, x0 + w, y0 + h), rect(u0, v0, u0 + 3, v0 + 4)
        cs, holes = O.clip_contours(p, q, "int")
        ix0, iy0 = max(x0, u0), max(y0, v0)
        ix1, iy1 = min(x0 + w, u0 + 3), min(y0 + h, v0 + 4)
        if ix0 < ix1 and iy0 < iy1:
            assert len(cs) == 1 and not holes[0]
            # counter-clockwise, starting at the local minimum's left point
            np.testing.assert_array_equal(cs[0], rect(ix0, iy0, ix1, iy1))
        else:
            assert cs == []


def test_self_intersection_is_the_polygon():
    rng = np.random.default_rng(5)
    for k in (3, 7, 30, 200):
        P = star(rng, k, 0.3, -0.2, 1.0)
        cs, holes = O.clip_contours(P, P, "int")
        assert len(cs) == 1 and not holes[0] and same_cycle(cs[0], P)
        cs, _ = O.clip_contours(P, P, "union")
        assert len(cs) == 1 and same_cycle(cs[0], P)
        assert O.clip_contours(P, P, "xor")[0] == []
        assert O.clip_contours(P, P, "diff")[0] == []


def test_disjoint_and_nested():
    rng = np.random.default_rng(6)
    P = star(rng, 25, 0.0, 0.0, 1.0)
    far = star(rng, 11, 4.0, 1.0, 1.0)
    cs, holes = O.clip_contours(P, far, "union")
    assert len(cs) == 2 and not holes.any()
    assert sorted(len(c) for c in cs) == [11, 25]
    got = {len(c): c for c in cs}
    assert same_cycle(got[25], P) and same_cycle(got[11], far)
    cs, holes = O.clip_contours(P, far, "diff")
    assert len(cs) == 1 and same_cycle(cs[0], P)
    inner = star(rng, 9, 0.0, 0.0, 0.3)          # radii >= 0.5 keep it inside P
    cs, holes = O.clip_contours(P, inner, "diff")
    assert len(cs) == 2 and sorted(holes.tolist()) == [False, True]
    hole = cs[int(np.argmax(holes))]
    outer = cs[int(np.argmin(holes))]
    assert same_cycle(outer, P)
    assert same_cycle(hole, inner[::-1])          # the hole runs clockwise
    cs, holes = O.clip_contours(P, inner, "int")
    assert len(cs) == 1 and same_cycle(cs[0], inner)
    cs, holes = O.clip_contours(inner, P, "diff")
    assert cs == []
    cs, holes = O.clip_contours(P, inner, "xor")
    assert len(cs) == 2 and sorted(holes.tolist()) == [False, True]


@pytest.mark.parametrize("op", ["int", "diff", "xor", "union"])
def test_shoelace_equals_area_and_orientation(op):
    rng = np.random.default_rng(10 + len(op))
    for _ in range(400):
        P = star(rng, int(rng.integers(3, 50)), 0.0, 0.0, 1.0)
        Q = star(rng, int(rng.integers(3, 50)), rng.uniform(-1, 1), rng.uniform(-1, 1),
                 rng.uniform(0.3, 1.5))
        cs, holes = O.clip_contours(P, Q, op)
        scale = abs(O.poly_area(P)) + abs(O.poly_area(Q))
        a = O.clip_area_op(P, Q, op)
        s = sum(shoelace(c) for c in cs)
        assert s == pytest.approx(a, rel=1e-12, abs=1e-13 * scale)
        for c, h in zip(cs, holes):
            assert len(c) >= 3
            assert (shoelace(c) < 0) == bool(h)
            # no vertex twice in a row (crossings are evaluated once)
            d = np.abs(np.diff(np.vstack([c, c[:1]]), axis=0)).max(axis=1)
            assert (d > 0).all()


def convex(rng, n, cx, cy, r):
    t = np.sort(rng.uniform(0, 2 * np.pi, n))
    return np.stack([cx + r * np.cos(t), cy + r * np.sin(t)], 1)


def cross2(u, v):
    return u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]


def seg_cross(p, q, a, b):
    d1 = cross2(q - p, a - p)
    d2 = cross2(q - p, b - p)
    d3 = cross2(b - a, p - a)
    d4 = cross2(b - a, q - a)
    return (d1 * d2 < 0) and (d3 * d4 < 0)


def inside_convex(pt, C, tol):
    e = np.roll(C, -1, axis=0) - C
    return (cross2(e, pt - C) >= -tol).all()


def dist_to_boundary(pt, C):
    a, b = C, np.roll(C, -1, axis=0)
    ab = b - a
    t = np.clip(((pt - a) * ab).sum(1) / (ab * ab).sum(1), 0, 1)
    return np.sqrt((((a + t[:, None] * ab) - pt) ** 2).sum(1)).min()


def test_convex_intersection_vertices():
    rng = np.random.default_rng(12)
    checked = 0
    for _ in range(300):
        P = convex(rng, int(rng.integers(3, 25)), 0, 0, 1)
        Q = convex(rng, int(rng.integers(3, 25)), *rng.uniform(-1, 1, 2), rng.uniform(0.5, 1.5))
        cs, holes = O.clip_contours(P, Q, "int")
        pin = [v for v in P if inside_convex(v, Q, 0) and dist_to_boundary(v, Q) > 1e-9]
        qin = [v for v in Q if inside_convex(v, P, 0) and dist_to_boundary(v, P) > 1e-9]
        ncross = sum(seg_cross(P[i], P[(i + 1) % len(P)], Q[j], Q[(j + 1) % len(Q)])
                     for i in range(len(P)) for j in range(len(Q)))
        if not cs:
            assert ncross == 0 and not pin and not qin
            continue
        assert len(cs) == 1 and not holes[0]
        c = cs[0]
        assert len(c) == len(pin) + len(qin) + ncross
        rows = {tuple(v) for v in c}
        assert all(tuple(v) in rows for v in pin) and all(tuple(v) in rows for v in qin)
        for v in c:
            assert inside_convex(v, P, 1e-12) and inside_convex(v, Q, 1e-12)
            assert min(dist_to_boundary(v, P), dist_to_boundary(v, Q)) < 1e-12
        assert shoelace(c) == pytest.approx(O.sh_convex_area(P, Q), rel=1e-11)
        checked += 1
    assert checked > 100


def test_empty_operands():
    rng = np.random.default_rng(3)
    P = star(rng, 12, 0, 0, 1)
    E = np.zeros((0, 2))
    assert O.clip_contours(P, E, "int")[0] == []
    cs, _ = O.clip_contours(P, E, "union")
    assert len(cs) == 1 and same_cycle(cs[0], P)
    cs, _ = O.clip_contours(E, P, "xor")
    assert len(cs) == 1 and same_cycle(cs[0], P)
    assert O.clip_contours(E, P, "diff")[0] == []
