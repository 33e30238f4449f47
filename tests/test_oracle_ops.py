"""Oracle: areas of GPC's other operations (DIFF, XOR, UNION) by the same
sweep, pinned by closed forms (rect pairs, disjoint, nested, identical,
empty operands) and by the set identities
    |A u B| = |A| + |B| - |A n B|,  |A - B| = |A| - |A n B|,
    |A ^ B| = |A u B| - |A n B|
on random simple stars, with |A| the shoelace area (an independent
formula).  The intersection itself is pinned in test_oracle_clip.py."""
import numpy as np
import pytest

import oracle as O

from test_oracle_clip import star


def rect(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], float)


def areas(P, Q):
    return {op: O.clip_area_op(P, Q, op) for op in ("diff", "int", "xor", "union")}


def test_rect_pairs_closed_form():
    rng = np.random.default_rng(11)
    for _ in range(300):
        a = np.sort(rng.uniform(0, 1, (2, 2)), 0)
        b = np.sort(rng.uniform(0, 1, (2, 2)), 0)
        P, Q = rect(a[0, 0], a[0, 1], a[1, 0], a[1, 1]), rect(b[0, 0], b[0, 1], b[1, 0], b[1, 1])
        ap = (a[1, 0] - a[0, 0]) * (a[1, 1] - a[0, 1])
        aq = (b[1, 0] - b[0, 0]) * (b[1, 1] - b[0, 1])
        w = max(0.0, min(a[1, 0], b[1, 0]) - max(a[0, 0], b[0, 0]))
        h = max(0.0, min(a[1, 1], b[1, 1]) - max(a[0, 1], b[0, 1]))
        i = w * h
        got = areas(P, Q)
        tol = dict(rel=1e-12, abs=1e-15)
        assert got["int"] == pytest.approx(i, **tol)
        assert got["union"] == pytest.approx(ap + aq - i, **tol)
        assert got["diff"] == pytest.approx(ap - i, **tol)
        assert got["xor"] == pytest.approx(ap + aq - 2 * i, **tol)


@pytest.mark.parametrize("k", [4, 9, 40, 150])
def test_set_identities_on_stars(k):
    rng = np.random.default_rng(k)
    for _ in range(60):
        P = star(rng, k, 0.0, 0.0, 1.0)
        Q = star(rng, int(rng.integers(4, 2 * k)), rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8),
                 rng.uniform(0.5, 1.5))
        ap, aq = abs(O.poly_area(P)), abs(O.poly_area(Q))
        got = areas(P, Q)
        scale = ap + aq
        assert got["union"] == pytest.approx(ap + aq - got["int"], rel=1e-12, abs=1e-13 * scale)
        assert got["diff"] == pytest.approx(ap - got["int"], rel=1e-12, abs=1e-13 * scale)
        assert got["xor"] == pytest.approx(got["union"] - got["int"], rel=1e-12, abs=1e-13 * scale)
        # B - A by swapping the operands
        assert O.clip_area_op(Q, P, "diff") == pytest.approx(aq - got["int"], rel=1e-12,
                                                             abs=1e-13 * scale)
        assert O.clip_area_op(Q, P, "union") == pytest.approx(got["union"], rel=1e-13,
                                                              abs=1e-14 * scale)


def test_disjoint_nested_identical_and_empty():
    rng = np.random.default_rng(3)
    P = star(rng, 30, 0.0, 0.0, 1.0)
    ap = abs(O.poly_area(P))
    far = star(rng, 12, 5.0, 0.3, 1.0)
    af = abs(O.poly_area(far))
    g = areas(P, far)
    assert g["int"] == 0.0 and g["diff"] == pytest.approx(ap, rel=1e-13)
    assert g["union"] == pytest.approx(ap + af, rel=1e-13)
    assert g["xor"] == pytest.approx(ap + af, rel=1e-13)
    inner = star(rng, 17, 0.0, 0.0, 0.3)          # radii >= 0.5 keep it inside P
    ai = abs(O.poly_area(inner))
    g = areas(P, inner)
    assert g["int"] == pytest.approx(ai, rel=1e-13)
    assert g["union"] == pytest.approx(ap, rel=1e-13)
    assert g["diff"] == pytest.approx(ap - ai, rel=1e-12)
    assert g["xor"] == pytest.approx(ap - ai, rel=1e-12)
    assert O.clip_area_op(inner, P, "diff") == pytest.approx(0.0, abs=1e-15)
    g = areas(P, P)
    assert g["int"] == pytest.approx(ap, rel=1e-13) and g["union"] == pytest.approx(ap, rel=1e-13)
    assert g["diff"] == pytest.approx(0.0, abs=1e-14) and g["xor"] == pytest.approx(0.0, abs=1e-14)
    E = np.zeros((0, 2))
    g = areas(P, E)
    assert g["int"] == 0.0 and g["diff"] == pytest.approx(ap, rel=1e-13)
    assert g["union"] == pytest.approx(ap, rel=1e-13) and g["xor"] == pytest.approx(ap, rel=1e-13)
    g = areas(E, P)
    assert g["int"] == 0.0 and g["diff"] == 0.0
    assert g["union"] == pytest.approx(ap, rel=1e-13) and g["xor"] == pytest.approx(ap, rel=1e-13)


def test_int_op_is_the_intersection():
    rng = np.random.default_rng(5)
    for _ in range(100):
        P = star(rng, 20, 0, 0, 1)
        Q = star(rng, 25, rng.uniform(-1, 1), rng.uniform(-1, 1), 1)
        assert O.clip_area_op(P, Q, "int") == O.clip_area(P, Q)
